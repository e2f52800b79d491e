#include "learn.hpp"

#include <algorithm>
#include <cmath>
#include <random>

namespace sc23ad {

void level_features(const HostCsr& a, int level, int depth, float* f) {
  const int n = a.n;
  double dom_sum = 0, dom_min = 1, ratio_sum = 0, absum = 0, asym = 0;
  long long off = 0, pos = 0;
  for (int i = 0; i < n; ++i) {
    double d = 0, s = 0, mo = 0;
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      const int c = a.cind[p];
      const double v = a.val[p];
      s += std::fabs(v);
      if (c == i) {
        d = std::fabs(v);
      } else {
        ++off;
        if (v > 0) ++pos;
        mo = std::max(mo, std::fabs(v));
        // a_ji by binary search in row c (columns are sorted)
        const int* b = a.cind.data() + a.rptr[c];
        const int* e = a.cind.data() + a.rptr[c + 1];
        const int* q = std::lower_bound(b, e, i);
        const double vt = (q != e && *q == i) ? a.val[q - a.cind.data()] : 0.0;
        asym += std::fabs(v - vt);
      }
    }
    absum += s;
    const double dom = s > 0 ? d / s : 0;
    dom_sum += dom;
    dom_min = std::min(dom_min, dom);
    ratio_sum += d > 0 ? mo / d : 0;
  }
  const double nn = std::max(n, 1);
  f[0] = static_cast<float>(std::log2(nn) / 24.0);
  f[1] = static_cast<float>(static_cast<double>(a.nnz()) / nn / 32.0);
  f[2] = static_cast<float>(dom_sum / nn);
  f[3] = static_cast<float>(n > 0 ? dom_min : 0);
  f[4] = static_cast<float>(ratio_sum / nn);
  f[5] = static_cast<float>(off > 0 ? static_cast<double>(pos) / off : 0);
  f[6] = static_cast<float>(absum > 0 ? asym / absum : 0);
  f[7] = static_cast<float>(depth > 0 ? static_cast<double>(level) / depth : 0);
}

void mlp_init(uint64_t seed, float* p) {
  std::mt19937_64 g(seed);
  auto uni = [&](double lim) { return static_cast<float>(((g() >> 11) * 0x1.0p-53 * 2 - 1) * lim); };
  float* w = p;
  auto layer = [&](int out, int in) {
    const double lim = std::sqrt(6.0 / (in + out));
    for (int k = 0; k < out * in; ++k) *w++ = uni(lim);
    for (int k = 0; k < out; ++k) *w++ = 0.0f;
  };
  layer(kMlpHidden, kMlpIn);
  layer(kMlpHidden, kMlpHidden);
  layer(kMlpOut, kMlpHidden);
}

}  // namespace sc23ad
