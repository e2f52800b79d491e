#include "csr.hpp"

#include <cmath>
#include <utility>

namespace sc23ad {

void lmax1_scaling(HostCsr& a, double* b) {
  for (int i = 0; i < a.n; ++i) {
    double mx = 0;
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p)
      if (std::fabs(a.val[p]) > mx) mx = std::fabs(a.val[p]);
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) a.val[p] /= mx;
    if (b) b[i] /= mx;
  }
}

std::vector<double> row_abs_max(const HostCsr& a) {
  std::vector<double> mx(a.n, 0.0);
  for (int i = 0; i < a.n; ++i)
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p)
      if (std::fabs(a.val[p]) > mx[i]) mx[i] = std::fabs(a.val[p]);
  return mx;
}

void hoare_sort(int left, int right, bool ascending, int* key,
                std::vector<int*> const& co) {
  hoare_sort_with(left, right, ascending, key, [&](int a, int b) {
    for (int* arr : co) std::swap(arr[a], arr[b]);
  });
}

int flex_bits(double max_abs) {
  if (max_abs < 1) return 30;
  const int whole = static_cast<int>(max_abs);
  return 32 - (static_cast<int>(std::ceil(std::log2(whole + 1))) + 1);
}

int tri_bits(double max_abs) {
  int b = 0;
  while (b < 30 && max_abs * (1 << b) < (1 << 30)) ++b;
  return b;
}

int32_t to_fixed(double v, int bit) {
  const double d = v * static_cast<double>(1 << bit);
  if (!(d > -2147483649.0 && d < 2147483648.0)) return INT32_MIN;
  return static_cast<int32_t>(d);
}

}  // namespace sc23ad
