#include "gen.hpp"

#include <algorithm>
#include <cmath>

namespace sc23ad {

double Rng::normal() {
  const double u1 = 1.0 - uniform();  // (0, 1]
  const double u2 = uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

HostCsr gen_poisson2d(int n) {
  HostCsr a;
  const int rows = n * n;
  a.reset(rows, rows);
  a.cind.reserve(static_cast<size_t>(rows) * 5);
  a.val.reserve(static_cast<size_t>(rows) * 5);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const int i = y * n + x;
      auto put = [&](int c, double v) {
        a.cind.push_back(c);
        a.val.push_back(v);
      };
      if (y > 0) put(i - n, -1.0);
      if (x > 0) put(i - 1, -1.0);
      put(i, 4.0);
      if (x < n - 1) put(i + 1, -1.0);
      if (y < n - 1) put(i + n, -1.0);
      a.rptr[i + 1] = static_cast<int>(a.cind.size());
    }
  return a;
}

HostCsr gen_laplace3d(int n) {
  HostCsr a;
  const int rows = n * n * n;
  const int nn = n * n;
  a.reset(rows, rows);
  a.cind.reserve(static_cast<size_t>(rows) * 7);
  a.val.reserve(static_cast<size_t>(rows) * 7);
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int i = (z * n + y) * n + x;
        auto put = [&](int c, double v) {
          a.cind.push_back(c);
          a.val.push_back(v);
        };
        if (z > 0) put(i - nn, -1.0);
        if (y > 0) put(i - n, -1.0);
        if (x > 0) put(i - 1, -1.0);
        put(i, 6.0);
        if (x < n - 1) put(i + 1, -1.0);
        if (y < n - 1) put(i + n, -1.0);
        if (z < n - 1) put(i + nn, -1.0);
        a.rptr[i + 1] = static_cast<int>(a.cind.size());
      }
  return a;
}

namespace {
inline double hmean(double a, double b) { return 2.0 * a * b / (a + b); }
}  // namespace

HostCsr gen_diffusion27(int n, int block, double ax, double ay, double az,
                        uint64_t seed) {
  const int nb = (n + block - 1) / block;
  std::vector<double> kblk(static_cast<size_t>(nb) * nb * nb);
  Rng rng(seed + 1000003ULL);
  for (auto& k : kblk) k = std::exp(2.0 * rng.normal());
  auto kat = [&](int x, int y, int z) {
    return kblk[(static_cast<size_t>(z / block) * nb + y / block) * nb + x / block];
  };
  HostCsr a;
  const int rows = n * n * n;
  a.reset(rows, rows);
  a.cind.reserve(static_cast<size_t>(rows) * 27);
  a.val.reserve(static_cast<size_t>(rows) * 27);
  const double axis[3] = {ax, ay, az};
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int i = (z * n + y) * n + x;
        const double ki = kat(x, y, z);
        double diag = 0;
        size_t diag_pos = 0;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              if (dx == 0 && dy == 0 && dz == 0) {
                diag_pos = a.cind.size();
                a.cind.push_back(i);
                a.val.push_back(0.0);
                continue;
              }
              const int d[3] = {dx, dy, dz};
              double w = 0;
              int m = 0;
              for (int t = 0; t < 3; ++t)
                if (d[t]) {
                  w += axis[t];
                  ++m;
                }
              w /= static_cast<double>(1 << (2 * (m - 1)));
              const int xx = x + dx, yy = y + dy, zz = z + dz;
              const bool inside = xx >= 0 && xx < n && yy >= 0 && yy < n &&
                                  zz >= 0 && zz < n;
              const double kij = inside ? hmean(ki, kat(xx, yy, zz)) : ki;
              diag += w * kij;
              if (inside) {
                a.cind.push_back((zz * n + yy) * n + xx);
                a.val.push_back(-w * kij);
              }
            }
        a.val[diag_pos] = diag;
        a.rptr[i + 1] = static_cast<int>(a.cind.size());
      }
  return a;
}

HostCsr gen_convdiff2d(int n, double pe, double angle) {
  const double h = 1.0 / (n + 1);
  const double bx = pe * std::cos(angle), by = pe * std::sin(angle);
  const double w = -1.0 - h * std::max(bx, 0.0);
  const double e = -1.0 - h * std::max(-bx, 0.0);
  const double s = -1.0 - h * std::max(by, 0.0);
  const double no = -1.0 - h * std::max(-by, 0.0);
  const double dg = 4.0 + h * (std::fabs(bx) + std::fabs(by));
  HostCsr a;
  const int rows = n * n;
  a.reset(rows, rows);
  a.cind.reserve(static_cast<size_t>(rows) * 5);
  a.val.reserve(static_cast<size_t>(rows) * 5);
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < n; ++x) {
      const int i = y * n + x;
      auto put = [&](int c, double v) {
        a.cind.push_back(c);
        a.val.push_back(v);
      };
      if (y > 0) put(i - n, s);
      if (x > 0) put(i - 1, w);
      put(i, dg);
      if (x < n - 1) put(i + 1, e);
      if (y < n - 1) put(i + n, no);
      a.rptr[i + 1] = static_cast<int>(a.cind.size());
    }
  return a;
}

HostCsr gen_inclusions3d(int n, int nsph, double rmin, double rmax, double kin,
                         uint64_t seed) {
  Rng rng(seed + 1000003ULL);
  std::vector<double> cx(nsph), cy(nsph), cz(nsph), rr(nsph);
  for (int s = 0; s < nsph; ++s) {
    cx[s] = rng.uniform(0.0, n);
    cy[s] = rng.uniform(0.0, n);
    cz[s] = rng.uniform(0.0, n);
    rr[s] = rng.uniform(rmin, rmax);
  }
  const size_t rows = static_cast<size_t>(n) * n * n;
  std::vector<double> k(rows, 1.0);
  for (int s = 0; s < nsph; ++s) {
    const int x0 = std::max(0, static_cast<int>(std::floor(cx[s] - rr[s])));
    const int x1 = std::min(n - 1, static_cast<int>(std::ceil(cx[s] + rr[s])));
    const int y0 = std::max(0, static_cast<int>(std::floor(cy[s] - rr[s])));
    const int y1 = std::min(n - 1, static_cast<int>(std::ceil(cy[s] + rr[s])));
    const int z0 = std::max(0, static_cast<int>(std::floor(cz[s] - rr[s])));
    const int z1 = std::min(n - 1, static_cast<int>(std::ceil(cz[s] + rr[s])));
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          const double dx = x + 0.5 - cx[s], dy = y + 0.5 - cy[s],
                       dz = z + 0.5 - cz[s];
          if (dx * dx + dy * dy + dz * dz <= rr[s] * rr[s])
            k[(static_cast<size_t>(z) * n + y) * n + x] = kin;
        }
  }
  HostCsr a;
  a.reset(static_cast<int>(rows), static_cast<int>(rows));
  a.cind.reserve(rows * 7);
  a.val.reserve(rows * 7);
  const int nn = n * n;
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        const int i = (z * n + y) * n + x;
        const double ki = k[i];
        double diag = 0;
        auto coup = [&](bool inside, int j) {
          const double kij = inside ? hmean(ki, k[j]) : ki;
          diag += kij;
          return kij;
        };
        // neighbours in ascending column order
        const int nb[6] = {i - nn, i - n, i - 1, i + 1, i + n, i + nn};
        const bool in[6] = {z > 0, y > 0, x > 0, x < n - 1, y < n - 1, z < n - 1};
        double c[6];
        for (int t = 0; t < 6; ++t) c[t] = coup(in[t], in[t] ? nb[t] : i);
        for (int t = 0; t < 3; ++t)
          if (in[t]) {
            a.cind.push_back(nb[t]);
            a.val.push_back(-c[t]);
          }
        a.cind.push_back(i);
        a.val.push_back(diag);
        for (int t = 3; t < 6; ++t)
          if (in[t]) {
            a.cind.push_back(nb[t]);
            a.val.push_back(-c[t]);
          }
        a.rptr[i + 1] = static_cast<int>(a.cind.size());
      }
  return a;
}

void gen_xstar(int n, uint64_t seed, double* xstar) {
  Rng rng(seed);
  for (int i = 0; i < n; ++i) xstar[i] = 2.0 * rng.uniform() - 1.0;
}

void csr_matvec(const HostCsr& a, const double* x, double* y) {
  for (int i = 0; i < a.n; ++i) {
    double t = 0;
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) t += a.val[p] * x[a.cind[p]];
    y[i] = t;
  }
}

std::vector<ConvDiffDraw> draw_convdiff_family(int count, uint64_t seed) {
  Rng rng(seed);
  std::vector<ConvDiffDraw> out(count);
  for (auto& d : out) {
    d.n = rng.uniform_int(32, 256);
    d.pe = rng.uniform(1.0, 100.0);
    d.angle = rng.uniform(0.0, 2.0 * M_PI);
  }
  return out;
}

}  // namespace sc23ad
