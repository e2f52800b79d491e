// One Ruge-Stueben coarsening step with its two heavy numeric phases on the
// GPU: the classical interpolation weights and the row-merge Galerkin
// products (R A, then (R A) P) of the large levels.  The strength graph, the
// two C/F passes (sequential greedy bookkeeping, helper_multigrid.hpp:138-266,
// 378-417) and the transposes stay on the host, and so do the products of
// levels whose rows are few or long (one GPU thread per row then runs a long
// dependent chain of table probes, and the host's threads are faster).
// Every FP64 value is computed in the same operation order as the host code,
// so the step's result is identical to rs_coarsen's bit for bit.
#include <algorithm>
#include <memory>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>

#include "device.hpp"
#include "rs_setup.hpp"

namespace sc23ad {

namespace {

struct DevCsr {
  DevBuf rptr, cind, val;
  int n = 0;
  explicit DevCsr(const HostCsr& a) : n(a.n) {
    rptr.alloc(sizeof(int) * (static_cast<size_t>(a.n) + 1));
    rptr.upload(a.rptr.data(), sizeof(int) * (static_cast<size_t>(a.n) + 1));
    const size_t nz = std::max<size_t>(a.cind.size(), 1);
    cind.alloc(sizeof(int) * nz);
    val.alloc(sizeof(double) * nz);
    if (!a.cind.empty()) {
      cind.upload(a.cind.data(), sizeof(int) * a.cind.size());
      val.upload(a.val.data(), sizeof(double) * a.val.size());
    }
  }
};

// Host interpolation's structure: C rows are unit rows; an F row holds one
// entry per strong C neighbour, numbered among the C points.  The values of
// F rows come from the GPU.
HostCsr interpolation_dev(const DevCsr& da, const HostCsr& cs, const HostCsr& ds,
                          const HostCsr& weak, const std::vector<int>& cf, int* warnings) {
  const int n = cs.n;
  std::vector<int> col(n, 0), fine;
  int nc = 0;
  for (int i = 0; i < n; ++i) {
    if (cf[i]) col[i] = nc++;
    else fine.push_back(i);
  }
  HostCsr p;
  p.reset(n, nc);
  for (int i = 0; i < n; ++i) {
    if (cf[i]) {
      p.cind.push_back(col[i]);
      p.val.push_back(1.0);
    } else {
      for (int q = cs.rptr[i]; q < cs.rptr[i + 1]; ++q) {
        p.cind.push_back(col[cs.cind[q]]);
        p.val.push_back(0.0);
      }
    }
    p.rptr[i + 1] = static_cast<int>(p.cind.size());
  }
  if (fine.empty()) return p;
  DevCsr dcs(cs), dds(ds), dwk(weak);
  DevBuf dfine(sizeof(int) * fine.size()), dpptr(sizeof(int) * (static_cast<size_t>(n) + 1)),
      dpval(sizeof(double) * std::max<size_t>(p.val.size(), 1)), dstat(2 * sizeof(int));
  dfine.upload(fine.data(), sizeof(int) * fine.size());
  dpptr.upload(p.rptr.data(), sizeof(int) * (static_cast<size_t>(n) + 1));
  dpval.upload(p.val.data(), sizeof(double) * p.val.size());
  const int zero[2] = {0, 0};
  dstat.upload(zero, sizeof(zero));
  hip_check(dev::launch_interp_weights(
                static_cast<int>(fine.size()), dfine.as<int>(), da.rptr.as<int>(),
                da.cind.as<int>(), da.val.as<double>(), dcs.rptr.as<int>(), dcs.cind.as<int>(),
                dcs.val.as<double>(), dds.rptr.as<int>(), dds.cind.as<int>(),
                dds.val.as<double>(), dwk.rptr.as<int>(), dwk.val.as<double>(),
                dpptr.as<int>(), dpval.as<double>(), dstat.as<int>(), lib_stream()),
            "interpolation weights");
  int stat[2];
  dstat.download(stat, sizeof(stat));
  if (stat[0])
    throw SetupError{kSetupFail,
                     "interpolation: F-F connection shares no C point (reference exits here)"};
  if (warnings) *warnings += stat[1];
  dpval.download(p.val.data(), sizeof(double) * p.val.size());
  return p;
}

// rowmerge_product on the GPU, rows processed in chunks whose hash tables fit
// a fixed scratch budget.
constexpr long long kScratchSlots = 64ll << 20;  // 64 Mi slots of (int, double)

// The GPU product pays off on many short rows: at 128^3 the finest level's
// products (1M rows, 7-25 terms each) ran 2-4x faster than the host's 16
// threads, the next levels' (262k rows and fewer, or 100+ terms) slower.
constexpr int kGpuProductMinRows = 1 << 20;
constexpr double kGpuProductMaxTerms = 64.0;

HostCsr rowmerge_auto(const HostCsr& l, const HostCsr& r, const DevCsr* dr_have) {
  const int n = l.n;
  std::vector<long long> toff(static_cast<size_t>(n) + 1, 0), ooff(toff);
  for (int i = 0; i < n; ++i) {
    long long terms = 0;
    for (int q = l.rptr[i]; q < l.rptr[i + 1]; ++q) {
      const int k = l.cind[q];
      terms += r.rptr[k + 1] - r.rptr[k];
    }
    toff[i + 1] = toff[i] + 2 * terms + 1;  // hash table: load factor <= 1/2
    ooff[i + 1] = ooff[i] + terms;
  }
  const double mean_terms = n > 0 ? static_cast<double>(ooff[n]) / n : 0.0;
  bool gpu = n >= kGpuProductMinRows && mean_terms <= kGpuProductMaxTerms;
  if (const char* f = std::getenv("SC23AD_SETUP_GPU_PRODUCTS"))  // 1: always, 0: never
    gpu = f[0] == '1';
  if (!gpu) return rowmerge_product(l, r);
  DevCsr dl(l);
  std::unique_ptr<DevCsr> own_r;
  if (!dr_have) own_r = std::make_unique<DevCsr>(r);
  const DevCsr& dr = dr_have ? *dr_have : *own_r;
  DevBuf dtoff(sizeof(long long) * toff.size()), dooff(sizeof(long long) * ooff.size());
  dtoff.upload(toff.data(), sizeof(long long) * toff.size());
  dooff.upload(ooff.data(), sizeof(long long) * ooff.size());
  HostCsr out;
  out.reset(n, r.m);
  DevBuf tkey, tval, ocol, oval, cnt, optr, gcol, gval;
  int row0 = 0;
  while (row0 < n) {
    // the largest chunk whose table fits the budget (at least one row)
    int row1 = row0 + 1;
    while (row1 < n && toff[row1 + 1] - toff[row0] <= kScratchSlots) ++row1;
    const long long nt = toff[row1] - toff[row0], no = ooff[row1] - ooff[row0];
    const int rows = row1 - row0;
    auto grow = [](DevBuf& b, size_t bytes) {
      if (b.bytes() < bytes) b.alloc(bytes);
    };
    grow(tkey, sizeof(int) * nt);
    grow(tval, sizeof(double) * nt);
    grow(ocol, sizeof(int) * std::max<long long>(no, 1));
    grow(oval, sizeof(double) * std::max<long long>(no, 1));
    grow(cnt, sizeof(int) * rows);
    hip_check(dev::launch_rowmerge(row0, row1, dl.rptr.as<int>(), dl.cind.as<int>(),
                                   dl.val.as<double>(), dr.rptr.as<int>(), dr.cind.as<int>(),
                                   dr.val.as<double>(), dtoff.as<long long>(),
                                   dooff.as<long long>(), toff[row0], ooff[row0],
                                   tkey.as<int>(), tval.as<double>(), ocol.as<int>(),
                                   oval.as<double>(), cnt.as<int>(), lib_stream()),
              "rowmerge");
    std::vector<int> c(rows), lptr(static_cast<size_t>(rows) + 1, 0);
    cnt.download(c.data(), sizeof(int) * rows);
    for (int k = 0; k < rows; ++k) lptr[k + 1] = lptr[k] + c[k];
    const int nz = lptr[rows];
    grow(optr, sizeof(int) * (static_cast<size_t>(rows) + 1));
    optr.upload(lptr.data(), sizeof(int) * lptr.size());
    grow(gcol, sizeof(int) * std::max(nz, 1));
    grow(gval, sizeof(double) * std::max(nz, 1));
    hip_check(dev::launch_rowmerge_gather(row0, row1, dooff.as<long long>(), ooff[row0],
                                          ocol.as<int>(), oval.as<double>(), optr.as<int>(),
                                          gcol.as<int>(), gval.as<double>(), lib_stream()),
              "rowmerge gather");
    const size_t at = out.cind.size();
    out.cind.resize(at + nz);
    out.val.resize(at + nz);
    if (nz > 0) {
      gcol.download(out.cind.data() + at, sizeof(int) * nz);
      gval.download(out.val.data() + at, sizeof(double) * nz);
    }
    for (int k = 0; k < rows; ++k) out.rptr[row0 + k + 1] = static_cast<int>(at) + lptr[k + 1];
    row0 = row1;
  }
  return out;
}

}  // namespace

CoarsenStep rs_coarsen_dev(const HostCsr& a, double theta) {
  if (theta <= 0 || 1 < theta)
    throw SetupError{kBadParam, "theta must satisfy 0 < theta <= 1"};
  const bool verbose = std::getenv("SC23AD_VERBOSE") != nullptr;
  auto t = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    if (!verbose) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[sc23ad setup/dev] n=%d %-14s %8.3f s\n", a.n, what,
                 std::chrono::duration<double>(now - t).count());
    t = now;
  };
  CoarsenStep step;
  HostCsr strong, weak;
  rs_strength(a, theta, strong, weak);
  HostCsr strong_t = pattern_transpose(strong);
  lap("strength");
  step.cf = rs_first_pass(strong, strong_t, &step.warnings);
  rs_second_pass(strong, step.cf);
  lap("C/F passes");
  HostCsr cs, ds;
  split_strong(strong, step.cf, cs, ds);
  DevCsr da(a);
  lap("split, upload");
  step.interp = interpolation_dev(da, cs, ds, weak, step.cf, &step.warnings);
  lap("interpolation");
  step.restrict_ = csr_transpose(step.interp);
  lap("transpose");
  // get_coarse_matrix (helper_multigrid.hpp:564-586): rows of P^T A first,
  // then (P^T A) P.
  HostCsr rt_a = rowmerge_auto(step.restrict_, a, &da);
  lap("R*A");
  step.coarse = rowmerge_auto(rt_a, step.interp, nullptr);
  lap("(RA)*P");
  return step;
}

}  // namespace sc23ad
