#include "device.hpp"

#include <algorithm>
#include <numeric>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace sc23ad {

using dev::Elem;
using dev::SpmvMode;
using dev::TriKind;

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw DevError{5, std::string(what) + ": " + hipGetErrorString(e)};
}

namespace {
thread_local hipStream_t tl_stream = nullptr;
}

StreamScope::StreamScope(hipStream_t s) : prev_(tl_stream) { tl_stream = s; }
StreamScope::~StreamScope() { tl_stream = prev_; }

hipStream_t lib_stream() {
  if (tl_stream) return tl_stream;
  static hipStream_t s = nullptr;
  if (!s) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
      throw DevError{8, "no HIP device available"};
    hip_check(hipSetDevice(0), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  }
  return s;
}

void DevBuf::alloc(size_t bytes) {
  release();
  lib_stream();
  if (bytes == 0) bytes = 4;
  hip_check(hipMalloc(&p_, bytes), "hipMalloc");
  bytes_ = bytes;
}

void DevBuf::release() {
  if (p_) (void)hipFree(p_);
  p_ = nullptr;
  bytes_ = 0;
}

void DevBuf::upload(const void* src, size_t bytes, size_t off) {
  if (bytes == 0) return;
  hip_check(hipMemcpyAsync(static_cast<char*>(p_) + off, src, bytes,
                           hipMemcpyHostToDevice, lib_stream()),
            "upload");
  hip_check(hipStreamSynchronize(lib_stream()), "upload sync");
}

void DevBuf::download(void* dst, size_t bytes, size_t off) const {
  if (bytes == 0) return;
  hip_check(hipMemcpyAsync(dst, static_cast<const char*>(p_) + off, bytes,
                           hipMemcpyDeviceToHost, lib_stream()),
            "download");
  hip_check(hipStreamSynchronize(lib_stream()), "download sync");
}

size_t DevMat::algo_bytes_spmv(size_t elem) const {
  // values + column indices + row pointers + gathered x (one read per column)
  // + y written once.
  return (elem + 4) * static_cast<size_t>(d.nnz) + 4 * static_cast<size_t>(d.n + 1) +
         elem * static_cast<size_t>(d.m) + elem * static_cast<size_t>(d.n);
}

namespace {

int pick_group(int n, int nnz, int forced) {
  if (forced > 0) return forced;
  const double avg = n > 0 ? static_cast<double>(nnz) / n : 1.0;
  int g = 1;
  while (g < 64 && g < avg) g <<= 1;
  return g;
}

template <typename V>
void upload_mat(DevMat& dm, int n, int m, int bit, int dbit,
                const std::vector<int>& rptr, const std::vector<int>& cind,
                const std::vector<V>& val, int group) {
  dm.rptr.alloc(sizeof(int) * rptr.size());
  dm.rptr.upload(rptr.data(), sizeof(int) * rptr.size());
  dm.cind.alloc(sizeof(int) * std::max<size_t>(cind.size(), 1));
  dm.cind.upload(cind.data(), sizeof(int) * cind.size());
  dm.val.alloc(sizeof(V) * std::max<size_t>(val.size(), 1));
  dm.val.upload(val.data(), sizeof(V) * val.size());
  dm.d.n = n;
  dm.d.m = m;
  dm.d.nnz = static_cast<int>(cind.size());
  dm.d.bit = bit;
  dm.d.dbit = dbit;
  dm.d.rptr = dm.rptr.as<int>();
  dm.d.cind = dm.cind.as<int>();
  dm.d.val = dm.val.get();
  dm.d.group = pick_group(n, dm.d.nnz, group);
  // SELL-64 copy when the rows of each 64-row slice have near-equal lengths
  // (padding at most 1/8 of the entries): one thread per row, coalesced.
  // Rows of irregular lengths (prolongation: one entry on coarse points,
  // several on fine ones) are sorted by length, longest first, inside windows
  // of kSellWindow rows, when that brings the padding under 1/2 of the
  // entries (the one-lane-per-row CSR kernel reads such rows uncoalesced).
  // Whether the copy is used is decided by tune_spmv.
  if (n >= 4096) {
    const int ns = (n + 63) / 64;
    auto slices = [&](const std::vector<int>& order, std::vector<long long>& soff, int num,
                      int den) {
      soff.assign(static_cast<size_t>(ns) + 1, 0);
      for (int sl = 0; sl < ns; ++sl) {
        int w = 0;
        for (int k = sl * 64; k < std::min(n, sl * 64 + 64); ++k) {
          const int r = order.empty() ? k : order[k];
          w = std::max(w, rptr[r + 1] - rptr[r]);
        }
        soff[sl + 1] = soff[sl] + 64LL * w;
      }
      return soff[ns] * den <= static_cast<long long>(cind.size()) * num;
    };
    std::vector<int> order;
    std::vector<long long> soff;
    bool ok = slices(order, soff, 9, 8);
    if (!ok) {
      order.resize(n);
      std::iota(order.begin(), order.end(), 0);
      for (int w0 = 0; w0 < n; w0 += dev::kSellWindow) {
        const int w1 = std::min(n, w0 + dev::kSellWindow);
        std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int a, int b) {
          return rptr[a + 1] - rptr[a] > rptr[b + 1] - rptr[b];
        });
      }
      ok = slices(order, soff, 3, 2);  // padding at most 1/2: the tuner decides
    }
    if (ok) {
      std::vector<int> scol(static_cast<size_t>(std::max<long long>(soff[ns], 1)), -1);
      std::vector<V> sval(scol.size(), V(0));
      for (int k = 0; k < n; ++k) {
        const int i = order.empty() ? k : order[k];
        const long long b = soff[k / 64] + (k % 64);
        for (int p = rptr[i]; p < rptr[i + 1]; ++p) {
          scol[static_cast<size_t>(b + 64LL * (p - rptr[i]))] = cind[p];
          sval[static_cast<size_t>(b + 64LL * (p - rptr[i]))] = val[p];
        }
      }
      dm.soff.alloc(sizeof(long long) * soff.size());
      dm.soff.upload(soff.data(), sizeof(long long) * soff.size());
      dm.scol.alloc(sizeof(int) * scol.size());
      dm.scol.upload(scol.data(), sizeof(int) * scol.size());
      dm.sval.alloc(sizeof(V) * sval.size());
      dm.sval.upload(sval.data(), sizeof(V) * sval.size());
      dm.d.soff = dm.soff.as<long long>();
      dm.d.scol = dm.scol.as<int>();
      dm.d.sval = dm.sval.get();
      if (!order.empty()) {
        dm.srow.alloc(sizeof(int) * order.size());
        dm.srow.upload(order.data(), sizeof(int) * order.size());
        dm.d.srow = dm.srow.as<int>();
      }
      // Integer square matrices whose entries lie within 32767 of the
      // diagonal also get 16-bit column offsets (a tuner candidate).
      if (std::is_same_v<V, int32_t> && order.empty() && n == m) {
        std::vector<int16_t> s16(scol.size(), 0);
        bool fits = true;
        for (int k = 0; k < n && fits; ++k) {
          const long long b = soff[k / 64] + (k % 64);
          for (int p = rptr[k]; p < rptr[k + 1]; ++p) {
            const int d = cind[p] - k;
            if (d < -32767 || d > 32767) {
              fits = false;
              break;
            }
            s16[static_cast<size_t>(b + 64LL * (p - rptr[k]))] = static_cast<int16_t>(d);
          }
        }
        if (fits) {
          dm.scol16.alloc(sizeof(int16_t) * s16.size());
          dm.scol16.upload(s16.data(), sizeof(int16_t) * s16.size());
        }
      }
    }
    // Rectangular matrices (restriction, prolongation) of a renumbered
    // hierarchy gather from a vector numbered in another level's order; a
    // second copy with rows ordered by their first column (then by length
    // inside windows of kSellWindow rows where lengths vary) keeps the
    // gathers of one slice close together; its writes scatter instead.  A
    // tuner candidate: at 128^3 the restriction took 21 us instead of 31.
    if (n != m && !cind.empty()) {
      std::vector<int> ord(n);
      std::iota(ord.begin(), ord.end(), 0);
      auto first = [&](int r) { return rptr[r] < rptr[r + 1] ? cind[rptr[r]] : INT32_MAX; };
      // inside windows of 4096 rows (SC23AD_SELL_COLWIN rows; 0: all rows),
      // so the scattered writes of one window merge in L2: at 128^3 the
      // prolongation took 39 us instead of 49 us with one global order
      static const int colwin = [] {
        const char* e = std::getenv("SC23AD_SELL_COLWIN");
        return e ? std::atoi(e) : 4096;
      }();
      const int cw = colwin > 0 ? colwin : n;
      for (int w0 = 0; w0 < n; w0 += cw)
        std::stable_sort(ord.begin() + w0, ord.begin() + std::min(n, w0 + cw),
                         [&](int a, int b) { return first(a) < first(b); });
      std::vector<long long> soff2;
      bool ok2 = slices(ord, soff2, 9, 8);
      if (!ok2) {  // then by length inside windows, as above
        for (int w0 = 0; w0 < n; w0 += dev::kSellWindow) {
          const int w1 = std::min(n, w0 + dev::kSellWindow);
          std::stable_sort(ord.begin() + w0, ord.begin() + w1, [&](int a, int b) {
            return rptr[a + 1] - rptr[a] > rptr[b + 1] - rptr[b];
          });
        }
        ok2 = slices(ord, soff2, 3, 2);
      }
      if (ok2) {
        std::vector<int> scol2(static_cast<size_t>(std::max<long long>(soff2[ns], 1)), -1);
        std::vector<V> sval2(scol2.size(), V(0));
        for (int k = 0; k < n; ++k) {
          const int i = ord[k];
          const long long b = soff2[k / 64] + (k % 64);
          for (int p = rptr[i]; p < rptr[i + 1]; ++p) {
            scol2[static_cast<size_t>(b + 64LL * (p - rptr[i]))] = cind[p];
            sval2[static_cast<size_t>(b + 64LL * (p - rptr[i]))] = val[p];
          }
        }
        dm.soff2.alloc(sizeof(long long) * soff2.size());
        dm.soff2.upload(soff2.data(), sizeof(long long) * soff2.size());
        dm.scol2.alloc(sizeof(int) * scol2.size());
        dm.scol2.upload(scol2.data(), sizeof(int) * scol2.size());
        dm.sval2.alloc(sizeof(V) * sval2.size());
        dm.sval2.upload(sval2.data(), sizeof(V) * sval2.size());
        dm.srow2.alloc(sizeof(int) * ord.size());
        dm.srow2.upload(ord.data(), sizeof(int) * ord.size());
        dm.d2 = dm.d;
        dm.d2.soff = dm.soff2.as<long long>();
        dm.d2.scol = dm.scol2.as<int>();
        dm.d2.sval = dm.sval2.get();
        dm.d2.srow = dm.srow2.as<int>();
        dm.d2.scol16 = nullptr;
        dm.has2 = true;
      }
    }
  }
}

// Level of every row of a triangular factor.  Dependencies of row i: every
// stored column except the stored inverse diagonal (last entry for LD, first
// for DU; L has none).
std::vector<int> tri_levels(TriKind kind, int n, const std::vector<int>& rptr,
                            const std::vector<int>& cind) {
  if (kind != TriKind::kL)
    for (int i = 0; i < n; ++i)
      if (rptr[i + 1] == rptr[i])
        throw DevError{2, "triangular factor row without a stored diagonal"};
  std::vector<int> lev(n, 0);
  auto dep_range = [&](int i, int& b, int& e) {
    b = rptr[i];
    e = rptr[i + 1];
    if (kind == TriKind::kLD) e = std::max(b, e - 1);
    if (kind == TriKind::kDU) b = std::min(e, b + 1);
  };
  if (kind == TriKind::kDU) {
    for (int i = n - 1; i >= 0; --i) {
      int b, e, L = 0;
      dep_range(i, b, e);
      for (int p = b; p < e; ++p) {
        const int c = cind[p];
        if (c <= i) throw DevError{2, "triangular factor has an entry on the wrong side"};
        L = std::max(L, lev[c] + 1);
      }
      lev[i] = L;
    }
  } else {
    for (int i = 0; i < n; ++i) {
      int b, e, L = 0;
      dep_range(i, b, e);
      for (int p = b; p < e; ++p) {
        const int c = cind[p];
        if (c >= i) throw DevError{2, "triangular factor has an entry on the wrong side"};
        L = std::max(L, lev[c] + 1);
      }
      lev[i] = L;
    }
  }
  return lev;
}

// Rows of each chain [chain_begin[c], chain_begin[c+1]) ordered by (level,
// row) (counting sort): perm[position] = row; lev_ptr = level boundaries
// over all chains; chain_lev = 2 level indices per chain.
int order_by_level(const std::vector<int>& lev, const std::vector<int>& chain_begin,
                   std::vector<int>& perm, std::vector<int>& lev_ptr,
                   std::vector<int>& chain_lev) {
  const int nchain = static_cast<int>(chain_begin.size()) - 1;
  perm.assign(lev.size(), 0);
  lev_ptr.assign(1, 0);
  chain_lev.clear();
  int maxlev_all = 0;
  for (int c = 0; c < nchain; ++c) {
    const int rb = chain_begin[c], re = chain_begin[c + 1];
    int maxlev = -1;
    for (int i = rb; i < re; ++i) maxlev = std::max(maxlev, lev[i]);
    const int nl = maxlev + 1;
    std::vector<int> cnt(nl + 1, 0);
    for (int i = rb; i < re; ++i) ++cnt[lev[i] + 1];
    for (int L = 0; L < nl; ++L) cnt[L + 1] += cnt[L];
    const int base = lev_ptr.back();
    const int first_lev = static_cast<int>(lev_ptr.size()) - 1;
    for (int L = 1; L <= nl; ++L) lev_ptr.push_back(base + cnt[L]);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int i = rb; i < re; ++i) perm[base + fill[lev[i]]++] = i;
    chain_lev.push_back(first_lev);
    chain_lev.push_back(first_lev + nl);
    maxlev_all = std::max(maxlev_all, nl);
  }
  return maxlev_all;
}

// Level schedule of a triangular factor.  `chain_begin` (nchain + 1 row
// boundaries) splits rows into independent contiguous chains (the BlockSOR
// partition), each walked by one workgroup in a single launch; with one
// chain, levels are walked in order, small ones inside one workgroup and
// large ones by a launch of their own.
template <typename V>
void build_plan(TriPlan& plan, TriKind kind, int n, const std::vector<int>& rptr,
                const std::vector<int>& cind, const std::vector<V>& val,
                const std::vector<int>& chain_begin, int single_wg_max) {
  plan.kind = kind;
  const std::vector<int> lev = tri_levels(kind, n, rptr, cind);
  const int nchain = static_cast<int>(chain_begin.size()) - 1;
  std::vector<int> perm, lev_ptr, chain_lev;
  const int maxlev_all = order_by_level(lev, chain_begin, perm, lev_ptr, chain_lev);
  plan.identity = true;
  for (int i = 0; i < n; ++i)
    if (perm[i] != i) plan.identity = false;
  plan.nlevels = maxlev_all;
  // segments
  std::vector<int> seg_chain_lev;  // chain_lev entries of single-chain segments
  struct PendingSeg {
    dev::TriSegment s;
    int chain_off;  // offset into the chain_lev buffer, or -1
  };
  std::vector<PendingSeg> pend;
  std::vector<int> all_chain_lev;
  if (nchain > 1) {
    all_chain_lev = chain_lev;
    PendingSeg ps{};
    ps.s.chain_kernel = 1;
    ps.s.nchain = nchain;
    ps.chain_off = 0;
    pend.push_back(ps);
  } else {
    const int nl = static_cast<int>(lev_ptr.size()) - 1;
    int L = 0;
    while (L < nl) {
      const int sz = lev_ptr[L + 1] - lev_ptr[L];
      if (sz > single_wg_max) {
        PendingSeg ps{};
        ps.s.chain_kernel = 0;
        ps.s.row_begin = lev_ptr[L];
        ps.s.row_end = lev_ptr[L + 1];
        ps.chain_off = -1;
        pend.push_back(ps);
        ++L;
        continue;
      }
      int L2 = L;
      while (L2 < nl && lev_ptr[L2 + 1] - lev_ptr[L2] <= single_wg_max) ++L2;
      PendingSeg ps{};
      ps.s.chain_kernel = 1;
      ps.s.nchain = 1;
      ps.chain_off = static_cast<int>(all_chain_lev.size());
      all_chain_lev.push_back(L);
      all_chain_lev.push_back(L2);
      pend.push_back(ps);
      L = L2;
    }
  }
  plan.perm.alloc(sizeof(int) * std::max(n, 1));
  plan.perm.upload(perm.data(), sizeof(int) * n);
  plan.lev_ptr.alloc(sizeof(int) * lev_ptr.size());
  plan.lev_ptr.upload(lev_ptr.data(), sizeof(int) * lev_ptr.size());
  plan.chain_lev.alloc(sizeof(int) * std::max<size_t>(all_chain_lev.size(), 2));
  plan.chain_lev.upload(all_chain_lev.data(), sizeof(int) * all_chain_lev.size());
  plan.segs.clear();
  for (auto& ps : pend) {
    if (ps.chain_off >= 0) ps.s.chain_lev = plan.chain_lev.as<int>() + ps.chain_off;
    plan.segs.push_back(ps.s);
  }
  plan.nlaunch = static_cast<int>(plan.segs.size());

  // Level-ordered layout: dependency entries in the order the sequential
  // solve subtracts them (ascending columns; descending for DU).
  std::vector<int> lbeg(static_cast<size_t>(n) + 1, 0), lcol;
  std::vector<V> lval, ldiag(static_cast<size_t>(std::max(n, 1)), V(0));
  lcol.reserve(cind.size());
  lval.reserve(cind.size());
  for (int idx = 0; idx < n; ++idx) {
    const int i = perm[idx];
    const int rb = rptr[i], re = rptr[i + 1];
    if (kind == TriKind::kLD) {
      for (int p = rb; p < re - 1; ++p) {
        lcol.push_back(cind[p]);
        lval.push_back(val[p]);
      }
      ldiag[idx] = val[re - 1];
    } else if (kind == TriKind::kL) {
      for (int p = rb; p < re; ++p) {
        lcol.push_back(cind[p]);
        lval.push_back(val[p]);
      }
    } else {
      for (int p = re - 1; p >= rb + 1; --p) {
        lcol.push_back(cind[p]);
        lval.push_back(val[p]);
      }
      ldiag[idx] = val[rb];
    }
    lbeg[idx + 1] = static_cast<int>(lcol.size());
  }
  plan.lrow.alloc(sizeof(int) * std::max(n, 1));
  plan.lrow.upload(perm.data(), sizeof(int) * n);
  plan.lbeg.alloc(sizeof(int) * lbeg.size());
  plan.lbeg.upload(lbeg.data(), sizeof(int) * lbeg.size());
  plan.lcol.alloc(sizeof(int) * std::max<size_t>(lcol.size(), 1));
  plan.lcol.upload(lcol.data(), sizeof(int) * lcol.size());
  plan.lval.alloc(sizeof(V) * std::max<size_t>(lval.size(), 1));
  plan.lval.upload(lval.data(), sizeof(V) * lval.size());
  plan.ldiag.alloc(sizeof(V) * ldiag.size());
  plan.ldiag.upload(ldiag.data(), sizeof(V) * ldiag.size());
  // Dependency entries as level-ordered positions, for the LDS-ring mode.
  std::vector<int> pos_of(static_cast<size_t>(std::max(n, 1)));
  for (int idx = 0; idx < n; ++idx) pos_of[perm[idx]] = idx;
  std::vector<int> lpos(lcol.size());
  for (size_t q = 0; q < lcol.size(); ++q) lpos[q] = pos_of[lcol[q]];
  // fixed-stride copies of the first 8 entries of each position
  constexpr int W = 8;
  std::vector<int> ecol(static_cast<size_t>(std::max(n, 1)) * W, 0),
      epos(static_cast<size_t>(std::max(n, 1)) * W, 0);
  std::vector<V> eval(static_cast<size_t>(std::max(n, 1)) * W, V(0));
  for (int idx = 0; idx < n; ++idx)
    for (int q = 0; q < W && lbeg[idx] + q < lbeg[idx + 1]; ++q) {
      ecol[static_cast<size_t>(idx) * W + q] = lcol[lbeg[idx] + q];
      epos[static_cast<size_t>(idx) * W + q] = lpos[lbeg[idx] + q];
      eval[static_cast<size_t>(idx) * W + q] = lval[lbeg[idx] + q];
    }
  plan.ecol.alloc(sizeof(int) * ecol.size());
  plan.ecol.upload(ecol.data(), sizeof(int) * ecol.size());
  plan.epos.alloc(sizeof(int) * epos.size());
  plan.epos.upload(epos.data(), sizeof(int) * epos.size());
  plan.eval.alloc(sizeof(V) * eval.size());
  plan.eval.upload(eval.data(), sizeof(V) * eval.size());
  plan.lpos.alloc(sizeof(int) * std::max<size_t>(lpos.size(), 1));
  plan.lpos.upload(lpos.data(), sizeof(int) * lpos.size());
  // Integer level launches: lanes per row from the level's mean entry count
  // (the smallest power of two >= mean, at most 64); FP rows stay one lane,
  // which keeps the reference's summation order.
  if (std::is_same_v<V, int32_t>) {
    for (auto& sg : plan.segs) {
      if (sg.chain_kernel) continue;
      const int rows = sg.row_end - sg.row_begin;
      const double mean =
          rows > 0 ? static_cast<double>(lbeg[sg.row_end] - lbeg[sg.row_begin]) / rows : 0.0;
      int g = 1;
      while (g < 64 && g < mean) g *= 2;
      sg.group = g;
    }
  }
  dev::TriLayout& L = plan.layout;
  L.lrow = plan.lrow.as<int>();
  L.lbeg = plan.lbeg.as<int>();
  L.lcol = plan.lcol.as<int>();
  L.lval = plan.lval.get();
  L.ldiag = plan.ldiag.get();
  L.ecol = plan.ecol.as<int>();
  L.eval = plan.eval.get();
  L.lev_ptr = plan.lev_ptr.as<int>();
  L.lpos = plan.lpos.as<int>();
  L.epos = plan.epos.as<int>();
  // Ring mode needs every chain walked by one workgroup in a single launch and
  // every dependency within the ring: a row of a level ending at position e
  // reads position q only if e - q <= ring.
  const bool one_launch_chains =
      plan.segs.size() == 1 && plan.segs[0].chain_kernel == 1 &&
      (nchain > 1 || (all_chain_lev.size() == 2 && all_chain_lev[0] == 0 &&
                      all_chain_lev[1] == static_cast<int>(lev_ptr.size()) - 1));
  int ring = 1;
  while (static_cast<size_t>(ring) * 2 * sizeof(V) <= 128 * 1024) ring *= 2;
  long long span = 0;
  for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
    for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx)
      for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q)
        span = std::max<long long>(span, static_cast<long long>(lev_ptr[Lv + 1]) - lpos[q]);
  L.ring = (one_launch_chains && span <= ring) ? ring : 0;
  L.n = n;
  // Entry-parallel mode (integer factors): a dependency is read from the LDS
  // ring when the ring holds it, else from global memory, which needs it at
  // least kEpFarLevels levels back.  The ring is the smallest power of two
  // that holds every dependency, at most 32768 values.
  size_t maxlev_rows = 0;
  for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
    maxlev_rows = std::max<size_t>(maxlev_rows, lev_ptr[Lv + 1] - lev_ptr[Lv]);
  int maxlev_chain = 0;
  for (size_t c = 0; c + 1 < all_chain_lev.size(); c += 2)
    maxlev_chain = std::max(maxlev_chain, all_chain_lev[c + 1] - all_chain_lev[c]);
  int ep_ring = 2;
  while (ep_ring < 32768 && ep_ring < span) ep_ring *= 2;
  // Accumulator copies (LD factors only) where one instruction's lanes pile
  // onto few rows: rows averaging at least 8 dependency entries and levels
  // averaging at least 512 entries per chain (the dense coarse levels of the
  // SOR ordering, 16-45 % faster there); on sparser or narrower levels the
  // extra reads of the finish cost more than the collisions (BlockSOR's
  // levels measured 5-15 % slower with them).  SC23AD_EP_COPIES=0/1 forces
  // them off/on where they fit.
  const size_t offdiag = lcol.size();
  const size_t nlev_all = std::max<size_t>(lev_ptr.size() - 1, 1);
  const char* env_cp = std::getenv("SC23AD_EP_COPIES");
  bool copies = kind == TriKind::kLD && offdiag >= 8 * static_cast<size_t>(std::max(n, 1)) &&
                offdiag >= 512 * nlev_all;
  if (env_cp) copies = kind == TriKind::kLD && env_cp[0] == '1';
  const int acc_rows = static_cast<int>((maxlev_rows + 63) / 64 * 64);
  if (copies && dev::ep_lds_bytes(ep_ring, maxlev_chain, acc_rows * dev::kEpCopies) >
                    dev::kTriBlockedMaxLds)
    copies = false;
  const int ep_acc = copies ? acc_rows * dev::kEpCopies : dev::kEpMaxRows;
  if (std::is_same_v<V, int32_t> && plan.identity && one_launch_chains &&
      maxlev_rows <= static_cast<size_t>(dev::kEpMaxRows) &&
      dev::ep_lds_bytes(ep_ring, maxlev_chain, ep_acc) <= dev::kTriBlockedMaxLds) {
    std::vector<int> lev_of(static_cast<size_t>(std::max(n, 1)), 0);
    std::vector<int> lev_eptr(lev_ptr.size(), 0);
    for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv) {
      lev_eptr[Lv] = lbeg[lev_ptr[Lv]];
      for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx) lev_of[idx] = static_cast<int>(Lv);
    }
    lev_eptr.back() = lbeg[n];
    std::vector<int> lent(std::max<size_t>(lcol.size(), 1) * 4, 0);
    bool fits = true, any_far = false;
    for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
      for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx)
        for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q) {
          const int dpos = lpos[q];
          const bool in_ring = static_cast<long long>(lev_ptr[Lv + 1]) - dpos <= ep_ring;
          const bool far = !in_ring && lev_of[idx] - lev_of[dpos] >= dev::kEpFarLevels;
          if (!in_ring && !far) fits = false;
          any_far = any_far || far;
          lent[4 * static_cast<size_t>(q)] = far ? -(lcol[q] + 1) : dpos;
          lent[4 * static_cast<size_t>(q) + 1] = static_cast<int>(lval[q]);
          lent[4 * static_cast<size_t>(q) + 2] = idx - lev_ptr[Lv];
        }
    if (fits) {
      plan.lent.alloc(sizeof(int) * lent.size());
      plan.lent.upload(lent.data(), sizeof(int) * lent.size());
      plan.lev_eptr.alloc(sizeof(int) * lev_eptr.size());
      plan.lev_eptr.upload(lev_eptr.data(), sizeof(int) * lev_eptr.size());
      L.lent = plan.lent.get();
      L.lev_eptr = plan.lev_eptr.as<int>();
      L.ring = ep_ring;
      // Waves: every row of a level held in registers, and at most one level
      // in 16 with more entries than the registers hold (the rest loop over
      // global memory, one round trip each).
      auto overflowing = [&](int waves, int per_lane) {
        size_t c = 0;
        const size_t cap = static_cast<size_t>(waves) * 64 * per_lane;
        for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
          if (static_cast<size_t>(lev_eptr[Lv + 1] - lev_eptr[Lv]) > cap) ++c;
        return c;
      };
      const size_t nlev = lev_ptr.size() - 1;
      int w = 1;
      while (w < 8 && static_cast<size_t>(w) * 64 * 2 < maxlev_rows) w *= 2;
      while (w < 8 && overflowing(w, dev::kEpEntriesPerLane) * 16 > nlev) w *= 2;
      if (const char* cap = std::getenv("SC23AD_EP_MAXWAVES"))  // tuning experiments
        while (w > 1 && w > std::atoi(cap)) w /= 2;
      if (const char* lo = std::getenv("SC23AD_EP_MINWAVES"))
        while (w < 8 && w < std::atoi(lo) && static_cast<size_t>(w) * 64 < maxlev_rows) w *= 2;
      L.ep_waves = w;
      L.ep_far = any_far ? 1 : 0;
      L.ep_maxlev = maxlev_chain;
      L.ep_copies = copies ? dev::kEpCopies : 1;
      L.ep_acc = ep_acc;
      // Narrow register layout (4 entries, 1 row per lane) when it holds
      // every row of a level and about as many entries.
      L.ep_narrow = (maxlev_rows <= static_cast<size_t>(w) * 64 && overflowing(w, 4) * 16 <= nlev)
                        ? 1 : 0;
      // Row-parallel mode for smoother factors with short rows: one lane per
      // row for at most kRpMaxEntries dependencies, else (no far entries)
      // 2-16 lanes of 8, 10 or 12 entries per row; no LDS accumulators.
      // SC23AD_TRI_RP=0 keeps the entry-parallel kernel, SC23AD_TRI_RP=1
      // uses one lane per row only; SC23AD_RP_ROWS=2 puts two rows on each
      // lane of the 4-entry layout (default one: at 128^3 level 0 took 157 us
      // with 8 waves of one row per lane, 174 us with 4 waves of two).
      int md = 0;
      for (int idx = 0; idx < n; ++idx) md = std::max(md, lbeg[idx + 1] - lbeg[idx]);
      const char* env_rp = std::getenv("SC23AD_TRI_RP");
      int e = 0, lanes = 1, rows = 1;
      const bool force_groups = env_rp && env_rp[0] == '2';  // tuning: lanes per row even for short rows
      if (md <= dev::kRpMaxEntries && !(force_groups && !any_far && md > 4)) {
        e = md <= 4 ? 4 : md <= 8 ? 8 : md <= 10 ? 10 : 12;
        if (const char* r = std::getenv("SC23AD_RP_ROWS"); r && e == 4 && r[0] == '2') rows = 2;
      } else if (!any_far && !(env_rp && env_rp[0] == '1')) {
        // the fewest waves, then the fewest entries per row
        size_t best = 0;
        for (int g = 2; g <= dev::kRpMaxLanes; g *= 2)
          for (int ge : {8, 10, 12}) {
            if (g * ge < md) continue;
            const size_t w = (maxlev_rows * g + 63) / 64;
            const size_t key = w * 4096 + static_cast<size_t>(g * ge);
            if (w <= 8 && (best == 0 || key < best)) {
              best = key;
              lanes = g;
              e = ge;
            }
          }
      }
      if (kind != TriKind::kLD || (env_rp && env_rp[0] == '0') ||
          dev::rp_lds_bytes(ep_ring, maxlev_chain) > dev::kTriBlockedMaxLds)
        e = 0;
      const size_t lanes_needed = (maxlev_rows + rows - 1) / rows * lanes;
      // the fewest waves, 1 to 8 (SC23AD_RP_POW2=1: powers of two; at 128^3
      // three waves instead of four took levels 1 and 3 from 94 and 109 us
      // to 87 and 95 us per solve)
      const bool pow2 = std::getenv("SC23AD_RP_POW2") != nullptr;
      int rw = 1;
      while (rw < 8 && static_cast<size_t>(rw) * 64 < lanes_needed) rw = pow2 ? rw * 2 : rw + 1;
      if (e > 0 && static_cast<size_t>(rw) * 64 >= lanes_needed) {
        std::vector<int> rent(static_cast<size_t>(std::max(n, 1)) * lanes * e * 2, 0);
        int farq = 0;
        for (int idx = 0; idx < n; ++idx) {
          // far entries first (a row's integer sum is order-free), so the
          // kernel loads far values for the first farq slots only
          std::vector<int> ord;
          for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q)
            if (lent[4 * static_cast<size_t>(q)] < 0) ord.push_back(q);
          farq = std::max(farq, static_cast<int>(ord.size()));
          for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q)
            if (lent[4 * static_cast<size_t>(q)] >= 0) ord.push_back(q);
          for (int m = 0; m < static_cast<int>(ord.size()); ++m) {
            const int q = ord[m];
            const size_t at =
                ((static_cast<size_t>(idx) * lanes + m % lanes) * e + m / lanes) * 2;
            rent[at] = lent[4 * static_cast<size_t>(q)];
            rent[at + 1] = lent[4 * static_cast<size_t>(q) + 1];
          }
        }
        plan.rent.alloc(sizeof(int) * rent.size());
        plan.rent.upload(rent.data(), sizeof(int) * rent.size());
        L.rent = plan.rent.get();
        L.rp_e = e;
        L.rp_rows = rows;
        L.rp_lanes = lanes;
        L.rp_waves = rw;
        L.rp_farq = farq;
      }
      // The block kernel's layout (four right-hand sides per workgroup,
      // block.hip), chosen on its own: its ring keeps four values per slot,
      // so it is at most kTriBlockedMaxLds / 16 slots, and it reads
      // dependencies at least kRpbFarLevels levels back from global memory;
      // one or two rows per lane, so that wide levels fit in 8 waves.
      if (kind == TriKind::kLD) {
        int md_b = 0;
        for (int idx = 0; idx < n; ++idx) md_b = std::max(md_b, lbeg[idx + 1] - lbeg[idx]);
        long long span_near = 0, span_all = 0;
        for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
          for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx)
            for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q) {
              const long long d = static_cast<long long>(lev_ptr[Lv + 1]) - lpos[q];
              span_all = std::max(span_all, d);
              if (lev_of[idx] - lev_of[lpos[q]] < dev::kRpbFarLevels)
                span_near = std::max(span_near, d);
            }
        auto pow2 = [](long long v) {
          int r = 2;
          while (r < v) r *= 2;
          return r;
        };
        // far entries per row for a ring of rb slots (one-lane layouts)
        auto far_count = [&](int rb) {
          int mx = 0;
          for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
            for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx) {
              int c = 0;
              for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q)
                if (static_cast<long long>(lev_ptr[Lv + 1]) - lpos[q] > rb) ++c;
              mx = std::max(mx, c);
            }
          return mx;
        };
        int best_e = 0, best_g = 0, best_r = 0, best_w = 0, best_ring = 0, best_fq = 0;
        size_t best_key = 0;
        for (int g : {1, 2, 4, 8, 16})
          for (int ge : {4, 8, 10, 12}) {
            if ((g > 1 && ge == 4) || g * ge < md_b) continue;
            // the whole span in the ring when it fits, else the largest ring
            // that fits (at least the near span), the rest read as far
            // dependencies
            int rb = pow2(span_all);
            if (dev::rpb_lds_bytes(rb, maxlev_chain) > dev::kTriBlockedMaxLds) {
              rb = pow2(span_near);
              while (dev::rpb_lds_bytes(2 * rb, maxlev_chain) <= dev::kTriBlockedMaxLds) rb *= 2;
            }
            if (dev::rpb_lds_bytes(rb, maxlev_chain) > dev::kTriBlockedMaxLds) continue;
            int fq = 0;
            if (span_all > rb) {
              fq = far_count(rb);
              // one lane: any number of far slots; g lanes: up to four per
              // lane (eight with 10 or 12 entries)
              if (fq > (g == 1 ? ge : (ge >= 10 ? 8 : 4) * g)) continue;
            }
            for (int rows_b : {1, 2}) {
              if (rows_b == 2 && g > 1 && ge > 8) continue;  // no such block kernel
              // two rows per lane: 1 (4 entries) or 2 far slots, one lane
              if (rows_b == 2 && fq > 0 && (g > 1 || fq > (ge == 4 ? 1 : 2))) continue;
              const size_t lanes = (maxlev_rows + rows_b - 1) / rows_b * g;
              const size_t w = (lanes + 63) / 64;
              if (w > 8) continue;
              const size_t key = (w * 4096 + static_cast<size_t>(g * ge)) * 4 + rows_b;
              if (best_e == 0 || key < best_key) {
                best_key = key;
                best_e = ge;
                best_g = g;
                best_r = rows_b;
                best_w = static_cast<int>(w);
                best_ring = rb;
                best_fq = fq;
              }
            }
          }
        if (best_e == 0 && std::getenv("SC23AD_VERBOSE")) {
          int rb_near = pow2(span_near);
          while (dev::rpb_lds_bytes(2 * rb_near, maxlev_chain) <= dev::kTriBlockedMaxLds)
            rb_near *= 2;
          std::fprintf(stderr,
                       "[sc23ad plan] no block layout: max deps %d, span %lld (near %lld), "
                       "max level rows %d, far entries per row %d at ring %d\n",
                       md_b, span_all, span_near, static_cast<int>(maxlev_rows),
                       far_count(rb_near), rb_near);
        }
        if (best_e > 0) {
          const int e_b = best_e, g_b = best_g;
          std::vector<int> rb(static_cast<size_t>(std::max(n, 1)) * g_b * e_b * 2, 0);
          for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
            for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx) {
              std::vector<int> code, vals;  // far entries first (one-lane layouts)
              for (int pass = 0; pass < 2; ++pass)
                for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q) {
                  const bool far =
                      static_cast<long long>(lev_ptr[Lv + 1]) - lpos[q] > best_ring;
                  if (far != (pass == 0)) continue;
                  code.push_back(far ? -(lcol[q] + 1) : lpos[q]);
                  vals.push_back(static_cast<int>(lval[q]));
                }
              for (size_t mm = 0; mm < code.size(); ++mm) {
                const size_t at =
                    ((static_cast<size_t>(idx) * g_b + mm % g_b) * e_b + mm / g_b) * 2;
                rb[at] = code[mm];
                rb[at + 1] = vals[mm];
              }
            }
          plan.rent_blk.alloc(sizeof(int) * rb.size());
          plan.rent_blk.upload(rb.data(), sizeof(int) * rb.size());
          L.rent_blk = plan.rent_blk.get();
          L.ring_blk = best_ring;
          L.far_blk = best_fq > 0 ? 1 : 0;
          L.farq_blk = best_fq;
          L.e_blk = e_b;
          L.lanes_blk = g_b;
          L.rows_blk = best_r;
          L.waves_blk = best_w;
          L.ep_maxlev = maxlev_chain;
        }
      }
    }
  }
  // FP smoother factors (solvers 0 and 1): the row-parallel kernel with one
  // lane per row keeps each row's stored order, so it applies to rows of at
  // most 8 dependencies, with a ring of at most 128 KB and the same far rule.
  // SC23AD_TRI_RP=0 keeps the level-pipelined kernel.
  if constexpr (!std::is_same_v<V, int32_t>) {
    int md = 0;
    for (int idx = 0; idx < n; ++idx) md = std::max(md, lbeg[idx + 1] - lbeg[idx]);
    const char* env_rp = std::getenv("SC23AD_TRI_RP");
    int fring = 2;
    while (fring * 2 * sizeof(V) <= 128 * 1024 && fring < span) fring *= 2;
    size_t lds = sizeof(V) * static_cast<size_t>(fring) +
                 sizeof(int) * (static_cast<size_t>(maxlev_chain) + 1);
    // one row per lane, or two where a level has more than 512 rows
    // (SC23AD_RP_ROWS=2: two rows per lane wherever the layout applies)
    const char* env_rows = std::getenv("SC23AD_RP_ROWS");
    const int fp_rows = (maxlev_rows > 8 * 64 || (env_rows && env_rows[0] == '2')) ? 2 : 1;
    if (kind == TriKind::kLD && plan.identity && one_launch_chains && md <= 8 &&
        maxlev_rows <= static_cast<size_t>(fp_rows) * 8 * 64 && lds <= dev::kTriBlockedMaxLds &&
        !(env_rp && env_rp[0] == '0')) {
      std::vector<int> lev_of(static_cast<size_t>(std::max(n, 1)), 0);
      for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
        for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx)
          lev_of[idx] = static_cast<int>(Lv);
      const int e = md <= 4 ? 4 : 8;
      std::vector<int> rdep(static_cast<size_t>(std::max(n, 1)) * e, 0),
          rcnt(static_cast<size_t>(std::max(n, 1)), 0);
      std::vector<V> rval(rdep.size(), V(0));
      bool fits = true, any_far = false;
      for (size_t Lv = 0; Lv + 1 < lev_ptr.size() && fits; ++Lv)
        for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx) {
          rcnt[idx] = lbeg[idx + 1] - lbeg[idx];
          for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q) {
            const int dpos = lpos[q];
            const bool in_ring = static_cast<long long>(lev_ptr[Lv + 1]) - dpos <= fring;
            const bool far = !in_ring && lev_of[idx] - lev_of[dpos] >= dev::kEpFarLevels;
            if (!in_ring && !far) fits = false;
            any_far = any_far || far;
            const size_t at = static_cast<size_t>(idx) * e + (q - lbeg[idx]);
            rdep[at] = far ? -(lcol[q] + 1) : dpos;
            rval[at] = lval[q];
          }
        }
      if (fits) {
        plan.rpf_dep.alloc(sizeof(int) * rdep.size());
        plan.rpf_dep.upload(rdep.data(), sizeof(int) * rdep.size());
        plan.rpf_val.alloc(sizeof(V) * rval.size());
        plan.rpf_val.upload(rval.data(), sizeof(V) * rval.size());
        plan.rpf_cnt.alloc(sizeof(int) * rcnt.size());
        plan.rpf_cnt.upload(rcnt.data(), sizeof(int) * rcnt.size());
        L.rpf_dep = plan.rpf_dep.as<int>();
        L.rpf_val = plan.rpf_val.get();
        L.rpf_cnt = plan.rpf_cnt.as<int>();
        L.ring = fring;
        L.ep_far = any_far ? 1 : 0;
        L.ep_maxlev = maxlev_chain;
        L.rp_e = e;
        L.rp_rows = fp_rows;
        int rw = 1;
        while (rw < 8 && static_cast<size_t>(rw) * 64 * fp_rows < maxlev_rows) rw *= 2;
        L.rp_waves = rw;
      }
    } else if (kind == TriKind::kLD && plan.identity && one_launch_chains && md > 8 &&
               md <= 16 * 8 && lds <= dev::kTriBlockedMaxLds &&
               !(env_rp && env_rp[0] == '0')) {
      // several lanes per row (k_tri_chain_rpg): kG chunks of kE entries, the
      // fewest waves (at most 8), then the fewest padded slots; every
      // dependency in the ring, or (rows over 2 or 4 lanes, shuffled) the
      // far rule of the one-lane kernel
      // FP32 chunks of 16 entries only where 4 or 8 do not fit in 8 waves
      // (wide levels of long rows: at 160^3 the FP32 solve went from 0.55 to
      // 0.43 s, while FP64 rows of 16 measured slower than the
      // level-pipelined kernel; SC23AD_RP_E16=1, tests: wherever they fit)
      const bool e16_first = std::getenv("SC23AD_RP_E16") != nullptr;
      int best_g = 0, best_e = 0;
      size_t best_key = 0;
      for (int g = 2; g <= 16; g *= 2)
        for (int ge : {4, 8, 16}) {
          if (g * ge < md || (ge == 16 && sizeof(V) == 8)) continue;
          const size_t w = (maxlev_rows * g + 63) / 64;
          size_t key = w * 4096 + static_cast<size_t>(g * ge);
          if ((ge == 16) != e16_first) key += static_cast<size_t>(1) << 30;
          if (w <= 8 && (best_g == 0 || key < best_key)) {
            best_key = key;
            best_g = g;
            best_e = ge;
          }
        }
      bool fits = best_g > 0;
      // FP64 rows over 16 lanes stage their products in LDS when the ring
      // covering the span and the staging area fit together; every other
      // row shuffles them to its lanes, which measured faster for FP32 rows
      // (one shuffle per product) and for FP64 rows over at most 8 lanes
      bool stage = false;
      if (fits && sizeof(V) == 8 && best_g == 16) {
        const int w = static_cast<int>((maxlev_rows * best_g + 63) / 64);
        stage = dev::rpg_lds_bytes(sizeof(V), fring, w, best_e, maxlev_chain, true) <=
                dev::kTriBlockedMaxLds;
        // FP64 rows of 16 chunks of 8 spill their registers when shuffled
        if (!stage && sizeof(V) == 8 && best_g == 16 && best_e == 8) fits = false;
      }
      const bool far_ok = !stage && best_g <= 4;
      std::vector<int> lev_of(static_cast<size_t>(std::max(n, 1)), 0);
      for (size_t Lv = 0; Lv + 1 < lev_ptr.size(); ++Lv)
        for (int idx = lev_ptr[Lv]; idx < lev_ptr[Lv + 1]; ++idx)
          lev_of[idx] = static_cast<int>(Lv);
      // SC23AD_FP_RING_FAR=1 (tests): the smallest ring the far rule admits,
      // so that smaller factors take the far path too
      if (far_ok && std::getenv("SC23AD_FP_RING_FAR")) {
        long long need = 1;
        for (int idx = 0; idx < n; ++idx)
          for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q)
            if (lev_of[idx] - lev_of[lpos[q]] < dev::kEpFarLevels)
              need = std::max(need, static_cast<long long>(lev_ptr[lev_of[idx] + 1]) - lpos[q]);
        int r = 2;
        while (r < need) r *= 2;
        fring = std::min(fring, r);
      }
      // a dependency outside the ring is read from global memory when it is
      // at least kEpFarLevels levels back
      auto is_far = [&](int idx, int dpos) {
        return static_cast<long long>(lev_ptr[lev_of[idx] + 1]) - dpos > fring;
      };
      bool any_far = false;
      for (int idx = 0; idx < n && fits; ++idx)
        for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q)
          if (is_far(idx, lpos[q])) {
            if (far_ok && lev_of[idx] - lev_of[lpos[q]] >= dev::kEpFarLevels)
              any_far = true;
            else
              fits = false;
          }
      if (fits) {
        const size_t slots = static_cast<size_t>(best_g) * best_e;
        std::vector<int> rdep(static_cast<size_t>(std::max(n, 1)) * slots, 0),
            rcnt(static_cast<size_t>(std::max(n, 1)), 0);
        std::vector<V> rval(rdep.size(), V(0));
        for (int idx = 0; idx < n; ++idx) {
          rcnt[idx] = lbeg[idx + 1] - lbeg[idx];
          for (int q = lbeg[idx]; q < lbeg[idx + 1]; ++q) {
            const size_t at = static_cast<size_t>(idx) * slots + (q - lbeg[idx]);
            // entry m: lane m / kE, slot m % kE; a far one names -(position + 1)
            rdep[at] = is_far(idx, lpos[q]) ? -(lpos[q] + 1) : lpos[q];
            rval[at] = lval[q];
          }
        }
        plan.rpf_dep.alloc(sizeof(int) * rdep.size());
        plan.rpf_dep.upload(rdep.data(), sizeof(int) * rdep.size());
        plan.rpf_val.alloc(sizeof(V) * rval.size());
        plan.rpf_val.upload(rval.data(), sizeof(V) * rval.size());
        plan.rpf_cnt.alloc(sizeof(int) * rcnt.size());
        plan.rpf_cnt.upload(rcnt.data(), sizeof(int) * rcnt.size());
        L.rpf_dep = plan.rpf_dep.as<int>();
        L.rpf_val = plan.rpf_val.get();
        L.rpf_cnt = plan.rpf_cnt.as<int>();
        L.ring = fring;
        L.ep_far = any_far ? 1 : 0;
        L.ep_maxlev = maxlev_chain;
        L.rp_e = best_e;
        L.rp_lanes = best_g;
        L.rp_waves = static_cast<int>((maxlev_rows * best_g + 63) / 64);
        L.rp_stage = stage ? 1 : 0;
      }
    }
  }
  int maxdeps = 0;
  for (int idx = 0; idx < n; ++idx) maxdeps = std::max(maxdeps, lbeg[idx + 1] - lbeg[idx]);
  if (std::getenv("SC23AD_VERBOSE")) {
    std::fprintf(stderr,
                 "[sc23ad plan] kind %d n %d nnz %zu chains %d levels(max chain) %d "
                 "max level rows %zu launches %d ring %d span %lld max deps %d ep waves %d far %d "
                 "narrow %d copies %d rp waves %d entries %d rows %d lanes %d identity %d "
                 "| block: waves %d entries %d rows %d lanes %d ring %d far %d\n",
                 static_cast<int>(kind), n, cind.size(), nchain, plan.nlevels, maxlev_rows,
                 plan.nlaunch, L.ring, span, maxdeps, L.ep_waves, L.ep_far, L.ep_narrow,
                 L.ep_copies, L.rp_waves, L.rp_e, L.rp_rows, L.rp_lanes, plan.identity ? 1 : 0,
                 L.waves_blk, L.e_blk, L.rows_blk, L.lanes_blk, L.ring_blk, L.farq_blk);
  }
}

// Per-row block boundaries of the dependency entries for the blocked
// single-workgroup solve; returns false when the factor does not fit in LDS.
bool build_blocked(TriPlan& plan, TriKind kind, int n, const std::vector<int>& rptr,
                   const std::vector<int>& cind, size_t elem, size_t acc) {
  if (n <= 0 || dev::tri_blocked_lds(n, elem, acc) > dev::kTriBlockedMaxLds) return false;
  const int W = dev::kTriBlockWidth;
  const int nb = (n + W - 1) / W;
  std::vector<int> bp(static_cast<size_t>(n) * (nb + 1));
  for (int i = 0; i < n; ++i) {
    int b0 = rptr[i], e0 = rptr[i + 1];
    if (kind == TriKind::kLD) e0 = std::max(b0, e0 - 1);
    if (kind == TriKind::kDU) b0 = std::min(e0, b0 + 1);
    int p = b0;
    for (int b = 0; b <= nb; ++b) {
      const int lim = b * W;
      while (p < e0 && cind[p] < lim) ++p;
      bp[static_cast<size_t>(i) * (nb + 1) + b] = p;
    }
  }
  plan.bptr.alloc(sizeof(int) * bp.size());
  plan.bptr.upload(bp.data(), sizeof(int) * bp.size());
  plan.nb = nb;
  plan.blocked = true;
  plan.nlaunch = 1;
  return true;
}

// Dense 64-wide column-block slices of an integer triangular factor.
// Tiles of an integer coarse factor for the one-workgroup-per-block solve:
// for block b, the 64 x 64 tiles of the blocks it depends on in processing
// order, then its diagonal tile; column-major inside a tile.
void build_tiles(TriPlan& plan, TriKind kind, int n, const std::vector<int>& rptr,
                 const std::vector<int>& cind, const std::vector<int32_t>& val) {
  const int W = dev::kTriBlockWidth;
  const int nb = (n + W - 1) / W;
  if (n <= 0 || nb > dev::kTileMaxBlocks) return;
  // the tiles replay the sequential order only for rows stored in ascending
  // column order (as the factorisation writes them); otherwise keep the
  // blocked solve, which follows the stored order as it is
  for (int i = 0; i < n; ++i)
    for (int p = rptr[i] + 1; p < rptr[i + 1]; ++p)
      if (cind[p] <= cind[p - 1]) return;
  const bool fwd = kind != TriKind::kDU;
  std::vector<long long> toff(nb + 1, 0);
  for (int b = 0; b < nb; ++b) toff[b + 1] = toff[b] + static_cast<long long>(fwd ? b + 1 : nb - b) * W * W;
  std::vector<int32_t> tl(static_cast<size_t>(toff[nb]), 0);
  std::vector<int32_t> dv(static_cast<size_t>(n), 0);
  for (int i = 0; i < n; ++i) {
    int b0 = rptr[i], e0 = rptr[i + 1];
    if (kind == TriKind::kLD) {
      dv[i] = val[e0 - 1];
      e0 = std::max(b0, e0 - 1);
    }
    if (kind == TriKind::kDU) {
      dv[i] = val[b0];
      b0 = std::min(e0, b0 + 1);
    }
    const int b = i / W;
    for (int p = b0; p < e0; ++p) {
      const int c = cind[p];
      const int bc = c / W;
      if (fwd ? bc > b : bc < b)
        throw DevError{1, "coarse factor entry on the wrong side of the diagonal"};
      const int slot = fwd ? bc : nb - 1 - bc;  // tile position in processing order
      tl[static_cast<size_t>(toff[b]) + static_cast<size_t>(slot) * W * W +
         static_cast<size_t>(c - bc * W) * W + (i - b * W)] = val[p];
    }
  }
  plan.slice.alloc(sizeof(int32_t) * tl.size());
  plan.slice.upload(tl.data(), sizeof(int32_t) * tl.size());
  plan.boff.alloc(sizeof(long long) * toff.size());
  plan.boff.upload(toff.data(), sizeof(long long) * toff.size());
  plan.ddinv.alloc(sizeof(int32_t) * dv.size());
  plan.ddinv.upload(dv.data(), sizeof(int32_t) * dv.size());
  plan.nb = nb;
  plan.dense = true;
}

// FP version of build_tiles: the same 64 x 64 tiles in processing order with
// FP values, plus for every tile row a 64-bit mask of the columns the factor
// stores, so that the tiled FP solve skips absent entries exactly as the
// sequential CSR solve does.
template <typename V>
void build_tiles_fp(TriPlan& plan, TriKind kind, int n, const std::vector<int>& rptr,
                    const std::vector<int>& cind, const std::vector<V>& val) {
  const int W = dev::kTriBlockWidth;
  const int nb = (n + W - 1) / W;
  if (n <= 0 || nb > dev::kTileMaxBlocks) return;
  // the tiles replay the sequential order only for rows stored in ascending
  // column order (as the factorisation writes them); otherwise keep the
  // blocked solve, which follows the stored order as it is
  for (int i = 0; i < n; ++i)
    for (int p = rptr[i] + 1; p < rptr[i + 1]; ++p)
      if (cind[p] <= cind[p - 1]) return;
  const bool fwd = kind != TriKind::kDU;
  std::vector<long long> toff(nb + 1, 0);
  for (int b = 0; b < nb; ++b)
    toff[b + 1] = toff[b] + static_cast<long long>(fwd ? b + 1 : nb - b) * W * W;
  std::vector<V> tl(static_cast<size_t>(toff[nb]), V(0));
  std::vector<uint64_t> mk(static_cast<size_t>(toff[nb] / W), 0);
  std::vector<V> dv(static_cast<size_t>(n), V(0));
  for (int i = 0; i < n; ++i) {
    int b0 = rptr[i], e0 = rptr[i + 1];
    if (kind == TriKind::kLD) {
      dv[i] = val[e0 - 1];
      e0 = std::max(b0, e0 - 1);
    }
    if (kind == TriKind::kDU) {
      dv[i] = val[b0];
      b0 = std::min(e0, b0 + 1);
    }
    const int b = i / W;
    for (int p = b0; p < e0; ++p) {
      const int c = cind[p];
      const int bc = c / W;
      if (fwd ? bc > b : bc < b)
        throw DevError{1, "coarse factor entry on the wrong side of the diagonal"};
      const int slot = fwd ? bc : nb - 1 - bc;
      const size_t tile = static_cast<size_t>(toff[b]) / (W * W) + slot;
      tl[tile * W * W + static_cast<size_t>(c - bc * W) * W + (i - b * W)] = val[p];
      mk[tile * W + (i - b * W)] |= uint64_t{1} << (c - bc * W);
    }
  }
  plan.slice.alloc(sizeof(V) * tl.size());
  plan.slice.upload(tl.data(), sizeof(V) * tl.size());
  plan.boff.alloc(sizeof(long long) * toff.size());
  plan.boff.upload(toff.data(), sizeof(long long) * toff.size());
  plan.tmask.alloc(sizeof(uint64_t) * mk.size());
  plan.tmask.upload(mk.data(), sizeof(uint64_t) * mk.size());
  plan.ddinv.alloc(sizeof(V) * dv.size());
  plan.ddinv.upload(dv.data(), sizeof(V) * dv.size());
  plan.nb = nb;
  plan.dense_fp = true;
}

// Rows and columns of a CSR matrix in a new numbering: new row r is old row
// row_perm[r] (identity when null), columns map through col_ipos (old -> new,
// identity when null).  The entries of a row keep their stored order, so FP
// row sums add in the same order as before.
template <typename V>
void renumber(const std::vector<int>& rptr, const std::vector<int>& cind,
              const std::vector<V>& val, const std::vector<int>* row_perm,
              const std::vector<int>* col_ipos, std::vector<int>& r2, std::vector<int>& c2,
              std::vector<V>& v2) {
  const int n = static_cast<int>(rptr.size()) - 1;
  r2.assign(static_cast<size_t>(n) + 1, 0);
  c2.resize(cind.size());
  v2.resize(val.size());
  for (int r = 0; r < n; ++r) {
    const int i = row_perm ? (*row_perm)[r] : r;
    r2[r + 1] = r2[r] + (rptr[i + 1] - rptr[i]);
  }
  for (int r = 0; r < n; ++r) {
    const int i = row_perm ? (*row_perm)[r] : r;
    int q = r2[r];
    for (int p = rptr[i]; p < rptr[i + 1]; ++p, ++q) {
      c2[q] = col_ipos ? (*col_ipos)[cind[p]] : cind[p];
      v2[q] = val[p];
    }
  }
}

// Row boundaries of the smoother's chains: BlockSOR's b_num blocks of
// ceil(n / b_num) rows (smoother/stationary.hpp:934-939), or one chain.
std::vector<int> chain_rows(int n, bool blocks, int b_num) {
  if (!blocks || b_num <= 1) return {0, n};
  const int size = (n + b_num - 1) / b_num;
  std::vector<int> cb(static_cast<size_t>(b_num) + 1, 0);
  for (int c = 0; c < b_num; ++c) cb[c + 1] = std::min(n, cb[c] + size);
  cb[b_num] = n;
  return cb;
}

// Uploads the levels of a hierarchy (IntLevel or FpLevel<T>) renumbered into
// the LD smoother's level order: level l's unknowns take their position in
// the level schedule of LD_l, so every level-scheduled kernel reads and
// writes contiguous ranges.  R and P map between the numberings of adjacent
// levels; the coarsest grid keeps its own order.
template <typename Level, typename BitsOf>
void upload_levels(DevAmg& a, const std::vector<Level>& levels, bool blocks, int b_num,
                   const LaunchCfg& c, BitsOf bits) {
  const int depth = static_cast<int>(levels.size());
  std::vector<std::vector<int>> perm(depth), ipos(depth), cbeg(depth);
  for (int l = 0; l < depth; ++l) {
    const auto& ld = levels[l].ld;
    const std::vector<int> lev = tri_levels(TriKind::kLD, ld.n, ld.rptr, ld.cind);
    std::vector<int> lp, cl;
    order_by_level(lev, chain_rows(ld.n, blocks, b_num), perm[l], lp, cl);
    ipos[l].assign(static_cast<size_t>(ld.n), 0);
    for (int k = 0; k < ld.n; ++k) ipos[l][perm[l][k]] = k;
    for (size_t q = 0; q + 1 < cl.size(); q += 2) cbeg[l].push_back(lp[cl[q]]);
    cbeg[l].push_back(ld.n);
  }
  for (int l = 0; l < depth; ++l) {
    const Level& L = levels[l];
    DevLevel& D = a.lv[l];
    D.n = L.a.n;
    const std::vector<int>* rp = &perm[l];
    const std::vector<int>* ci = &ipos[l];
    const std::vector<int>* nrp = l + 1 < depth ? &perm[l + 1] : nullptr;
    const std::vector<int>* nci = l + 1 < depth ? &ipos[l + 1] : nullptr;
    using V = typename std::decay_t<decltype(L.a.val)>::value_type;
    std::vector<int> r2, c2;
    std::vector<V> v2;
    auto put = [&](DevMat& dm, const auto& m, const std::vector<int>* rows,
                   const std::vector<int>* cols, int group) {
      renumber(m.rptr, m.cind, m.val, rows, cols, r2, c2, v2);
      int b, db;
      bits(m, b, db);
      upload_mat(dm, m.n, m.m, b, db, r2, c2, v2, group);
    };
    put(D.a, L.a, rp, ci, c.spmv_group);
    put(D.u, L.u, rp, ci, c.spmv_group);
    put(D.r, L.r, nrp, ci, c.spmv_group);
    put(D.p, L.p, rp, nci, c.spmv_group);
    put(D.ld, L.ld, rp, ci, 0);  // last: r2/c2/v2 keep LD for the plan
    build_plan(D.ld_plan, TriKind::kLD, L.ld.n, r2, c2, v2, cbeg[l], c.single_wg_max_rows);
    D.perm.alloc(sizeof(int) * std::max(L.a.n, 1));
    D.perm.upload(perm[l].data(), sizeof(int) * L.a.n);
  }
  if (depth > 0) {
    const auto& a0 = levels[0].a;
    int b, db;
    bits(a0, b, db);
    upload_mat(a.a0, a0.n, a0.m, b, db, a0.rptr, a0.cind, a0.val, c.spmv_group);
  }
}

// SpMV launch-shape autotuner: times every candidate kernel for this
// matrix (integer CSR with 1..64 lanes per row, SELL-64 when built; FP: CSR
// thread per row, SELL-64) on the library stream and keeps the fastest.
// Every candidate computes the same result (integer sums are order-free; the
// FP kernels add a row's products in stored order), so the choice is free.
// SC23AD_SPMV_TUNE=0 keeps the defaults (lanes from nnz/row, SELL if built).
// Candidates are timed in the mode the V-cycle uses the matrix in (the fused
// subtract or add reads a second vector; the add runs in place).
void tune_spmv(DevMat& m, Elem e, SpmvMode mode) {
  const char* env = std::getenv("SC23AD_SPMV_TUNE");
  if ((env && env[0] == '0') || m.d.n <= 0) return;
  const size_t es = e == Elem::kF64 ? 8 : 4;
  DevBuf x(es * std::max(m.d.m, 1)), y(es * std::max(m.d.n, 1)), z(es * std::max(m.d.n, 1));
  hip_check(hipMemsetAsync(x.get(), 0, es * std::max(m.d.m, 1), lib_stream()), "memset");
  hip_check(hipMemsetAsync(y.get(), 0, es * std::max(m.d.n, 1), lib_stream()), "memset");
  hip_check(hipMemsetAsync(z.get(), 0, es * std::max(m.d.n, 1), lib_stream()), "memset");
  const void* zp = mode == SpmvMode::kPlain ? nullptr
                   : mode == SpmvMode::kAdd ? y.get()
                                            : z.get();
  std::vector<dev::DCsr> cands;
  dev::DCsr csr = m.d;
  csr.soff = nullptr;
  csr.srow = nullptr;
  if (e == Elem::kI32) {
    const double avg = static_cast<double>(m.d.nnz) / m.d.n;
    for (int g = 1; g <= 64; g *= 2) {
      if (g > 2 * avg + 2) break;
      csr.group = g;
      cands.push_back(csr);
    }
  } else {
    cands.push_back(csr);
  }
  if (m.d.soff) cands.push_back(m.d);
  if (m.d.soff && m.scol16.get()) {
    dev::DCsr c16 = m.d;
    c16.scol16 = m.scol16.as<int16_t>();
    cands.push_back(c16);
  }
  if (m.has2) cands.push_back(m.d2);
  // row-sorted SELL copies (rectangular R / P): also with XCD-contiguous rows
  for (size_t i = 0, n0 = cands.size(); i < n0; ++i)
    if (cands[i].soff && cands[i].srow && !std::getenv("SC23AD_XCD_REMAP")) {
      dev::DCsr c = cands[i];
      c.remap = 1;
      cands.push_back(c);
    }
  // SC23AD_SPMV_FORMAT=csr|sell|sell16 keeps only that format's candidates
  // where the matrix has it (tests run every kernel this way).
  if (const char* f = std::getenv("SC23AD_SPMV_FORMAT")) {
    const std::string want(f);
    std::vector<dev::DCsr> keep;
    for (const auto& c : cands) {
      const char* k = !c.soff                            ? "csr"
                      : c.scol16                         ? "sell16"
                      : (m.has2 && c.soff == m.d2.soff) ? "sellcol"
                                                         : "sell";
      if (want == k) keep.push_back(c);
    }
    if (!keep.empty()) cands = keep;
  }
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  float best = 1e30f;
  size_t best_i = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    for (int r = 0; r < 2; ++r)
      hip_check(dev::launch_spmv(e, cands[i], mode, x.get(), zp, y.get(), lib_stream()),
                "tune spmv");
    hip_check(hipEventRecord(e0, lib_stream()), "record");
    for (int r = 0; r < 5; ++r)
      hip_check(dev::launch_spmv(e, cands[i], mode, x.get(), zp, y.get(), lib_stream()),
                "tune spmv");
    hip_check(hipEventRecord(e1, lib_stream()), "record");
    hip_check(hipEventSynchronize(e1), "sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (std::getenv("SC23AD_VERBOSE"))
      std::fprintf(stderr, "[sc23ad spmv tune] n %d nnz %d %s%s%s%s lanes %d: %.2f us\n",
                   m.d.n, m.d.nnz, cands[i].soff ? "SELL" : "CSR",
                   (m.has2 && cands[i].soff == m.d2.soff) ? " column-ordered"
                   : cands[i].srow                        ? " sorted"
                                                          : "",
                   cands[i].scol16 ? " 16-bit" : "", cands[i].remap == 1 ? " XCD rows" : "",
                   cands[i].group, ms * 1000.0f / 5);
    if (ms < best) {
      best = ms;
      best_i = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  m.d = cands[best_i];
  if (m.soff.get() && m.d.soff != m.soff.as<long long>()) {
    m.soff = DevBuf();
    m.scol = DevBuf();
    m.sval = DevBuf();
    m.srow = DevBuf();
  }
  if (!m.d.scol16) m.scol16 = DevBuf();
  if (m.has2 && m.d.soff != m.soff2.as<long long>()) {
    m.soff2 = DevBuf();
    m.scol2 = DevBuf();
    m.sval2 = DevBuf();
    m.srow2 = DevBuf();
    m.has2 = false;
  }
}

}  // namespace

void DevAmg::tune() {
  for (DevLevel& L : lv) {
    tune_spmv(L.a, elem, SpmvMode::kSub);  // residual
    tune_spmv(L.u, elem, SpmvMode::kSub);  // smoother
    tune_spmv(L.r, elem, SpmvMode::kPlain);
    tune_spmv(L.p, elem, SpmvMode::kAdd);  // prolongation added in place
  }
  if (depth > 0) tune_spmv(a0, elem, SpmvMode::kPlain);  // the Arnoldi step's operator
}

std::unique_ptr<DevAmg::Work> DevAmg::make_work() const {
  auto w = std::make_unique<Work>();
  const size_t es = esize();
  w->ee.resize(depth + 1);
  w->xx.resize(depth + 1);
  w->rr.resize(depth + 1);
  w->tmp.resize(depth + 1);
  for (int l = 1; l <= depth; ++l) {
    const size_t n = static_cast<size_t>(std::max(level_n(l), 1));
    w->ee[l].alloc(es * n);
    w->xx[l].alloc(es * n);
    w->rr[l].alloc(es * n);
    w->tmp[l].alloc(es * n);
  }
  const size_t n0s = static_cast<size_t>(std::max(n0(), 1));
  w->r0.alloc(es * n0s);
  w->vin0.alloc(es * n0s);
  w->vout0.alloc(es * n0s);
  w->tmp[0].alloc(es * n0s);
  w->tmpc.alloc(es * std::max(n_coarse, 1));
  // published-value words of the tiled coarse solves: epoch 0 = never written
  std::vector<unsigned long long> zero(static_cast<size_t>(std::max(n_coarse, 1)), 0);
  for (DevBuf* f : {&w->lflags, &w->uflags}) {
    f->alloc(sizeof(unsigned long long) * zero.size());
    f->upload(zero.data(), sizeof(unsigned long long) * zero.size());
  }
  w->err.alloc(sizeof(int));
  w->err.upload(zero.data(), sizeof(int));
  w->tickets.alloc(4 * sizeof(unsigned));
  w->tickets.upload(zero.data(), 4 * sizeof(unsigned));
  return w;
}

std::unique_ptr<DevAmg> DevAmg::upload(const IntHierarchy& h, const LaunchCfg& c) {
  auto a = std::make_unique<DevAmg>();
  a->elem = Elem::kI32;
  a->cfg = c;
  a->depth = h.depth;
  a->iter = h.prm.iter;
  a->n_coarse = h.n_coarse;
  a->lv.resize(h.depth);
  const bool bj = h.prm.smoother == SmootherKind::kBlockSOR;
  upload_levels(*a, h.lv, bj, h.prm.b_num, c, [](const IntCsr& m, int& b, int& db) {
    b = m.bit;
    db = m.dbit;
  });
  upload_mat(a->lc, h.lc.n, h.lc.m, h.lc.bit, -1, h.lc.rptr, h.lc.cind, h.lc.val, 0);
  upload_mat(a->uc, h.uc.n, h.uc.m, h.uc.bit, h.uc.dbit, h.uc.rptr, h.uc.cind, h.uc.val, 0);
  build_plan(a->lc_plan, TriKind::kL, h.lc.n, h.lc.rptr, h.lc.cind, h.lc.val, {0, h.lc.n},
             c.single_wg_max_rows);
  build_plan(a->uc_plan, TriKind::kDU, h.uc.n, h.uc.rptr, h.uc.cind, h.uc.val, {0, h.uc.n},
             c.single_wg_max_rows);
  build_blocked(a->lc_plan, TriKind::kL, h.lc.n, h.lc.rptr, h.lc.cind, 4, 8);
  build_blocked(a->uc_plan, TriKind::kDU, h.uc.n, h.uc.rptr, h.uc.cind, 4, 8);
  if (12 * static_cast<size_t>(h.lc.n) + 16 <= dev::kTriBlockedMaxLds) {
    build_tiles(a->lc_plan, TriKind::kL, h.lc.n, h.lc.rptr, h.lc.cind, h.lc.val);
    build_tiles(a->uc_plan, TriKind::kDU, h.uc.n, h.uc.rptr, h.uc.cind, h.uc.val);
  }
  a->work_ = a->make_work();
  a->tune();
  return a;
}

template <typename T>
std::unique_ptr<DevAmg> DevAmg::upload_fp(const FpHierarchy<T>& h, const LaunchCfg& c) {
  auto a = std::make_unique<DevAmg>();
  a->elem = sizeof(T) == 8 ? Elem::kF64 : Elem::kF32;
  a->cfg = c;
  a->depth = h.depth;
  a->iter = h.prm.iter;
  a->n_coarse = h.n_coarse;
  a->lv.resize(h.depth);
  a->jacobi = h.prm.smoother == SmootherKind::kJacobi;
  const bool bj = h.prm.smoother == SmootherKind::kBlockSOR;
  LaunchCfg cf = c;
  cf.spmv_group = 1;
  upload_levels(*a, h.lv, bj, h.prm.b_num, cf, [](const FpCsr<T>&, int& b, int& db) {
    b = 0;
    db = -1;
  });
  upload_mat(a->lc, h.lc.n, h.lc.m, 0, -1, h.lc.rptr, h.lc.cind, h.lc.val, 1);
  upload_mat(a->uc, h.uc.n, h.uc.m, 0, -1, h.uc.rptr, h.uc.cind, h.uc.val, 1);
  build_plan(a->lc_plan, TriKind::kL, h.lc.n, h.lc.rptr, h.lc.cind, h.lc.val, {0, h.lc.n},
             c.single_wg_max_rows);
  build_plan(a->uc_plan, TriKind::kDU, h.uc.n, h.uc.rptr, h.uc.cind, h.uc.val, {0, h.uc.n},
             c.single_wg_max_rows);
  build_blocked(a->lc_plan, TriKind::kL, h.lc.n, h.lc.rptr, h.lc.cind, sizeof(T), sizeof(T));
  build_blocked(a->uc_plan, TriKind::kDU, h.uc.n, h.uc.rptr, h.uc.cind, sizeof(T), sizeof(T));
  if (!std::getenv("SC23AD_FP_COARSE_BLOCKED")) {
    build_tiles_fp(a->lc_plan, TriKind::kL, h.lc.n, h.lc.rptr, h.lc.cind, h.lc.val);
    build_tiles_fp(a->uc_plan, TriKind::kDU, h.uc.n, h.uc.rptr, h.uc.cind, h.uc.val);
  }
  a->work_ = a->make_work();
  a->tune();
  return a;
}

template std::unique_ptr<DevAmg> DevAmg::upload_fp<double>(const FpHierarchy<double>&,
                                                           const LaunchCfg&);
template std::unique_ptr<DevAmg> DevAmg::upload_fp<float>(const FpHierarchy<float>&,
                                                          const LaunchCfg&);

void DevAmg::spmv(const DevMat& m, SpmvMode mode, const void* x, const void* z,
                  void* y) const {
  hip_check(dev::launch_spmv(elem, m.d, mode, x, z, y, lib_stream()), "spmv");
}

// Whether trsv() solves the factor with one launch that binds armed
// dispatch events (the row- and entry-parallel integer kernels).
bool DevAmg::trsv_timing_aware(const TriPlan& plan) const {
  return elem == Elem::kI32 && !plan.dense && !plan.blocked && plan.segs.size() == 1 &&
         (plan.layout.rp_waves > 0 || plan.layout.ep_waves > 0);
}

void DevAmg::trsv(const TriPlan& plan, const DevMat& m, const void* in, void* out,
                  unsigned long long* flags, int* epoch, int* err, unsigned* tickets) const {
  if (plan.dense && elem == Elem::kI32 && flags) {
    *epoch = *epoch == INT32_MAX ? 1 : *epoch + 1;
    hip_check(dev::launch_trsv_tiles(plan.kind, m.d.n, plan.nb, plan.slice.as<int32_t>(),
                                     plan.boff.as<long long>(), plan.ddinv.as<int32_t>(),
                                     m.d.bit, m.d.dbit, in, out, flags, *epoch, err, tickets,
                                     lib_stream()),
              "trsv tiles");
    return;
  }
  if (plan.dense_fp && elem != Elem::kI32 && flags) {
    *epoch = *epoch == INT32_MAX ? 1 : *epoch + 1;
    hip_check(dev::launch_trsv_tiles_fp(elem, plan.kind, m.d.n, plan.nb, plan.slice.get(),
                                        plan.boff.as<long long>(), plan.tmask.as<uint64_t>(),
                                        plan.ddinv.get(), in, out, flags, *epoch, err, tickets,
                                        lib_stream()),
              "trsv tiles fp");
    return;
  }
  if (plan.blocked) {
    hip_check(dev::launch_trsv_blocked(elem, plan.kind, m.d, plan.bptr.as<int>(), plan.nb,
                                       in, out, lib_stream()),
              "trsv blocked");
    return;
  }
  if (plan.layout.rp_waves > 0 && elem == Elem::kI32 && plan.segs.size() == 1) {
    hip_check(dev::launch_trsv_rp(plan.layout, plan.segs[0], m.d.bit, m.d.dbit, in, out,
                                  lib_stream()),
              "trsv rp");
    return;
  }
  if (plan.layout.rpf_dep && elem != Elem::kI32 && plan.segs.size() == 1) {
    hip_check(dev::launch_trsv_rpf(elem, plan.layout, plan.segs[0], in, out, lib_stream()),
              "trsv rp fp");
    return;
  }
  if (plan.layout.ep_waves > 0 && elem == Elem::kI32 && plan.segs.size() == 1) {
    hip_check(dev::launch_trsv_ep(plan.kind, plan.layout, plan.segs[0], m.d.bit, m.d.dbit, in,
                                  out, lib_stream()),
              "trsv ep");
    return;
  }
  for (const auto& seg : plan.segs)
    hip_check(dev::launch_trsv_layout(elem, plan.kind, plan.layout, seg, m.d.bit, m.d.dbit,
                                      in, out, lib_stream()),
              "trsv");
}

// SOR / BlockSOR / JacobiSOR Smooth (Library/smoother/stationary.hpp:415-430,
// 997-1012, 647-660): inited: iter x { t = in - U out; out = LD^-1 t };
// otherwise out = LD^-1 in first.  Jacobi (stationary.hpp:50-74): LD^-1 is
// the diagonal d_inv, U is A, and the update is out += d_inv t.
void DevAmg::smooth(int l, const void* in, void* out, bool inited, Work& w) const {
  const DevLevel& L = lv[l];
  int k0 = 0;
  EventTimer* tm = l == 0 ? w.trsv_timer : nullptr;
  if (!inited) {
    if (tm) tm->begin(trsv_timing_aware(L.ld_plan));
    trsv(L.ld_plan, L.ld, in, out, nullptr, nullptr, nullptr, nullptr);
    if (tm) tm->end();
    k0 = 1;
  }
  void* t = w.tmp[l].get();
  for (int k = k0; k < iter; ++k) {
    spmv(L.u, SpmvMode::kSub, out, in, t);
    if (jacobi) {
      hip_check(dev::launch_jacobi_update(elem, L.n, L.ld.d.val, t, out, lib_stream()),
                "jacobi");
      continue;
    }
    if (tm) tm->begin(trsv_timing_aware(L.ld_plan));
    trsv(L.ld_plan, L.ld, t, out, nullptr, nullptr, nullptr, nullptr);
    if (tm) tm->end();
  }
}

void DevAmg::check_errors(Work& w) const {
  if (!lc_plan.dense && !uc_plan.dense && !lc_plan.dense_fp && !uc_plan.dense_fp) return;
  int e = 0;
  hip_check(hipMemcpyAsync(&e, w.err.get(), sizeof(int), hipMemcpyDeviceToHost, lib_stream()),
            "error flag");
  hip_check(hipStreamSynchronize(lib_stream()), "sync");
  if (e != 0) {
    // clear the word so that later calls on this work context start clean
    const int zero = 0;
    hip_check(hipMemcpyAsync(w.err.get(), &zero, sizeof(int), hipMemcpyHostToDevice,
                             lib_stream()),
              "error flag reset");
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    throw DevError{5, "coarse solve: a block waited too long for its dependencies"};
  }
}

// direct::LUflex / LU<T>::Solve (Library/solver/lu.hpp:91-94, 33-36).
void DevAmg::coarse(const void* in, void* out, Work& w) const {
  unsigned* tk = w.tickets.as<unsigned>();
  trsv(lc_plan, lc, in, w.tmpc.get(), w.lflags.as<unsigned long long>(), &w.lepoch,
       w.err.as<int>(), tk);
  trsv(uc_plan, uc, w.tmpc.get(), out, w.uflags.as<unsigned long long>(), &w.uepoch,
       w.err.as<int>(), tk + 2);
}

// intAMGflex::Precondition / dfAMG::Precondition (INTAMG.hpp:269-293,
// 108-133).  Coarse work vectors are indexed by level (1..depth).
void DevAmg::vcycle(const void* in, void* out, Work& w) const {
  const int n = n0();
  hip_check(dev::launch_permute(elem, n, perm_of(0), in, w.vin0.get(), true, lib_stream()),
            "permute in");
  vcycle_lo(w.vin0.get(), w.vout0.get(), w);
  hip_check(dev::launch_permute(elem, n, perm_of(0), w.vout0.get(), out, false, lib_stream()),
            "permute out");
}

void DevAmg::level_op(int level, int op, const void* in, void* out, Work& w, void* ti_p,
                      void* to_p) const {
  // Numbering of the operand and of the result (nullptr: original order).
  const int* pin = op == 7 ? nullptr : op == 3 ? perm_of(level + 1) : perm_of(level);
  const int* pout = op == 7 ? nullptr : op == 2 ? perm_of(level + 1) : perm_of(level);
  const int nin = op == 3 ? level_n(level + 1) : level_n(level);
  const int nout = op == 2 ? level_n(level + 1) : level_n(level);
  const size_t es = esize();
  struct Staging {  // caller's buffer, or one allocated for this call
    DevBuf own;
    void* p;
    Staging(void* given, size_t bytes) : p(given) {
      if (!p) {
        own.alloc(bytes);
        p = own.get();
      }
    }
    void* get() const { return p; }
  } ti(ti_p, es * std::max(nin, 1)), to(to_p, es * std::max(nout, 1));
  auto to_level = [&](const int* p, int n, const void* src, void* dst) {
    if (p) hip_check(dev::launch_permute(elem, n, p, src, dst, true, lib_stream()), "permute");
    else hip_check(hipMemcpyAsync(dst, src, es * n, hipMemcpyDeviceToDevice, lib_stream()), "copy");
  };
  to_level(pin, nin, in, ti.get());
  if (op == 6) to_level(pout, nout, out, to.get());  // Smooth(true) reads out too
  if (op == 7) {
    coarse(ti.get(), to.get(), w);
  } else {
    const DevLevel& L = lv[level];
    switch (op) {
      case 0: spmv(L.a, SpmvMode::kPlain, ti.get(), nullptr, to.get()); break;
      case 1: spmv(L.u, SpmvMode::kPlain, ti.get(), nullptr, to.get()); break;
      case 2: spmv(L.r, SpmvMode::kPlain, ti.get(), nullptr, to.get()); break;
      case 3: spmv(L.p, SpmvMode::kPlain, ti.get(), nullptr, to.get()); break;
      case 4: trsv(L.ld_plan, L.ld, ti.get(), to.get(), nullptr, nullptr, nullptr, nullptr); break;
      case 5: smooth(level, ti.get(), to.get(), false, w); break;
      case 6: smooth(level, ti.get(), to.get(), true, w); break;
      default: throw DevError{1, "level_op: bad op"};
    }
  }
  if (pout) hip_check(dev::launch_permute(elem, nout, pout, to.get(), out, false, lib_stream()), "permute");
  else hip_check(hipMemcpyAsync(out, to.get(), es * nout, hipMemcpyDeviceToDevice, lib_stream()), "copy");
}

// intAMGflex::Precondition / dfAMG::Precondition on level-ordered vectors.
void DevAmg::vcycle_lo(const void* in, void* out, Work& w) const {
  smooth(0, in, out, false, w);
  spmv(lv[0].a, SpmvMode::kSub, out, in, w.r0.get());  // r = in - A out
  if (depth == 0) return;
  spmv(lv[0].r, SpmvMode::kPlain, w.r0.get(), nullptr, w.ee[1].get());
  for (int l = 1; l < depth; ++l) {
    smooth(l, w.ee[l].get(), w.xx[l].get(), false, w);
    spmv(lv[l].a, SpmvMode::kSub, w.xx[l].get(), w.ee[l].get(), w.rr[l].get());
    spmv(lv[l].r, SpmvMode::kPlain, w.rr[l].get(), nullptr, w.ee[l + 1].get());
  }
  coarse(w.ee[depth].get(), w.xx[depth].get(), w);
  for (int l = depth - 1; l >= 1; --l) {
    spmv(lv[l].p, SpmvMode::kAdd, w.xx[l + 1].get(), w.xx[l].get(), w.xx[l].get());
    smooth(l, w.ee[l].get(), w.xx[l].get(), true, w);
  }
  spmv(lv[0].p, SpmvMode::kAdd, w.xx[1].get(), out, out);
  smooth(0, in, out, true, w);
}

double DevAmg::vcycle_bytes() const {
  const size_t es = esize();
  double b = 0;
  auto trsv_bytes = [&](const DevMat& m) {
    return static_cast<double>((es + 4) * m.d.nnz + 4 * (m.d.n + 1) + 2 * es * m.d.n);
  };
  for (int l = 0; l < depth; ++l) {
    const DevLevel& L = lv[l];
    const double n = L.n;
    // pre-smooth (LD solve) + post-smooth (U spmv-sub + LD solve)
    b += 2 * trsv_bytes(L.ld) + L.u.algo_bytes_spmv(es) + es * n;
    // residual (A spmv-sub), restriction, prolongation-add
    b += L.a.algo_bytes_spmv(es) + es * n;
    b += L.r.algo_bytes_spmv(es);
    b += L.p.algo_bytes_spmv(es) + es * n;
  }
  b += trsv_bytes(lc) + trsv_bytes(uc);
  return b;
}

// ---------------------------------------------------------------------------

void upload_fp64_operator(DevMat& dm, const HostCsr& a) {
  upload_mat(dm, a.n, a.m, 0, -1, a.rptr, a.cind, a.val, 1);
}

DevSolver::DevSolver(Elem e, const HostCsr& a, DevAmg* amg_, int m_, double eps_,
                     int bit_, int max_outer_)
    : elem(e), m(m_), eps(eps_), bit(bit_), max_outer(max_outer_), amg(amg_) {
  lib_stream();  // device set up
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  StreamScope scope(stream);
  work = amg->make_work();
  work->trsv_timer = &trsv0_timer;
  const int n = a.n;
  upload_mat(a64, a.n, a.m, 0, -1, a.rptr, a.cind, a.val, 1);
  const size_t es = amg->esize();
  V.alloc(es * static_cast<size_t>(n) * (m + 1));
  Z.alloc(es * static_cast<size_t>(n) * (m + 1));
  r.alloc(sizeof(double) * std::max(n, 1));
  // state layout (bytes, 8-aligned)
  const size_t H = 8 * static_cast<size_t>(m + 1) * m;
  const size_t sz = H + 8 * m * 2 + 8 * (m + 1) + 8 * m + 4 * m + 8 * dev::kRedBlocks * 3 +
                    8 * dev::kRedBlocks * 2 + 4 * dev::kRedBlocks * 2 + 64 + 8 * 8 + 8 * 4 +
                    4 * 8 + 16 + 128;
  state.alloc(sz);
  hip_check(hipMemsetAsync(state.get(), 0, sz, lib_stream()), "memset state");
  char* p = state.as<char>();
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 7) & ~static_cast<size_t>(7);
    return q;
  };
  k.m = m;
  k.bit = bit;
  k.eps = eps;
  k.H = take(H);
  k.c = take(8 * m);
  k.s = take(8 * m);
  k.e = take(8 * (m + 1));
  k.y = reinterpret_cast<double*>(take(8 * m));
  k.yf = reinterpret_cast<float*>(take(4 * m));
  k.ipart = reinterpret_cast<int64_t*>(take(8 * 3 * dev::kRedBlocks));
  k.dpart = reinterpret_cast<double*>(take(8 * 2 * dev::kRedBlocks));
  k.fpart = reinterpret_cast<float*>(take(4 * 2 * dev::kRedBlocks));
  k.counter = reinterpret_cast<unsigned*>(take(64));
  k.scal = reinterpret_cast<double*>(take(8 * 8));
  k.iscal = reinterpret_cast<int64_t*>(take(8 * 4));
  k.flags = reinterpret_cast<int*>(take(4 * 8));
  k.trace_len = reinterpret_cast<int*>(take(16));
  k.trace = nullptr;
  k.trace_cap = 0;
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_flags), 8 * sizeof(int) + 16,
                          hipHostMallocDefault),
            "hipHostMalloc");
  hip_check(hipStreamSynchronize(lib_stream()), "solver init");
}

DevSolver::~DevSolver() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (host_flags) (void)hipHostFree(host_flags);
  work.reset();
  if (stream) (void)hipStreamDestroy(stream);
}

void EventTimer::begin(bool aware) {
  if (!on) return;
  while (static_cast<int>(ev.size()) < 2 * used + 2) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "event");
    ev.push_back(e);
  }
  // A timing-aware launcher takes the armed pair and binds it to its own
  // dispatch; a marker recorded on the stream before it would cost a gap of
  // several microseconds per launch, so the start is recorded only for the
  // other launchers (the stream interval is timed then).
  if (!aware) hip_check(hipEventRecord(ev[2 * used], lib_stream()), "record");
  dev::arm_dispatch_events(ev[2 * used], ev[2 * used + 1]);
  armed_aware = aware;
}

void EventTimer::end() {
  if (!on) return;
  if (dev::take_dispatch_events().start) {
    if (armed_aware) return;  // no launch took the pair: nothing timed
    // no launcher took the events: time the stream interval instead
    hip_check(hipEventRecord(ev[2 * used + 1], lib_stream()), "record");
    ++stream_timed;
  }
  ++used;
}

void EventTimer::collect() {
  for (int q = 0; q < used; ++q) {
    hip_check(hipEventSynchronize(ev[2 * q + 1]), "sync");
    float t = 0;
    hip_check(hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]), "elapsed");
    ms_total += t;
    ++launches;
  }
  used = 0;
}

void EventTimer::reset() {
  used = 0;
  launches = 0;
  stream_timed = 0;
  ms_total = 0;
}

EventTimer::~EventTimer() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

void DevSolver::read_flags() {
  hip_check(hipMemcpyAsync(host_flags, k.flags, 4 * sizeof(int), hipMemcpyDeviceToHost,
                           lib_stream()),
            "flags");
  hip_check(hipMemcpyAsync(host_flags + 8, k.scal + 5, sizeof(double),
                           hipMemcpyDeviceToHost, lib_stream()),
            "residual");
  hip_check(hipStreamSynchronize(lib_stream()), "flags sync");
}

void DevSolver::set_trace(int cap) {
  StreamScope scope(stream);
  if (cap <= 0) {
    tracing = false;
    return;
  }
  trace_cap = cap;
  pinned_dummy.alloc(sizeof(uint64_t) * cap);
  k.trace = pinned_dummy.as<uint64_t>();
  k.trace_cap = cap;
  tracing = true;
  hip_check(hipMemsetAsync(k.trace_len, 0, sizeof(int), lib_stream()), "trace reset");
}

void DevSolver::read_trace(uint64_t* out, int cap, int* len) {
  StreamScope scope(stream);
  int l = 0;
  hip_check(hipMemcpyAsync(&l, k.trace_len, sizeof(int), hipMemcpyDeviceToHost, lib_stream()),
            "trace len");
  hip_check(hipStreamSynchronize(lib_stream()), "sync");
  const int cnt = std::min(std::min(l, cap), trace_cap);
  if (cnt > 0) pinned_dummy.download(out, sizeof(uint64_t) * cnt);
  *len = cnt;
}

DevSolver::Result DevSolver::solve(const double* b_dev, double nrm_b, double* x_dev) {
  StreamScope scope(stream);
  Result res;
  hipStream_t s = lib_stream();
  const int n = amg->n0();
  const size_t es = amg->esize();
  k.nrm_b = nrm_b;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  history.clear();
  // Restarted: prev = 0 before the first pass
  hip_check(hipMemsetAsync(k.scal, 0, 8 * sizeof(double), s), "reset scal");
  hip_check(hipMemsetAsync(k.flags, 0, 8 * sizeof(int), s), "reset flags");
  hip_check(hipEventRecord(e0, s), "record");
  const int seq = opts.seq_reduction;
  char* Vb = V.as<char>();
  char* Zb = Z.as<char>();
  int iter = 0;
  res.status = 6;
  StallWatch stall;
  bool stalled = false;
  for (int outer = 0; outer < max_outer && !stalled; ++outer) {
    hip_check(dev::launch_outer_residual(a64.d, b_dev, x_dev, r.as<double>(), k, elem, seq, s),
              "outer residual");
    read_flags();
    if (host_flags[2]) {
      res.status = 4;
      break;
    }
    ++res.restarts;
    hip_check(dev::launch_v0(elem, n, r.as<double>(), Vb, k, s), "v0");
    int steps = 0;
    bool conv = false, ovf = false;
    for (int j = 0; j < m; ++j) {
      void* vj = Vb + es * static_cast<size_t>(j) * n;
      void* zj = Zb + es * static_cast<size_t>(j) * n;
      void* w = Vb + es * static_cast<size_t>(j + 1) * n;
      amg->vcycle(vj, zj, *work);
      if (tracing && elem == Elem::kI32) {
        hip_check(dev::launch_checksum(n, static_cast<const int32_t*>(vj), k, s), "cs");
        hip_check(dev::launch_checksum(n, static_cast<const int32_t*>(zj), k, s), "cs");
      }
      spmv_timer.begin(true);  // every SpMV launcher binds the events
      amg->spmv(amg->a0, SpmvMode::kPlain, zj, nullptr, w);
      spmv_timer.end();
      hip_check(dev::launch_dot_first(elem, n, Vb, w, k, j, seq, s), "dot");
      for (int kk = 0; kk <= j; ++kk)
        hip_check(dev::launch_axpy_dot(elem, n, Vb, w, k, kk, j, seq, s), "axpy_dot");
      hip_check(dev::launch_scal(elem, n, w, k, j, seq, s), "scal");
      read_flags();
      const double rel = *reinterpret_cast<double*>(host_flags + 8);
      history.push_back(rel);
      steps = j + 1;
      if (host_flags[1]) {
        ovf = true;
        break;
      }
      if (host_flags[0]) {
        conv = true;
        break;
      }
      if (stall.step(rel, opts.stall_steps)) {
        stalled = true;
        break;
      }
    }
    if (ovf) {
      res.status = 3;
      res.overflow = 1;
      break;
    }
    if (stalled) {  // abandoned: no update of x, STAGNATION
      res.status = 4;
      break;
    }
    hip_check(dev::launch_update_x(elem, n, Zb, x_dev, k, steps, s), "update x");
    iter += conv ? steps : m;
    if (conv) {
      res.status = 0;
      res.converged = 1;
      res.iter = iter;
      break;
    }
  }
  hip_check(hipEventRecord(e1, s), "record");
  hip_check(hipEventSynchronize(e1), "sync");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  res.solve_ms = ms;
  double sc[8];
  hip_check(hipMemcpy(sc, k.scal, sizeof sc, hipMemcpyDeviceToHost), "scal");
  res.implicit_relres = sc[5];
  spmv_timer.collect();
  trsv0_timer.collect();
  amg->check_errors(*work);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return res;
}

double DevSolver::relres(const double* b_dev, double nrm_b, const double* x_dev) {
  StreamScope scope(stream);
  hip_check(dev::launch_relres(a64.d, b_dev, x_dev, r.as<double>(), k.dpart,
                               k.counter + 4, k.scal + 6, lib_stream()),
            "relres");
  double v = 0;
  hip_check(hipMemcpyAsync(&v, k.scal + 6, sizeof(double), hipMemcpyDeviceToHost,
                           lib_stream()),
            "relres copy");
  hip_check(hipStreamSynchronize(lib_stream()), "relres sync");
  return v / nrm_b;
}

}  // namespace sc23ad
