#include <algorithm>
#include "hierarchy.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace sc23ad {

namespace {
struct Lap {
  bool on = std::getenv("SC23AD_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void operator()(const char* what) {
    if (!on) return;
    auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[sc23ad setup] %-24s %8.3f s\n", what,
                 std::chrono::duration<double>(now - t).count());
    t = now;
  }
};
}  // namespace

AmgChain build_chain(const HostCsr& fine, const AmgParams& prm) {
  if (prm.max_levels < 1 || prm.iter < 1 || !(prm.omega > 0) ||
      (prm.smoother == SmootherKind::kBlockSOR && prm.b_num < 1))
    throw SetupError{kBadParam, "invalid AMG parameters"};
  Lap lap;
  AmgChain ch;
  // INTAMG.hpp:162-195 / 38-51: coarsen until max_levels is reached or a
  // level has at most coarse_n rows.  Each step consumes the unscaled
  // previous coarse operator.
  auto coarsen = [&](const HostCsr& a) {
    return prm.device_setup ? rs_coarsen_dev(a, prm.theta) : rs_coarsen(a, prm.theta);
  };
  CoarsenStep st = coarsen(fine);
  ch.warnings += st.warnings;
  ch.coarse.push_back(std::move(st.coarse));
  ch.interp.push_back(std::move(st.interp));
  ch.restrict_.push_back(std::move(st.restrict_));
  ch.depth = 1;
  while (prm.max_levels > ch.depth && ch.coarse[ch.depth - 1].n > prm.coarse_n) {
    ++ch.depth;
    CoarsenStep nx = coarsen(ch.coarse[ch.depth - 2]);
    ch.warnings += nx.warnings;
    ch.coarse.push_back(std::move(nx.coarse));
    ch.interp.push_back(std::move(nx.interp));
    ch.restrict_.push_back(std::move(nx.restrict_));
  }
  lap("Galerkin chain");
  return ch;
}

// SOR: split "LD-U" (helper_matrix.hpp:530-604), then move (1 - 1/omega) * d
// to the front of each U row and scale the last LD entry by 1/omega
// (stationary.hpp:385-396).  BlockSOR: b_num contiguous blocks of
// ceil(n/b_num) rows; LD keeps the diagonal / omega and the lower entries of
// the row's own block, U everything else (stationary.hpp:934-978).
void smoother_split(const HostCsr& a, const AmgParams& prm, HostCsr& ld,
                    HostCsr& u, int level) {
  const int n = a.n;
  const double omega = static_cast<size_t>(level) < prm.level_omega.size()
                           ? prm.level_omega[level]
                           : prm.omega;
  ld.reset(n, a.m);
  u.reset(n, a.m);
  if (prm.smoother == SmootherKind::kSOR) {
    for (int i = 0; i < n; ++i) {
      const size_t ld_begin = ld.cind.size();
      for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p)
        if (a.cind[p] <= i) {
          ld.cind.push_back(a.cind[p]);
          ld.val.push_back(a.val[p]);
        }
      if (ld.cind.size() == ld_begin)
        throw SetupError{kSetupFail,
                         "SOR split: row " + std::to_string(i) +
                             " has no diagonal or lower entry"};
      double& d = ld.val.back();
      u.cind.push_back(ld.cind.back());
      u.val.push_back((1 - 1.0 / omega) * d);
      d = 1.0 / omega * d;
      for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p)
        if (a.cind[p] > i) {
          u.cind.push_back(a.cind[p]);
          u.val.push_back(a.val[p]);
        }
      ld.rptr[i + 1] = static_cast<int>(ld.cind.size());
      u.rptr[i + 1] = static_cast<int>(u.cind.size());
    }
    return;
  }
  if (prm.smoother == SmootherKind::kJacobiSOR) {
    // stationary.hpp:586-631: left = start of the row's 1200-row block.
    for (int i = 0; i < n; ++i) {
      const int left = (i / kJacobiSorBlock) * kJacobiSorBlock;
      for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
        const int c = a.cind[p];
        const double v = a.val[p];
        if (c == i) {
          ld.cind.push_back(c);
          ld.val.push_back((1.0 / omega) * v);
          u.cind.push_back(c);
          u.val.push_back((1 - 1.0 / omega) * v);
        } else if (c < left) {
          ld.cind.push_back(c);
          ld.val.push_back(v);
        } else {
          u.cind.push_back(c);
          u.val.push_back(v);
        }
      }
      ld.rptr[i + 1] = static_cast<int>(ld.cind.size());
      u.rptr[i + 1] = static_cast<int>(u.cind.size());
    }
    return;
  }
  const int bn = prm.b_num;
  const int size = (n + bn - 1) / bn;
  std::vector<int> bj(bn + 1);
  bj[0] = 0;
  for (int k = 0; k < bn; ++k) bj[k + 1] = bj[k] + size;
  bj[bn] = n;
  for (int i = 0; i < n; ++i) {
    const int left = bj[i / size];
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      const int c = a.cind[p];
      const double v = a.val[p];
      if (c == i) {
        ld.cind.push_back(c);
        ld.val.push_back((1.0 / omega) * v);
        u.cind.push_back(c);
        u.val.push_back((1 - 1.0 / omega) * v);
      } else if (left <= c && c < i) {
        ld.cind.push_back(c);
        ld.val.push_back(v);
      } else {
        u.cind.push_back(c);
        u.val.push_back(v);
      }
    }
    ld.rptr[i + 1] = static_cast<int>(ld.cind.size());
    u.rptr[i + 1] = static_cast<int>(u.cind.size());
  }
}

HostCsr extract_ld_inv(const HostCsr& a) {
  HostCsr o;
  o.reset(a.n, a.m);
  for (int i = 0; i < a.n; ++i) {
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      if (a.cind[p] == i) {
        o.cind.push_back(i);
        o.val.push_back(1 / a.val[p]);
      } else if (a.cind[p] < i) {
        o.cind.push_back(a.cind[p]);
        o.val.push_back(a.val[p]);
      }
    }
    o.rptr[i + 1] = static_cast<int>(o.cind.size());
  }
  return o;
}

HostCsr extract_l(const HostCsr& a) {
  HostCsr o;
  o.reset(a.n, a.m);
  for (int i = 0; i < a.n; ++i) {
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p)
      if (a.cind[p] < i) {
        o.cind.push_back(a.cind[p]);
        o.val.push_back(a.val[p]);
      }
    o.rptr[i + 1] = static_cast<int>(o.cind.size());
  }
  return o;
}

HostCsr extract_dinv_u(const HostCsr& a) {
  HostCsr o;
  o.reset(a.n, a.m);
  for (int i = 0; i < a.n; ++i) {
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      if (a.cind[p] == i) {
        o.cind.push_back(i);
        o.val.push_back(1 / a.val[p]);
      } else if (a.cind[p] > i) {
        o.cind.push_back(a.cind[p]);
        o.val.push_back(a.val[p]);
      }
    }
    o.rptr[i + 1] = static_cast<int>(o.cind.size());
  }
  return o;
}

// Complete sparse LU without pivoting, row by row (IKJ), keeping all fill and
// explicit zeros (helper_matrix.hpp:978-1060).  Row i is eliminated against
// each earlier row k in ascending column order, including fill created on the
// way: the multiplier is a_ik / u_kk, stored at column k, and every column c
// of row k's upper part becomes a_ic - m * u_kc (or -m * u_kc where row i had
// no entry).  A row k without a stored diagonal reuses the previous
// multiplier and contributes nothing, as the reference does.  The row is
// held in a dense accumulator; each entry sees the same operations in the
// same order as in the reference's sorted merge.
HostCsr complete_lu(const HostCsr& a) {
  const int n = a.n;
  HostCsr f;
  f.reset(n, a.m);
  if (n == 0) return f;
  std::vector<int> diag_pos(n, -1);
  std::vector<double> acc(static_cast<size_t>(a.m), 0.0);
  std::vector<char> present(static_cast<size_t>(a.m), 0);
  double pivot = 0;
  for (int i = 0; i < n; ++i) {
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      acc[a.cind[p]] = a.val[p];
      present[a.cind[p]] = 1;
    }
    if (i > 0) {
      for (int k = 0; k < i; ++k) {
        if (!present[k]) continue;
        const int dp = diag_pos[k];
        if (dp >= 0) pivot = acc[k] / f.val[dp];
        acc[k] = pivot;
        if (dp < 0) continue;
        const int end = f.rptr[k + 1];
        for (int j = dp + 1; j < end; ++j) {
          const int c = f.cind[j];
          if (present[c]) {
            acc[c] = acc[c] - pivot * f.val[j];
          } else {
            acc[c] = -pivot * f.val[j];
            present[c] = 1;
          }
        }
      }
    }
    for (int c = 0; c < a.m; ++c) {
      if (!present[c]) continue;
      if (c == i) diag_pos[i] = static_cast<int>(f.cind.size());
      f.cind.push_back(c);
      f.val.push_back(acc[c]);
      present[c] = 0;
    }
    f.rptr[i + 1] = static_cast<int>(f.cind.size());
  }
  return f;
}

namespace {

IntCsr int_shell(const HostCsr& a) {
  IntCsr o;
  o.n = a.n;
  o.m = a.m;
  o.rptr = a.rptr;
  o.cind = a.cind;
  o.val.resize(a.val.size());
  return o;
}

void check_bits(int bit, const char* what) {
  if (bit < 0 || bit > 30)
    throw SetupError{kOverflow, std::string(what) +
                                    ": values too large for int32 fixed point"};
}

}  // namespace

IntCsr to_flex(const HostCsr& a, int cap, int* saturated) {
  IntCsr o = int_shell(a);
  double mx = 0;
  for (double v : a.val)
    if (std::fabs(v) > mx) mx = std::fabs(v);
  if (mx >= 2147483647.0) throw SetupError{kOverflow, "CSRflex: |a| too large"};
  o.bit = flex_bits(mx);
  if (cap > 0) {
    if (cap > o.bit && saturated) *saturated = 1;
    o.bit = std::min(o.bit, cap);
  }
  check_bits(o.bit, "CSRflex");
  for (size_t p = 0; p < a.val.size(); ++p) o.val[p] = to_fixed(a.val[p], o.bit);
  return o;
}

IntCsr to_fixed_csr(const HostCsr& a, int bit) {
  IntCsr o = int_shell(a);
  o.bit = bit;
  for (size_t p = 0; p < a.val.size(); ++p) o.val[p] = to_fixed(a.val[p], bit);
  return o;
}

// ldtri::CSRflex / BJCSRflex: separate bit counts for the stored inverse
// diagonal (dbit) and the off-diagonal entries (bit), both by the flex rule.
IntCsr to_ld_flex(const HostCsr& l, int cap, int* saturated) {
  IntCsr o = int_shell(l);
  double mx = 0, dmx = 0;
  for (int i = 0; i < l.n; ++i)
    for (int p = l.rptr[i]; p < l.rptr[i + 1]; ++p) {
      const double v = std::fabs(l.val[p]);
      if (l.cind[p] == i) {
        if (v > dmx) dmx = v;
      } else if (v > mx) {
        mx = v;
      }
    }
  if (mx >= 2147483647.0 || dmx >= 2147483647.0)
    throw SetupError{kOverflow, "ldtri::CSRflex: |a| too large"};
  o.dbit = flex_bits(dmx);
  o.bit = flex_bits(mx);
  if (cap > 0) {
    if (cap > o.bit && saturated) *saturated = 1;
    o.bit = std::min(o.bit, cap);
  }
  check_bits(o.bit, "ldtri::CSRflex");
  check_bits(o.dbit, "ldtri::CSRflex");
  for (int i = 0; i < l.n; ++i)
    for (int p = l.rptr[i]; p < l.rptr[i + 1]; ++p)
      o.val[p] = to_fixed(l.val[p], l.cind[p] == i ? o.dbit : o.bit);
  return o;
}

IntCsr to_l_flex(const HostCsr& l) {
  IntCsr o = int_shell(l);
  double mx = 0;
  for (double v : l.val)
    if (std::fabs(v) > mx) mx = std::fabs(v);
  o.bit = tri_bits(mx);
  for (size_t p = 0; p < l.val.size(); ++p) o.val[p] = to_fixed(l.val[p], o.bit);
  return o;
}

IntCsr to_dinvu_flex(const HostCsr& du) {
  IntCsr o = int_shell(du);
  double mx = 0, dmx = 0;
  for (int i = 0; i < du.n; ++i)
    for (int p = du.rptr[i]; p < du.rptr[i + 1]; ++p) {
      const double v = std::fabs(du.val[p]);
      if (du.cind[p] == i) {
        if (v > dmx) dmx = v;
      } else if (v > mx) {
        mx = v;
      }
    }
  o.dbit = tri_bits(dmx);
  o.bit = tri_bits(mx);
  for (int i = 0; i < du.n; ++i)
    for (int p = du.rptr[i]; p < du.rptr[i + 1]; ++p)
      o.val[p] = to_fixed(du.val[p], du.cind[p] == i ? o.dbit : o.bit);
  return o;
}


IntHierarchy build_int_hierarchy(const HostCsr& fine, const AmgChain& chain,
                                 const AmgParams& prm) {
  if (prm.smoother == SmootherKind::kJacobi)
    throw SetupError{kBadParam,
                     "Jacobi smoother: FP precisions only (the reference's integer Jacobi "
                     "needs Fixed<bit> matrices, which intAMGflex does not use)"};
  Lap lap;
  IntHierarchy h;
  h.prm = prm;
  h.depth = chain.depth;
  const int d = chain.depth;
  std::vector<HostCsr> ca = chain.coarse;     // scaled in place below
  std::vector<HostCsr> rt = chain.restrict_;  // scaled in place below
  // INTAMG.hpp:176-211: divide each coarse operator's rows (and the matching
  // rows of R) by the row's max |a|, then multiply R's columns by the
  // previous level's row maxima.  `prev` starts as zeros, so a hierarchy with
  // one coarse level gets R_0 = 0 exactly as the reference does.
  lap("copies");
  std::vector<double> prev(static_cast<size_t>(fine.n), 0.0);
  for (int l = 0; l < d; ++l) {
    HostCsr& c = ca[l];
    HostCsr& r = rt[l];
    const bool last = (l == d - 1);
    std::vector<double> mx = row_abs_max(c);
    for (int i = 0; i < c.n; ++i) {
      for (int p = c.rptr[i]; p < c.rptr[i + 1]; ++p) c.val[p] /= mx[i];
      for (int p = r.rptr[i]; p < r.rptr[i + 1]; ++p) r.val[p] /= mx[i];
      if (!last && l == 0) continue;
      for (int p = r.rptr[i]; p < r.rptr[i + 1]; ++p) r.val[p] *= prev[r.cind[p]];
    }
    if (!last)
      for (int i = 0; i < c.n; ++i) prev[i] = mx[i];
  }
  h.lv.resize(d);
  for (int l = 0; l < d; ++l) {
    IntLevel& L = h.lv[l];
    const HostCsr& src = (l == 0) ? fine : ca[l - 1];
    const int cap = static_cast<size_t>(l) < prm.level_bits.size() ? prm.level_bits[l] : 0;
    int sat = 0;
    L.a = (l == 0) ? to_fixed_csr(fine, 30) : to_flex(src, cap, &sat);
    lap("  A");
    HostCsr ld, u;
    smoother_split(src, prm, ld, u, l);
    lap("  split");
    L.ld = to_ld_flex(extract_ld_inv(ld), cap, &sat);
    lap("  LD");
    L.u = to_flex(u, cap, &sat);
    h.saturated += sat;
    L.r = to_flex(rt[l]);
    L.p = to_flex(chain.interp[l]);
    // per-level vector bits: R_l's result is shifted d bits less, P_l's d
    // bits more (kept within the shifts' range [0, 62]; a step outside it
    // counts as saturated and is cut to the range's edge)
    if (!prm.level_vbits.empty()) {
      auto vb = [&](int q) {
        return static_cast<size_t>(q) < prm.level_vbits.size() ? prm.level_vbits[q] : 0;
      };
      int dv = vb(l + 1) - vb(l);
      const int lo = std::max(L.p.bit - 62, -62), hi = L.r.bit;
      if (dv < lo || dv > hi) {
        dv = std::clamp(dv, lo, hi);
        h.saturated += 1;
      }
      L.r.bit -= dv;
      L.p.bit += dv;
    }
    lap("  U R P");
  }
  lap("int levels");
  HostCsr lu = complete_lu(ca[d - 1]);
  lap("coarse LU");
  h.lc = to_l_flex(extract_l(lu));
  h.uc = to_dinvu_flex(extract_dinv_u(lu));
  h.n_coarse = ca[d - 1].n;
  return h;
}

namespace {
template <typename T>
FpCsr<T> to_fp(const HostCsr& a) {
  FpCsr<T> o;
  o.n = a.n;
  o.m = a.m;
  o.rptr = a.rptr;
  o.cind = a.cind;
  o.val.resize(a.val.size());
  for (size_t p = 0; p < a.val.size(); ++p) o.val[p] = static_cast<T>(a.val[p]);
  return o;
}
}  // namespace

namespace {
// Jacobi's d_inv (stationary.hpp:40-47): omega converted to T, divided in
// double by the row's first diagonal entry, stored as T, as one diagonal
// entry per row.
template <typename T>
FpCsr<T> jacobi_dinv(const HostCsr& a, double omega) {
  FpCsr<T> o;
  o.n = a.n;
  o.m = a.m;
  o.rptr.resize(static_cast<size_t>(a.n) + 1);
  o.cind.resize(a.n);
  o.val.resize(a.n);
  const double w = static_cast<double>(static_cast<T>(omega));
  for (int i = 0; i < a.n; ++i) {
    int p = a.rptr[i];
    while (p < a.rptr[i + 1] && a.cind[p] != i) ++p;
    if (p == a.rptr[i + 1])
      throw SetupError{kSetupFail, "Jacobi: row " + std::to_string(i) + " has no diagonal"};
    o.rptr[i] = i;
    o.cind[i] = i;
    o.val[i] = static_cast<T>(w / a.val[p]);
  }
  o.rptr[a.n] = a.n;
  return o;
}
}  // namespace

template <typename T>
FpHierarchy<T> build_fp_hierarchy(const HostCsr& fine, const AmgChain& chain,
                                  const AmgParams& prm) {
  FpHierarchy<T> h;
  h.prm = prm;
  h.depth = chain.depth;
  const int d = chain.depth;
  h.lv.resize(d);
  for (int l = 0; l < d; ++l) {
    FpLevel<T>& L = h.lv[l];
    const HostCsr& src = (l == 0) ? fine : chain.coarse[l - 1];
    L.a = to_fp<T>(src);
    if (prm.smoother == SmootherKind::kJacobi) {
      const double omega =
          static_cast<size_t>(l) < prm.level_omega.size() ? prm.level_omega[l] : prm.omega;
      L.ld = jacobi_dinv<T>(src, omega);
      L.u = L.a;
    } else {
      HostCsr ld, u;
      smoother_split(src, prm, ld, u, l);
      L.ld = to_fp<T>(extract_ld_inv(ld));
      L.u = to_fp<T>(u);
    }
    L.r = to_fp<T>(chain.restrict_[l]);
    L.p = to_fp<T>(chain.interp[l]);
  }
  HostCsr lu = complete_lu(chain.coarse[d - 1]);
  h.lc = to_fp<T>(extract_l(lu));
  h.uc = to_fp<T>(extract_dinv_u(lu));
  h.n_coarse = chain.coarse[d - 1].n;
  return h;
}

template FpHierarchy<double> build_fp_hierarchy<double>(const HostCsr&,
                                                        const AmgChain&,
                                                        const AmgParams&);
template FpHierarchy<float> build_fp_hierarchy<float>(const HostCsr&,
                                                      const AmgChain&,
                                                      const AmgParams&);

}  // namespace sc23ad
