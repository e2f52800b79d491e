#include "rs_setup.hpp"

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <utility>
#include <algorithm>

namespace sc23ad {

namespace {
constexpr int kUnassigned = -38;  // first pass's "not yet visited" marker
}

// get_strong_weak_set (helper_multigrid.hpp:67-108).  `weak` keeps the
// diagonal together with the weak off-diagonals, in row order.
void rs_strength(const HostCsr& a, double theta, HostCsr& strong,
                 HostCsr& weak) {
  const int n = a.n;
  strong.reset(n, n);
  weak.reset(n, n);
  strong.cind.reserve(a.nnz());
  strong.val.reserve(a.nnz());
  weak.cind.reserve(a.nnz());
  weak.val.reserve(a.nnz());
  for (int i = 0; i < n; ++i) {
    double mx = 0;
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      if (a.cind[p] == i) continue;
      if (std::fabs(a.val[p]) > mx) mx = std::fabs(a.val[p]);
    }
    for (int p = a.rptr[i]; p < a.rptr[i + 1]; ++p) {
      const int c = a.cind[p];
      if (c != i && std::fabs(a.val[p]) >= theta * mx) {
        strong.cind.push_back(c);
        strong.val.push_back(a.val[p]);
      } else {
        weak.cind.push_back(c);
        weak.val.push_back(a.val[p]);
      }
    }
    strong.rptr[i + 1] = static_cast<int>(strong.cind.size());
    weak.rptr[i + 1] = static_cast<int>(weak.cind.size());
  }
}

// get_influence_set (helper_multigrid.hpp:110-136): transpose of the strength
// pattern, each row listing its sources in ascending order.
HostCsr pattern_transpose(const HostCsr& s) {
  HostCsr t;
  t.reset(s.m, s.n);
  for (int p = 0; p < s.nnz(); ++p) ++t.rptr[s.cind[p] + 1];
  for (int i = 0; i < s.m; ++i) t.rptr[i + 1] += t.rptr[i];
  t.cind.resize(s.nnz());
  std::vector<int> fill(t.rptr.begin(), t.rptr.end() - 1);
  for (int i = 0; i < s.n; ++i)
    for (int p = s.rptr[i]; p < s.rptr[i + 1]; ++p)
      t.cind[fill[s.cind[p]]++] = i;
  return t;
}

// transpose_interpolation_matrix (helper_multigrid.hpp:534-562).
HostCsr csr_transpose(const HostCsr& p) {
  HostCsr t;
  t.reset(p.m, p.n);
  for (int q = 0; q < p.nnz(); ++q) ++t.rptr[p.cind[q] + 1];
  for (int i = 0; i < p.m; ++i) t.rptr[i + 1] += t.rptr[i];
  t.cind.resize(p.nnz());
  t.val.resize(p.nnz());
  std::vector<int> fill(t.rptr.begin(), t.rptr.end() - 1);
  for (int i = 0; i < p.n; ++i)
    for (int q = p.rptr[i]; q < p.rptr[i + 1]; ++q) {
      const int dst = fill[p.cind[q]]++;
      t.cind[dst] = i;
      t.val[dst] = p.val[q];
    }
  return t;
}

// first_pass (helper_multigrid.hpp:138-266).  Points are ordered by how many
// points they strongly influence (descending), ties by "influence"
// (ascending) inside each cardinality bucket; then the greedy selection keeps
// the order up to date by swapping points between buckets.  The bucket
// bookkeeping is reproduced step for step, because it decides which point is
// picked next.
std::vector<int> rs_first_pass(const HostCsr& s, const HostCsr& st,
                               int* warnings) {
  const int n = s.n;
  std::vector<int> cf(n), visited(n, 0), card(n), outdeg(n), infl(n, 0),
      order(n), where(n);
  for (int i = 0; i < n; ++i) {
    card[i] = st.rptr[i + 1] - st.rptr[i];
    outdeg[i] = s.rptr[i + 1] - s.rptr[i];
    order[i] = i;
    if (card[i] == 0 && outdeg[i] == 0) {
      cf[i] = 0;
      visited[i] = 1;
    } else {
      cf[i] = kUnassigned;
    }
    for (int p = st.rptr[i]; p < st.rptr[i + 1]; ++p) {
      const int src = st.cind[p];
      infl[i] += s.rptr[src + 1] - s.rptr[src];
    }
  }
  if (n > 0)
    hoare_sort(0, n - 1, false, card.data(),
               {outdeg.data(), infl.data(), order.data()});

  int max_card = n > 0 ? card[0] : 0;
  // bucket_end[c] is one past the last sorted position holding cardinality c;
  // bucket c spans [bucket_end[c+1], bucket_end[c]).
  std::vector<int> bucket_end(max_card + 2, 0);
  {
    int start = 0;
    for (int c = max_card; c >= 0; --c) {
      int count = 0;
      bool stopped = false;
      for (int q = start; q < n; ++q) {
        if (card[q] == c) {
          ++count;
        } else {
          bucket_end[c] = bucket_end[c + 1] + count;
          start = q;
          stopped = true;
          break;
        }
      }
      if (!stopped) bucket_end[c] = bucket_end[c + 1] + count;
    }
  }
  for (int c = max_card; c >= 0; --c)
    if (bucket_end[c] - bucket_end[c + 1] > 1)
      hoare_sort(bucket_end[c + 1], bucket_end[c] - 1, true, infl.data(),
                 {order.data()});

  for (int q = 0; q < n; ++q) where[order[q]] = q;

  auto swap_slots = [&](int a, int b) {
    std::swap(card[a], card[b]);
    std::swap(order[a], order[b]);
    std::swap(where[order[a]], where[order[b]]);
  };

  int pos = 0;
  while (pos < n) {
    while (pos < n && visited[order[pos]]) ++pos;
    if (pos >= n) break;
    const int c_pt = order[pos];
    visited[c_pt] = 1;
    cf[c_pt] = 1;
    for (int p = st.rptr[c_pt]; p < st.rptr[c_pt + 1]; ++p) {
      const int f_pt = st.cind[p];
      if (visited[f_pt]) continue;
      visited[f_pt] = 1;
      cf[f_pt] = 0;
      for (int q = s.rptr[f_pt]; q < s.rptr[f_pt + 1]; ++q) {
        if (visited[s.cind[q]]) continue;
        const int slot = where[s.cind[q]];
        if (slot <= pos && warnings) ++*warnings;
        ++card[slot];
        int target;
        if (card[slot] > max_card) {
          max_card = card[slot];
          if (bucket_end[card[slot]] > pos + 1 && warnings) ++*warnings;
          target = pos + 1;
          bucket_end.push_back(pos + 1);
          bucket_end[max_card] = pos + 2;
        } else if (bucket_end[card[slot]] > pos + 1) {
          target = bucket_end[card[slot]];
          bucket_end[card[slot]] += 1;
        } else {
          target = pos + 1;
          bucket_end[card[slot]] = pos + 2;
        }
        swap_slots(slot, target);
      }
    }
    for (int p = s.rptr[c_pt]; p < s.rptr[c_pt + 1]; ++p) {
      if (visited[s.cind[p]]) continue;
      const int slot = where[s.cind[p]];
      if (card[slot] == 0) continue;
      --card[slot];
      const int target = bucket_end[card[slot] + 1] - 1;
      bucket_end[card[slot] + 1] -= 1;
      swap_slots(slot, target);
    }
    ++pos;
  }
  return cf;
}

// second_pass (helper_multigrid.hpp:378-417): an F point with a strong F
// neighbour that shares no strong C point becomes a C point, in place, in
// ascending order.
int rs_second_pass(const HostCsr& s, std::vector<int>& cf) {
  const int n = s.n;
  for (int i = 0; i < n; ++i) {
    if (cf[i] == kUnassigned)
      throw SetupError{kSetupFail, "second pass found an unassigned point"};
    if (cf[i]) continue;
    for (int p = s.rptr[i]; p < s.rptr[i + 1]; ++p) {
      const int f = s.cind[p];
      if (cf[f]) continue;
      bool shared = false;
      int a = s.rptr[i], b = s.rptr[f];
      while (a < s.rptr[i + 1] && b < s.rptr[f + 1]) {
        if (s.cind[b] < s.cind[a]) {
          ++b;
        } else if (s.cind[b] == s.cind[a]) {
          if (cf[s.cind[a]]) {
            shared = true;
            break;
          }
          ++a;
          ++b;
        } else {
          ++a;
        }
      }
      if (shared) continue;
      cf[i] = 1;
      break;
    }
  }
  int nc = 0;
  for (int i = 0; i < n; ++i)
    if (cf[i]) ++nc;
  return nc;
}

// sepalate_strong_to_coarse_fine (helper_multigrid.hpp:419-455).
void split_strong(const HostCsr& s, const std::vector<int>& cf, HostCsr& cs,
                  HostCsr& ds) {
  cs.reset(s.n, s.n);
  ds.reset(s.n, s.n);
  for (int i = 0; i < s.n; ++i) {
    for (int p = s.rptr[i]; p < s.rptr[i + 1]; ++p) {
      HostCsr& dst = cf[s.cind[p]] ? cs : ds;
      dst.cind.push_back(s.cind[p]);
      dst.val.push_back(s.val[p]);
    }
    cs.rptr[i + 1] = static_cast<int>(cs.cind.size());
    ds.rptr[i + 1] = static_cast<int>(ds.cind.size());
  }
}

namespace {

// get_interpolation_matrix (helper_multigrid.hpp:457-532).  For an F point i
// and each strong C neighbour j:
//   w_ij = -(a_ij + sum_{m in D^s_i} a_im * a_mj / sum_{k in C^s_i} a_mk)
//          / sum_{n in weak_i (incl. diagonal)} a_in
// evaluated left to right in row order.  C rows are unit rows.
HostCsr interpolation(const HostCsr& a, const HostCsr& cs, const HostCsr& ds,
                      const HostCsr& weak, const std::vector<int>& cf,
                      int* warnings) {
  const int n = a.n;
  std::vector<int> col(n, 0);
  int nc = 0;
  for (int i = 0; i < n; ++i)
    if (cf[i]) col[i] = nc++;
  HostCsr p;
  p.reset(n, nc);
  for (int i = 0; i < n; ++i) {
    if (cf[i]) {
      p.cind.push_back(col[i]);
      p.val.push_back(1.0);
    } else {
      double deno = 0;
      for (int q = weak.rptr[i]; q < weak.rptr[i + 1]; ++q) deno += weak.val[q];
      for (int q = cs.rptr[i]; q < cs.rptr[i + 1]; ++q) {
        double aij = cs.val[q];
        const int cj = cs.cind[q];
        for (int r = ds.rptr[i]; r < ds.rptr[i + 1]; ++r) {
          const int mrow = ds.cind[r];
          double amj = 0;
          for (int t = a.rptr[mrow]; t < a.rptr[mrow + 1]; ++t)
            if (a.cind[t] == cj) {
              amj = a.val[t];
              break;
            }
          bool hit = false;
          double amk = 0;
          int mp = a.rptr[mrow], kp = cs.rptr[i];
          while (mp < a.rptr[mrow + 1] && kp < cs.rptr[i + 1]) {
            if (a.cind[mp] < cs.cind[kp]) {
              ++mp;
            } else if (a.cind[mp] == cs.cind[kp]) {
              amk += a.val[mp];
              hit = true;
              ++mp;
              ++kp;
            } else {
              ++kp;
            }
          }
          if (amk == 0) {
            if (!hit)
              throw SetupError{kSetupFail,
                               "interpolation: F-F connection shares no C "
                               "point (reference exits here)"};
            if (warnings) ++*warnings;
            amk = 0.00001;
          }
          aij += ds.val[r] * amj / amk;
        }
        p.cind.push_back(col[cj]);
        p.val.push_back(-aij / deno);
      }
    }
    p.rptr[i + 1] = static_cast<int>(p.cind.size());
  }
  return p;
}

}  // namespace

// Each output row is built with a dense accumulator: for every column c the
// terms alpha_k * R(k,c) arrive in the order of k exactly as in the sorted
// merge (a column that is absent takes alpha*x, a present one y + alpha*x),
// and a value that becomes exactly zero is removed, so a later term restarts
// it from alpha*x just as the merge would.  Rows are independent and are
// computed in parallel; the result does not depend on the thread count.
HostCsr rowmerge_product(const HostCsr& l, const HostCsr& r) {
  HostCsr out;
  out.reset(l.n, r.m);
  const int n = l.n;
  std::vector<std::vector<int>> rc(n);
  std::vector<std::vector<double>> rv(n);
#pragma omp parallel
  {
    std::vector<double> acc(static_cast<size_t>(r.m), 0.0);
    std::vector<char> present(static_cast<size_t>(r.m), 0);
    std::vector<int> touched;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      touched.clear();
      for (int q = l.rptr[i]; q < l.rptr[i + 1]; ++q) {
        const double alpha = l.val[q];
        const int k = l.cind[q];
        for (int p = r.rptr[k]; p < r.rptr[k + 1]; ++p) {
          const int c = r.cind[p];
          double v;
          if (present[c]) {
            v = acc[c] + alpha * r.val[p];
          } else {
            v = alpha * r.val[p];
            touched.push_back(c);
          }
          if (std::fabs(v) == 0.0) {
            present[c] = 0;
          } else {
            present[c] = 1;
            acc[c] = v;
          }
        }
      }
      std::sort(touched.begin(), touched.end());
      touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
      std::vector<int>& oc = rc[i];
      std::vector<double>& ov = rv[i];
      for (int c : touched) {
        if (present[c]) {
          oc.push_back(c);
          ov.push_back(acc[c]);
        }
        present[c] = 0;
      }
    }
  }
  for (int i = 0; i < n; ++i) out.rptr[i + 1] = out.rptr[i] + static_cast<int>(rc[i].size());
  out.cind.resize(out.rptr[n]);
  out.val.resize(out.rptr[n]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    std::copy(rc[i].begin(), rc[i].end(), out.cind.begin() + out.rptr[i]);
    std::copy(rv[i].begin(), rv[i].end(), out.val.begin() + out.rptr[i]);
  }
  return out;
}

namespace {
struct PhaseClock {
  bool on = std::getenv("SC23AD_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void lap(const char* what, int n) {
    if (!on) return;
    auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[sc23ad setup] n=%d %-14s %8.3f s\n", n, what,
                 std::chrono::duration<double>(now - t).count());
    t = now;
  }
};
}  // namespace

CoarsenStep rs_coarsen(const HostCsr& a, double theta) {
  if (theta <= 0 || 1 < theta)
    throw SetupError{kBadParam, "theta must satisfy 0 < theta <= 1"};
  PhaseClock clk;
  CoarsenStep step;
  HostCsr strong, weak;
  rs_strength(a, theta, strong, weak);
  HostCsr strong_t = pattern_transpose(strong);
  clk.lap("strength", a.n);
  step.cf = rs_first_pass(strong, strong_t, &step.warnings);
  clk.lap("first pass", a.n);
  rs_second_pass(strong, step.cf);
  clk.lap("second pass", a.n);
  HostCsr cs, ds;
  split_strong(strong, step.cf, cs, ds);
  step.interp = interpolation(a, cs, ds, weak, step.cf, &step.warnings);
  clk.lap("interpolation", a.n);
  step.restrict_ = csr_transpose(step.interp);
  // get_coarse_matrix (helper_multigrid.hpp:564-586): rows of P^T A first,
  // then (P^T A) P.
  HostCsr rt_a = rowmerge_product(step.restrict_, a);
  clk.lap("R*A", a.n);
  step.coarse = rowmerge_product(rt_a, step.interp);
  clk.lap("(RA)*P", a.n);
  return step;
}

}  // namespace sc23ad
