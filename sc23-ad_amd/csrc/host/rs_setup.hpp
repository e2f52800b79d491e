// Classical Ruge-Stueben coarsening: one coarsening step of the reference's
// helper::multigrid::get_amg_set (Library/helper/helper_multigrid.hpp:34-63).
//
// Reproduced exactly: strength by |a_ij| >= theta * max_{k!=i} |a_ik|, the
// first pass's sort + bucket bookkeeping, the in-place second pass, the
// classical interpolation with its FP64 operation order, R = P^T, and the
// row-merge Galerkin product P^T A P that drops exact zeros.
#pragma once

#include "csr.hpp"

namespace sc23ad {

struct CoarsenStep {
  HostCsr coarse;     // C_A = P^T A P (unscaled)
  HostCsr interp;     // P  (n_fine x n_coarse)
  HostCsr restrict_;  // R = P^T (n_coarse x n_fine)
  std::vector<int> cf;  // 1 = coarse point, 0 = fine point
  int warnings = 0;     // count of the reference's "Warning 1/2" and "1 amk = 0"
};

// Throws SetupError{kBadParam} for theta outside (0,1] and
// SetupError{kSetupFail} where the reference calls exit().
CoarsenStep rs_coarsen(const HostCsr& a, double theta);
// The same step with the interpolation weights and both Galerkin products
// computed on the GPU (setup_dev.cpp); identical results.
CoarsenStep rs_coarsen_dev(const HostCsr& a, double theta);

// Building blocks, exposed for the unit tests.
void rs_strength(const HostCsr& a, double theta, HostCsr& strong,
                 HostCsr& weak);
std::vector<int> rs_first_pass(const HostCsr& strong, const HostCsr& strong_t,
                               int* warnings);
int rs_second_pass(const HostCsr& strong, std::vector<int>& cf);
HostCsr pattern_transpose(const HostCsr& s);
// Strong connections to C points (cs) and to F points (ds), in row order.
void split_strong(const HostCsr& s, const std::vector<int>& cf, HostCsr& cs,
                  HostCsr& ds);
HostCsr csr_transpose(const HostCsr& p);
// out = L * R computed row by row as sum_k L(i,k) * R(k,:), merging in
// ascending column order, y + alpha*x per shared column, dropping exact zeros
// (Library/helper/helper_sparse.hpp:213-285).
HostCsr rowmerge_product(const HostCsr& l, const HostCsr& r);

}  // namespace sc23ad
