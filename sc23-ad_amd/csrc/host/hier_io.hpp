// Binary hierarchy file ("SC23AMG1"), used to hand a built hierarchy to the
// CPU oracle and to compare it with the one the reference builds.
//
// Layout (little endian):
//   char[8] "SC23AMG1"
//   int32  precision (0 FP64, 1 FP32, 2 INT), depth, smoother, iter, b_num,
//          max_levels, coarse_n, n_coarse
//   double omega, theta
//   for l in 0..depth-1: matrices A, LD, U, R, P
//   matrices Lc, Uc (coarsest-grid factors)
// each matrix: int32 n, m, nnz, bit, dbit; int32 rptr[n+1]; int32 cind[nnz];
//              values[nnz] as int32 / float / double by precision.
#pragma once

#include <string>

#include "hierarchy.hpp"

namespace sc23ad {

bool write_hierarchy(const std::string& path, const IntHierarchy& h);
bool write_hierarchy(const std::string& path, const FpHierarchy<double>& h);
bool write_hierarchy(const std::string& path, const FpHierarchy<float>& h);

}  // namespace sc23ad
