// Synthetic linear systems for the benchmark configurations.
//
// The reference reads SuiteSparse matrices (Library/io/mm.hpp:24-98) that are
// not available offline, so every workload here is generated.  Random numbers
// come from std::mt19937_64 (fully specified by the C++ standard) through an
// explicit map u = (x >> 11) * 2^-53 in [0,1), never through
// std::uniform_real_distribution, whose output is implementation-defined.
// Rows list their columns in ascending order.
#pragma once

#include <cstdint>
#include <random>
#include <vector>

#include "csr.hpp"

namespace sc23ad {

struct Rng {
  std::mt19937_64 eng;
  explicit Rng(uint64_t seed) : eng(seed) {}
  double uniform() { return static_cast<double>(eng() >> 11) * 0x1.0p-53; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  int uniform_int(int lo, int hi) {  // inclusive
    return lo + static_cast<int>(uniform() * (hi - lo + 1));
  }
  double normal();  // Box-Muller from two uniforms
};

// 2D 5-point Poisson, n x n grid, Dirichlet: 4 on the diagonal, -1 off it.
HostCsr gen_poisson2d(int n);
// 3D 7-point Laplacian, n^3 grid, Dirichlet: 6 / -1.
HostCsr gen_laplace3d(int n);
// 3D 27-point variable-coefficient diffusion on n^3 nodes.  k = exp(2 g) with
// g ~ N(0,1) constant on block^3-node blocks, anisotropy (ax, ay, az); the
// coupling of nodes i and j = i + d is -w(d) * hm(k_i, k_j), where hm is the
// harmonic mean and w(d) = (sum over the nonzero axes of d of a_axis) /
// 4^(m-1), m = number of nonzero components of d.  The diagonal is the sum of
// all 26 couplings, with k_i standing in for neighbours outside the domain
// (Dirichlet), so A is a symmetric M-matrix.
HostCsr gen_diffusion27(int n, int block, double ax, double ay, double az,
                        uint64_t seed);
// 2D 5-point first-order upwind convection-diffusion -lap u + b . grad u on an
// n x n grid, h = 1/(n+1), b = pe * (cos angle, sin angle), scaled by h^2.
HostCsr gen_convdiff2d(int n, double pe, double angle);
// 3D 7-point diffusion on n^3 nodes: k = 1 with `nsph` spheres of k = kin,
// radius uniform in [rmin, rmax] cells, centres uniform in the box; coupling
// -hm(k_i, k_j), diagonal = sum of couplings with k_i for outside neighbours.
HostCsr gen_inclusions3d(int n, int nsph, double rmin, double rmax, double kin,
                         uint64_t seed);

// x* ~ U[-1,1] from mt19937_64(seed); b = A x* with each row summed in column
// order.
void gen_xstar(int n, uint64_t seed, double* xstar);
void csr_matvec(const HostCsr& a, const double* x, double* y);

// Parameters of one system of the batched convection-diffusion family
// (256 systems, seed 3): side uniform in {32..256}, Peclet uniform in
// [1,100], angle uniform in [0, 2 pi), drawn in that order per system.
struct ConvDiffDraw {
  int n;
  double pe, angle;
};
std::vector<ConvDiffDraw> draw_convdiff_family(int count, uint64_t seed);

}  // namespace sc23ad
