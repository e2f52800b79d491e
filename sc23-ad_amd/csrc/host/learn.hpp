// Learned per-level smoother parameters: the level features the network
// reads, and the network's fixed shape.  The network itself runs on the GPU
// (csrc/kernels/mlp.hip); this header is host-only.
#pragma once

#include <cmath>

#include <cstdint>

#include "csr.hpp"

namespace sc23ad {

// Features of one level's operator A (unscaled, as the coarsening produced it):
//   0 log2(rows) / 24           1 entries per row / 32
//   2 mean |a_ii| / sum_j |a_ij| 3 min  |a_ii| / sum_j |a_ij|
//   4 mean max_{j!=i} |a_ij| / |a_ii|
//   5 fraction of off-diagonal entries > 0
//   6 asymmetry sum |a_ij - a_ji| / sum |a_ij|
//   7 level / depth
constexpr int kLevelFeatures = 8;
void level_features(const HostCsr& a, int level, int depth, float* f);

// Network: 8 -> 32 (tanh) -> 32 (tanh) -> 2 (sigmoid); outputs map to
// omega = 0.6 + 1.3 y0 and bits = round(8 + 22 y1).  Parameters in one float
// array: W1[32][8], b1[32], W2[32][32], b2[32], W3[2][32], b3[2].
constexpr int kMlpIn = kLevelFeatures, kMlpHidden = 32, kMlpOut = 2;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpOut * kMlpHidden + kMlpOut;
// Glorot-uniform weights from mt19937_64(seed), zero biases.
void mlp_init(uint64_t seed, float* params);
inline double mlp_omega(float y0) { return 0.6 + 1.3 * static_cast<double>(y0); }
inline int mlp_bits(float y1) { return static_cast<int>(8.0 + 22.0 * static_cast<double>(y1) + 0.5); }
// Second output -> fractional bits of the next coarser level's vectors
// relative to the fine level's F: [-21, 1] (F_{l+1} in [8, 30] for F = 29).
constexpr int kVbitsLo = -21, kVbitsHi = 1;
inline int mlp_vbits(float y1) {
  const int v = static_cast<int>(std::floor(kVbitsLo + (kVbitsHi - kVbitsLo) *
                                                           static_cast<double>(y1) + 0.5));
  return v < kVbitsLo ? kVbitsLo : v > kVbitsHi ? kVbitsHi : v;
}

}  // namespace sc23ad
