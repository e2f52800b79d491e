// Device helpers of the triangular solves shared by the single and the block
// (several right-hand sides) kernels: the integer row finish of each factor
// kind, the LDS-only barrier of the chain kernels and the DPP group sums.
#pragma once

#include "dev_common.hpp"
#include "launch.hpp"

namespace sc23ad::dev {
namespace {

constexpr int kTB = 64;
constexpr int kBlkThreads = 1024;

template <typename T>
struct AccOf {
  using type = T;
};
template <>
struct AccOf<int32_t> {
  using type = uint64_t;
};

template <typename T>
__device__ __forceinline__ typename AccOf<T>::type acc_init(T v, int bit) {
  return v;
}
template <>
__device__ __forceinline__ uint64_t acc_init<int32_t>(int32_t v, int bit) {
  return static_cast<uint64_t>(static_cast<int64_t>(v)) << bit;
}

template <typename T>
__device__ __forceinline__ typename AccOf<T>::type acc_sub(typename AccOf<T>::type t, T v,
                                                           T x) {
  return t - v * x;
}
template <>
__device__ __forceinline__ uint64_t acc_sub<int32_t>(uint64_t t, int32_t v, int32_t x) {
  return t - prod64(v, x);
}

template <typename T, TriKind K>
__device__ __forceinline__ T acc_final(typename AccOf<T>::type t, T dinv, int bit, int dbit) {
  if (K == TriKind::kL) return t;
  return t * dinv;
}
template <>
__device__ __forceinline__ int32_t acc_final<int32_t, TriKind::kL>(uint64_t t, int32_t, int bit,
                                                                   int) {
  return low32(as_i64(t) >> bit);
}
template <>
__device__ __forceinline__ int32_t acc_final<int32_t, TriKind::kLD>(uint64_t t, int32_t dinv,
                                                                    int bit, int dbit) {
  return low32(mul_wrap(as_i64(t) >> bit, dinv) >> dbit);
}
template <>
__device__ __forceinline__ int32_t acc_final<int32_t, TriKind::kDU>(uint64_t t, int32_t dinv,
                                                                    int bit, int dbit) {
  return low32(mul_wrap(as_i64(t) >> bit, dinv) >> dbit);
}

constexpr int kTileThreads = 256;

// Wait budget of the tiled coarse solves: one launch gives up its waits once
// 2^27 ticks of the constant 100 MHz clock (1.3 s) have passed since the
// workgroup started, or as soon as any workgroup has recorded an error.
// Workgroups take their block from a ticket counter in processing order, so
// every wait is on a workgroup that is already running and ends within
// microseconds; the budget only bounds a launch when something else is wrong.
constexpr long long kTileMaxTicks = 1ll << 27;

__device__ __forceinline__ bool tile_wait_over(long long t0, int* err) {
  return wall_clock64() - t0 > kTileMaxTicks ||
         __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// The processing position of this workgroup (0, 1, ... in the order
// workgroups start), shared through LDS; tickets[0] counts starts.
__device__ __forceinline__ int tile_ticket(unsigned* tickets, int* sh) {
  if (threadIdx.x == 0) *sh = static_cast<int>(atomicAdd(&tickets[0], 1u));
  __syncthreads();
  return *sh;
}

// The last of `total` workgroups to finish resets both counters for the
// next launch on the stream (tickets[1] counts finishes).
__device__ __forceinline__ void tile_done(unsigned* tickets, int total) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned done = atomicAdd(&tickets[1], 1u);
    if (done == static_cast<unsigned>(total) - 1) {
      atomicExch(&tickets[0], 0u);
      atomicExch(&tickets[1], 0u);
    }
  }
}

// Workgroup barrier that orders LDS only: the solution is exchanged through
// the LDS ring, so outstanding global loads (the prefetch of the next levels)
// and stores (the result) may stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int kWaves>
__device__ __forceinline__ void ep_sync() {
  if (kWaves == 1) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
  } else {
    lds_barrier();
  }
}

// Sum of a 64-bit value over aligned groups of kG <= 16 lanes, with DPP
// moves inside rows of 16 lanes: swap neighbours (quad_perm 1,0,3,2), swap
// pairs (quad_perm 2,3,0,1), then lane i with lane 7 - i of its 8 (row half
// mirror) and lane 15 - i of its 16 (row mirror).  Every lane ends with its
// group's sum.  All lanes must be active.
template <int kCtrl>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint32_t>(v)), kCtrl,
                                             0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(
      0, static_cast<int>(static_cast<uint32_t>(v >> 32)), kCtrl, 0xF, 0xF, false);
  return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) | static_cast<uint32_t>(lo);
}

template <int kG>
__device__ __forceinline__ uint64_t group_sum(uint64_t v) {
  static_assert(kG == 1 || kG == 2 || kG == 4 || kG == 8 || kG == 16, "groups of 1-16 lanes");
  if (kG >= 2) v += dpp_u64<0xB1>(v);   // quad_perm(1, 0, 3, 2)
  if (kG >= 4) v += dpp_u64<0x4E>(v);   // quad_perm(2, 3, 0, 1)
  if (kG >= 8) v += dpp_u64<0x141>(v);  // row_half_mirror
  if (kG >= 16) v += dpp_u64<0x140>(v); // row_mirror
  return v;
}

}  // namespace
}  // namespace sc23ad::dev
