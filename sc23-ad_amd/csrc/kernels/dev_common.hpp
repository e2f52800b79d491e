// Device helpers shared by the HIP translation units (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sc23ad::dev {

// Integer semantics of the reference under C++20: wrap-around arithmetic,
// arithmetic right shift, truncation to the low 32 bits.
__device__ __forceinline__ int64_t as_i64(uint64_t v) { return static_cast<int64_t>(v); }
__device__ __forceinline__ int32_t low32(int64_t v) {
  return static_cast<int32_t>(static_cast<uint32_t>(static_cast<uint64_t>(v)));
}
__device__ __forceinline__ int64_t mul_wrap(int64_t a, int64_t b) {
  return as_i64(static_cast<uint64_t>(a) * static_cast<uint64_t>(b));
}
__device__ __forceinline__ uint64_t prod64(int32_t a, int32_t b) {
  return static_cast<uint64_t>(static_cast<int64_t>(a) * static_cast<int64_t>(b));
}
__device__ __forceinline__ int32_t add32(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b));
}
__device__ __forceinline__ int32_t sub32(int32_t a, int32_t b) {
  return static_cast<int32_t>(static_cast<uint32_t>(a) - static_cast<uint32_t>(b));
}

// double -> int conversions with the x86 cvttsd2si result out of range.
__device__ __forceinline__ int32_t trunc_i32(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
  return static_cast<int32_t>(v);
}
__device__ __forceinline__ int64_t trunc_i64(double v) {
  if (!(v > -9223372036854775808.0 && v < 9223372036854775808.0)) return INT64_MIN;
  return static_cast<int64_t>(v);
}

// The dispatcher hands workgroups to the 8 XCDs round-robin (workgroup b runs
// on XCD b % 8), and each XCD has its own L2.  This renumbering gives every
// XCD a contiguous range of workgroup indices (a bijection of [0, gridDim.x)),
// so neighbouring rows, whose gathers share cache lines, run under one L2.
__device__ __forceinline__ int xcd_block() {
  constexpr int kXcd = 8;
  const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
  const int q = nb / kXcd, r = nb % kXcd, x = b % kXcd, k = b / kXcd;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask) {
  const int lo = __shfl_xor(static_cast<int>(static_cast<uint32_t>(v)), mask);
  const int hi = __shfl_xor(static_cast<int>(static_cast<uint32_t>(v >> 32)), mask);
  return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
         static_cast<uint32_t>(lo);
}
__device__ __forceinline__ uint64_t shfl_down_u64(uint64_t v, int d) {
  const int lo = __shfl_down(static_cast<int>(static_cast<uint32_t>(v)), d);
  const int hi = __shfl_down(static_cast<int>(static_cast<uint32_t>(v >> 32)), d);
  return (static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
         static_cast<uint32_t>(lo);
}

// Wave-wide (64 lanes) sum of an order-free uint64 value.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int m = 32; m > 0; m >>= 1) v += shfl_xor_u64(v, m);
  return v;
}

// Sum over a block of an order-free uint64 value (result valid in thread 0).
template <int kThreads>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* sh) {
  v = wave_sum_u64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  uint64_t r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) r += sh[w];
  __syncthreads();
  return r;
}

// The fixed halving tree of the blocked FP reduction over kThreads values
// held one per thread; the result is valid in thread 0.  s[t] += s[t+stride]
// for stride = kThreads/2 .. 1, strides below 64 inside the first wave with
// __shfl_down, which pairs lane t with lane t + stride exactly like the tree.
template <int kThreads, typename T>
__device__ __forceinline__ T block_tree(T v, T* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int st = kThreads / 2; st >= 64; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + st];
    __syncthreads();
  }
  T r = 0;
  if (threadIdx.x < 64) {
    r = sh[threadIdx.x];
#pragma unroll
    for (int st = 32; st > 0; st >>= 1) r = r + __shfl_down(r, st);
  }
  __syncthreads();
  return r;
}

// Last-block-done handshake: every block publishes its partial, then the last
// block to arrive sees all of them.  Returns true in the last block (all
// threads).  Resets the counter for the next launch.
__device__ __forceinline__ bool last_block(unsigned* counter, int* sh_flag) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned prev = atomicAdd(counter, 1u);
    *sh_flag = (prev == gridDim.x - 1);
  }
  __syncthreads();
  const bool last = *sh_flag != 0;
  if (last) {
    __threadfence();
    if (threadIdx.x == 0) atomicExch(counter, 0u);
  }
  return last;
}

}  // namespace sc23ad::dev
