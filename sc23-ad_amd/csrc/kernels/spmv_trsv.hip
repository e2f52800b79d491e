// Sparse matrix-vector products and level-scheduled triangular solves.
//
// Integer kernels (CSR int32 values, int64 accumulation, arithmetic shift;
// Library/matrix/spmat.hpp:77-86,128-137) are order-free, so a row may be
// split across lanes.  FP kernels (CSR<T>::SpMV spmat.hpp:43-52 and the
// triangular solves) keep each row's sum strictly in column order, one thread
// per row, so that the device rounds exactly like the CPU.
#include "dev_common.hpp"
#include "launch.hpp"

namespace sc23ad::dev {

namespace {

constexpr int kBlock = 256;

template <int G, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_int(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const int32_t* __restrict__ val, int bit, const int32_t* __restrict__ x,
    const int32_t* z, int32_t* y) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int row = gid / G;
  const int lane = gid & (G - 1);
  uint64_t acc = 0;
  if (row < n) {
    const int rb = rptr[row], re = rptr[row + 1];
    for (int p = rb + lane; p < re; p += G) acc += prod64(val[p], x[cind[p]]);
  }
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) acc += shfl_xor_u64(acc, m);
  if (row < n && lane == 0) {
    const int32_t s = low32(as_i64(acc) >> bit);
    if (MODE == 0) y[row] = s;
    else if (MODE == 1) y[row] = sub32(z[row], s);
    else y[row] = add32(z[row], s);
  }
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_fp(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const T* __restrict__ val, const T* __restrict__ x, const T* z, T* y) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  T t = 0;
  const int re = rptr[row + 1];
  for (int p = rptr[row]; p < re; ++p) t += val[p] * x[cind[p]];
  if (MODE == 0) y[row] = t;
  else if (MODE == 1) y[row] = z[row] - t;
  else y[row] = z[row] + t;
}

template <int G, int MODE>
hipError_t spmv_int_g(const DCsr& a, const void* x, const void* z, void* y,
                      hipStream_t s) {
  const long long threads = static_cast<long long>(a.n) * G;
  const int grid = static_cast<int>((threads + kBlock - 1) / kBlock);
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_spmv_int<G, MODE>), dim3(grid), dim3(kBlock), 0, s, a.n,
                     a.rptr, a.cind, static_cast<const int32_t*>(a.val), a.bit,
                     static_cast<const int32_t*>(x), static_cast<const int32_t*>(z),
                     static_cast<int32_t*>(y));
  return hipGetLastError();
}

template <int MODE>
hipError_t spmv_int_mode(const DCsr& a, const void* x, const void* z, void* y,
                         hipStream_t s) {
  switch (a.group) {
    case 1: return spmv_int_g<1, MODE>(a, x, z, y, s);
    case 2: return spmv_int_g<2, MODE>(a, x, z, y, s);
    case 4: return spmv_int_g<4, MODE>(a, x, z, y, s);
    case 8: return spmv_int_g<8, MODE>(a, x, z, y, s);
    case 16: return spmv_int_g<16, MODE>(a, x, z, y, s);
    case 32: return spmv_int_g<32, MODE>(a, x, z, y, s);
    default: return spmv_int_g<64, MODE>(a, x, z, y, s);
  }
}

template <typename T, int MODE>
hipError_t spmv_fp(const DCsr& a, const void* x, const void* z, void* y,
                   hipStream_t s) {
  const int grid = (a.n + kBlock - 1) / kBlock;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((k_spmv_fp<T, MODE>), dim3(grid), dim3(kBlock), 0, s, a.n,
                     a.rptr, a.cind, static_cast<const T*>(a.val),
                     static_cast<const T*>(x), static_cast<const T*>(z),
                     static_cast<T*>(y));
  return hipGetLastError();
}

// ---- triangular rows --------------------------------------------------------

template <TriKind K>
__device__ __forceinline__ int32_t tri_row_int(int i, const int* __restrict__ rptr,
                                               const int* __restrict__ cind,
                                               const int32_t* __restrict__ v,
                                               int bit, int dbit,
                                               const int32_t* __restrict__ in,
                                               const int32_t* out) {
  const int rb = rptr[i], re = rptr[i + 1];
  uint64_t t = static_cast<uint64_t>(static_cast<int64_t>(in[i])) << bit;
  if (K == TriKind::kLD) {
    for (int p = rb; p < re - 1; ++p) t -= prod64(v[p], out[cind[p]]);
    return low32(mul_wrap(as_i64(t) >> bit, v[re - 1]) >> dbit);
  } else if (K == TriKind::kL) {
    for (int p = rb; p < re; ++p) t -= prod64(v[p], out[cind[p]]);
    return low32(as_i64(t) >> bit);
  } else {
    for (int p = re - 1; p >= rb + 1; --p) t -= prod64(v[p], out[cind[p]]);
    return low32(mul_wrap(as_i64(t) >> bit, v[rb]) >> dbit);
  }
}

template <TriKind K, typename T>
__device__ __forceinline__ T tri_row_fp(int i, const int* __restrict__ rptr,
                                        const int* __restrict__ cind,
                                        const T* __restrict__ v,
                                        const T* __restrict__ in, const T* out) {
  const int rb = rptr[i], re = rptr[i + 1];
  T t = in[i];
  if (K == TriKind::kLD) {
    int p = rb;
    for (; p < re - 1; ++p) t -= v[p] * out[cind[p]];
    return t * v[p];
  } else if (K == TriKind::kL) {
    for (int p = rb; p < re; ++p) t -= v[p] * out[cind[p]];
    return t;
  } else {
    int p = re - 1;
    for (; p >= rb + 1; --p) t -= v[p] * out[cind[p]];
    return t * v[p];
  }
}

template <typename T, TriKind K>
struct Row;
template <TriKind K>
struct Row<int32_t, K> {
  __device__ static int32_t eval(int i, const int* rptr, const int* cind,
                                 const void* v, int bit, int dbit,
                                 const int32_t* in, const int32_t* out) {
    return tri_row_int<K>(i, rptr, cind, static_cast<const int32_t*>(v), bit,
                          dbit, in, out);
  }
};
template <TriKind K>
struct Row<double, K> {
  __device__ static double eval(int i, const int* rptr, const int* cind,
                                const void* v, int, int, const double* in,
                                const double* out) {
    return tri_row_fp<K, double>(i, rptr, cind, static_cast<const double*>(v), in,
                                 out);
  }
};
template <TriKind K>
struct Row<float, K> {
  __device__ static float eval(int i, const int* rptr, const int* cind,
                               const void* v, int, int, const float* in,
                               const float* out) {
    return tri_row_fp<K, float>(i, rptr, cind, static_cast<const float*>(v), in,
                                out);
  }
};

constexpr int kChainBlock = 512;

// One workgroup per chain; levels of the chain one after another, separated
// by a workgroup barrier (rows of a level only read rows of earlier levels).
template <typename T, TriKind K>
__global__ __launch_bounds__(kChainBlock) void k_tri_chain(
    const int* __restrict__ rptr, const int* __restrict__ cind, const void* val,
    int bit, int dbit, const int* __restrict__ perm,
    const int* __restrict__ lev_ptr, const int* __restrict__ chain_lev,
    const T* __restrict__ in, T* out) {
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  for (int L = lb; L < le; ++L) {
    const int s = lev_ptr[L], e = lev_ptr[L + 1];
    for (int idx = s + threadIdx.x; idx < e; idx += kChainBlock) {
      const int i = perm[idx];
      out[i] = Row<T, K>::eval(i, rptr, cind, val, bit, dbit, in, out);
    }
    __syncthreads();
  }
}

// All rows of one level, many workgroups.
template <typename T, TriKind K>
__global__ __launch_bounds__(kBlock) void k_tri_level(
    const int* __restrict__ rptr, const int* __restrict__ cind, const void* val,
    int bit, int dbit, const int* __restrict__ perm, int row_begin, int row_end,
    const T* __restrict__ in, T* out) {
  const int idx = row_begin + blockIdx.x * kBlock + threadIdx.x;
  if (idx >= row_end) return;
  const int i = perm[idx];
  out[i] = Row<T, K>::eval(i, rptr, cind, val, bit, dbit, in, out);
}

template <typename T, TriKind K>
hipError_t tri_seg(const DCsr& a, const int* perm, const int* lev_ptr,
                   const TriSegment& seg, const void* in, void* out,
                   hipStream_t s) {
  if (seg.chain_kernel) {
    if (seg.nchain <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_tri_chain<T, K>), dim3(seg.nchain), dim3(kChainBlock), 0,
                       s, a.rptr, a.cind, a.val, a.bit, a.dbit, perm, lev_ptr,
                       seg.chain_lev, static_cast<const T*>(in),
                       static_cast<T*>(out));
  } else {
    const int rows = seg.row_end - seg.row_begin;
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_tri_level<T, K>), dim3((rows + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, a.rptr, a.cind, a.val, a.bit, a.dbit,
                       perm, seg.row_begin, seg.row_end, static_cast<const T*>(in),
                       static_cast<T*>(out));
  }
  return hipGetLastError();
}

template <typename T>
hipError_t tri_kind(TriKind kind, const DCsr& a, const int* perm,
                    const int* lev_ptr, const TriSegment& seg, const void* in,
                    void* out, hipStream_t s) {
  switch (kind) {
    case TriKind::kLD: return tri_seg<T, TriKind::kLD>(a, perm, lev_ptr, seg, in, out, s);
    case TriKind::kL: return tri_seg<T, TriKind::kL>(a, perm, lev_ptr, seg, in, out, s);
    default: return tri_seg<T, TriKind::kDU>(a, perm, lev_ptr, seg, in, out, s);
  }
}

}  // namespace

hipError_t launch_spmv(Elem e, const DCsr& a, SpmvMode mode, const void* x,
                       const void* z, void* y, hipStream_t s) {
  const int md = static_cast<int>(mode);
  if (e == Elem::kI32) {
    if (md == 0) return spmv_int_mode<0>(a, x, z, y, s);
    if (md == 1) return spmv_int_mode<1>(a, x, z, y, s);
    return spmv_int_mode<2>(a, x, z, y, s);
  }
  if (e == Elem::kF64) {
    if (md == 0) return spmv_fp<double, 0>(a, x, z, y, s);
    if (md == 1) return spmv_fp<double, 1>(a, x, z, y, s);
    return spmv_fp<double, 2>(a, x, z, y, s);
  }
  if (md == 0) return spmv_fp<float, 0>(a, x, z, y, s);
  if (md == 1) return spmv_fp<float, 1>(a, x, z, y, s);
  return spmv_fp<float, 2>(a, x, z, y, s);
}

hipError_t launch_trsv_segment(Elem e, TriKind kind, const DCsr& a,
                               const int* perm, const int* lev_ptr,
                               const TriSegment& seg, const void* in, void* out,
                               hipStream_t s) {
  if (e == Elem::kI32) return tri_kind<int32_t>(kind, a, perm, lev_ptr, seg, in, out, s);
  if (e == Elem::kF64) return tri_kind<double>(kind, a, perm, lev_ptr, seg, in, out, s);
  return tri_kind<float>(kind, a, perm, lev_ptr, seg, in, out, s);
}

}  // namespace sc23ad::dev
