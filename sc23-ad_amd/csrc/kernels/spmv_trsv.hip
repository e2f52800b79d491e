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

// ---- blocked single-workgroup solve for the coarsest grid ------------------
//
// The coarsest factors (n <= a few thousand, dense-ish after the complete LU)
// form one long dependency chain.  One workgroup keeps the solution and the
// row accumulators in LDS and walks 64-column blocks: the block's diagonal
// tile is solved by one wave as a column sweep (lane r owns row r; after
// column c is final it is broadcast and every dependent row subtracts its
// term), then all waves subtract the block's contributions from the rows
// still to come.  Each row thus receives its terms in the same column order
// as the sequential solve (ascending for L / LD, descending for DU), so FP
// results round exactly as the CPU's; integer sums are order-free anyway.
// bptr[i * (nb + 1) + b] is the first dependency entry of row i whose column
// is >= 64 b.

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

// Broadcast lane `src` (wave-uniform) to the whole wave with v_readlane.
template <typename T>
__device__ __forceinline__ T bcast_lane(T v, int src) {
  if constexpr (sizeof(T) == 4) {
    int x;
    __builtin_memcpy(&x, &v, 4);
    x = __builtin_amdgcn_readlane(x, src);
    T r;
    __builtin_memcpy(&r, &x, 4);
    return r;
  } else {
    int x[2];
    __builtin_memcpy(x, &v, 8);
    x[0] = __builtin_amdgcn_readlane(x[0], src);
    x[1] = __builtin_amdgcn_readlane(x[1], src);
    T r;
    __builtin_memcpy(&r, x, 8);
    return r;
  }
}

template <typename T, TriKind K>
__global__ __launch_bounds__(kBlkThreads) void k_tri_blocked(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind, const void* valv,
    int bit, int dbit, const int* __restrict__ bptr, int nb, const T* __restrict__ in,
    T* __restrict__ out) {
  using Acc = typename AccOf<T>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  Acc* acc = reinterpret_cast<Acc*>(smem);
  T* xs = reinterpret_cast<T*>(acc + n);
  T* dinv = xs + n;
  uint64_t* msk = reinterpret_cast<uint64_t*>(
      (reinterpret_cast<uintptr_t>(dinv + n) + 15) & ~static_cast<uintptr_t>(15));
  T* tile = reinterpret_cast<T*>(msk + kTB);
  const T* val = static_cast<const T*>(valv);
  constexpr bool fwd = (K != TriKind::kDU);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int nw = kBlkThreads / 64;
  const int stride_b = nb + 1;
  for (int i = tid; i < n; i += kBlkThreads) {
    acc[i] = acc_init<T>(in[i], bit);
    if (K == TriKind::kLD) dinv[i] = val[rptr[i + 1] - 1];
    if (K == TriKind::kDU) dinv[i] = val[rptr[i]];
  }
  __syncthreads();
  for (int q = 0; q < nb; ++q) {
    const int b = fwd ? q : nb - 1 - q;
    const int r0 = b * kTB;
    const int r1 = min(r0 + kTB, n);
    const int rows = r1 - r0;
    // (a) the diagonal tile: dependency entries of rows r0..r1-1 in block b
    for (int rr = wid; rr < rows; rr += nw) {
      const int i = r0 + rr;
      const int pb = bptr[i * stride_b + b], pe = bptr[i * stride_b + b + 1];
      const int p = pb + lane;
      uint64_t bitv = 0;
      if (p < pe) {
        const int c = cind[p] - r0;
        tile[c * kTB + rr] = val[p];
        bitv = 1ull << c;
      }
#pragma unroll
      for (int m = 32; m > 0; m >>= 1) bitv |= shfl_xor_u64(bitv, m);
      if (lane == 0) msk[rr] = bitv;
    }
    __syncthreads();
    // (b) one wave: column sweep over the tile
    if (wid == 0) {
      Acc t = lane < rows ? acc[r0 + lane] : Acc(0);
      const uint64_t mine = lane < rows ? msk[lane] : 0;
      const T dv = lane < rows ? dinv[r0 + lane] : T(0);
      T xv = 0;
      for (int st = 0; st < rows; ++st) {
        const int c = fwd ? st : rows - 1 - st;
        T xl = 0;
        if (lane == c) {
          xl = acc_final<T, K>(t, dv, bit, dbit);
          xv = xl;
        }
        const T xc = bcast_lane<T>(xl, c);
        if ((mine >> c) & 1ull) t = acc_sub<T>(t, tile[c * kTB + lane], xc);
      }
      if (lane < rows) {
        xs[r0 + lane] = xv;
        out[r0 + lane] = xv;
      }
    }
    __syncthreads();
    // (c) all waves: rows still to come take block b's terms
    const int lo = fwd ? r1 : 0;
    const int hi = fwd ? n : r0;
    for (int i = lo + tid; i < hi; i += kBlkThreads) {
      const int pb = bptr[i * stride_b + b], pe = bptr[i * stride_b + b + 1];
      if (pb == pe) continue;
      Acc t = acc[i];
      // batches of kBatch independent loads, applied in the row's order
      constexpr int kBatch = 8;
      const int cnt = pe - pb;
      for (int q0 = 0; q0 < cnt; q0 += kBatch) {
        T vv[kBatch];
        int cc[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
          const int q = q0 + u;
          if (q < cnt) {
            const int p = fwd ? pb + q : pe - 1 - q;
            vv[u] = val[p];
            cc[u] = cind[p];
          }
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
          if (q0 + u < cnt) t = acc_sub<T>(t, vv[u], xs[cc[u]]);
      }
      acc[i] = t;
    }
    __syncthreads();
  }
}

template <typename T, TriKind K>
hipError_t tri_blocked(const DCsr& a, const int* bptr, int nb, const void* in, void* out,
                       hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const size_t lds = tri_blocked_lds(a.n, sizeof(T), sizeof(typename AccOf<T>::type));
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tri_blocked<T, K>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(kTriBlockedMaxLds));
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_tri_blocked<T, K>), dim3(1), dim3(kBlkThreads), lds, s, a.n, a.rptr,
                     a.cind, a.val, a.bit, a.dbit, bptr, nb, static_cast<const T*>(in),
                     static_cast<T*>(out));
  return hipGetLastError();
}

// ---- level-ordered chain solve with prefetch ---------------------------------
//
// The factor is re-laid out in level order: position idx holds row lrow[idx]
// with its dependency entries lcol/lval[lbeg[idx] .. lbeg[idx+1]) in the order
// the sequential solve subtracts them, and its stored inverse diagonal in
// ldiag[idx].  One workgroup walks the levels of one chain; while level L is
// being computed, the row data of level L+1 (which does not depend on the
// solution) is already loaded into registers, so the only loads left on the
// critical path are the solution values of earlier levels.  When the chain's
// rows fit, its solution is also kept in LDS.

constexpr int kPre = 8;  // dependency entries held in registers per row

template <typename T>
struct PreRow {
  int row = -1, beg = 0, end = 0;
  T in = 0, dinv = 0;
  int col[kPre];
  T val[kPre];
};

template <typename T, TriKind K>
__device__ __forceinline__ void pre_load(PreRow<T>& r, int idx, int lim,
                                         const int* __restrict__ lrow,
                                         const int* __restrict__ lbeg,
                                         const int* __restrict__ lcol,
                                         const T* __restrict__ lval,
                                         const T* __restrict__ ldiag,
                                         const T* __restrict__ in) {
  if (idx >= lim) {
    r.row = -1;
    return;
  }
  r.row = lrow[idx];
  r.beg = lbeg[idx];
  r.end = lbeg[idx + 1];
  r.in = in[r.row];
  if (K != TriKind::kL) r.dinv = ldiag[idx];
#pragma unroll
  for (int q = 0; q < kPre; ++q) {
    const int p = r.beg + q;
    if (p < r.end) {
      r.col[q] = lcol[p];
      r.val[q] = lval[p];
    }
  }
}

template <typename T, TriKind K, bool kLds>
__device__ __forceinline__ T pre_eval(const PreRow<T>& r, int bit, int dbit,
                                      const int* __restrict__ lcol,
                                      const T* __restrict__ lval, const T* out,
                                      const T* xs, int r0) {
  using Acc = typename AccOf<T>::type;
  Acc t = acc_init<T>(r.in, bit);
  auto X = [&](int c) -> T { return kLds ? xs[c - r0] : out[c]; };
#pragma unroll
  for (int q = 0; q < kPre; ++q)
    if (r.beg + q < r.end) t = acc_sub<T>(t, r.val[q], X(r.col[q]));
  for (int p = r.beg + kPre; p < r.end; ++p) t = acc_sub<T>(t, lval[p], X(lcol[p]));
  return acc_final<T, K>(t, r.dinv, bit, dbit);
}

template <typename T, TriKind K, bool kLds>
__global__ __launch_bounds__(kChainBlock) void k_tri_chain_pre(
    const int* __restrict__ lrow, const int* __restrict__ lbeg,
    const int* __restrict__ lcol, const T* __restrict__ lval,
    const T* __restrict__ ldiag, int bit, int dbit, const int* __restrict__ lev_ptr,
    const int* __restrict__ chain_lev, const int* __restrict__ chain_row0,
    const T* __restrict__ in, T* out) {
  extern __shared__ __align__(16) unsigned char smem_c[];
  T* xs = reinterpret_cast<T*>(smem_c);
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int r0 = kLds ? chain_row0[blockIdx.x] : 0;
  PreRow<T> nxt;
  if (lb < le)
    pre_load<T, K>(nxt, lev_ptr[lb] + threadIdx.x, lev_ptr[lb + 1], lrow, lbeg, lcol,
                   lval, ldiag, in);
  for (int L = lb; L < le; ++L) {
    const PreRow<T> cur = nxt;
    const int s = lev_ptr[L], e = lev_ptr[L + 1];
    if (L + 1 < le)
      pre_load<T, K>(nxt, lev_ptr[L + 1] + threadIdx.x, lev_ptr[L + 2], lrow, lbeg, lcol,
                     lval, ldiag, in);
    if (cur.row >= 0) {
      const T v = pre_eval<T, K, kLds>(cur, bit, dbit, lcol, lval, out, xs, r0);
      out[cur.row] = v;
      if (kLds) xs[cur.row - r0] = v;
    }
    // rows beyond one per thread (large levels)
    for (int idx = s + kChainBlock + threadIdx.x; idx < e; idx += kChainBlock) {
      PreRow<T> r;
      pre_load<T, K>(r, idx, e, lrow, lbeg, lcol, lval, ldiag, in);
      const T v = pre_eval<T, K, kLds>(r, bit, dbit, lcol, lval, out, xs, r0);
      out[r.row] = v;
      if (kLds) xs[r.row - r0] = v;
    }
    __syncthreads();
  }
}

// One level, many workgroups, level-ordered layout.
template <typename T, TriKind K>
__global__ __launch_bounds__(kBlock) void k_tri_level_pre(
    const int* __restrict__ lrow, const int* __restrict__ lbeg,
    const int* __restrict__ lcol, const T* __restrict__ lval,
    const T* __restrict__ ldiag, int bit, int dbit, int row_begin, int row_end,
    const T* __restrict__ in, T* out) {
  const int idx = row_begin + blockIdx.x * kBlock + threadIdx.x;
  if (idx >= row_end) return;
  PreRow<T> r;
  pre_load<T, K>(r, idx, row_end, lrow, lbeg, lcol, lval, ldiag, in);
  out[r.row] = pre_eval<T, K, false>(r, bit, dbit, lcol, lval, out, nullptr, 0);
}

template <typename T, TriKind K>
hipError_t tri_pre_seg(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                       const void* in, void* out, hipStream_t s) {
  const T* vin = static_cast<const T*>(in);
  T* vout = static_cast<T*>(out);
  const T* lval = static_cast<const T*>(L.lval);
  const T* ldiag = static_cast<const T*>(L.ldiag);
  if (seg.chain_kernel) {
    if (seg.nchain <= 0) return hipSuccess;
    if (L.lds_rows > 0) {
      const size_t lds = sizeof(T) * static_cast<size_t>(L.lds_rows);
      static bool attr = false;
      if (!attr) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&k_tri_chain_pre<T, K, true>),
            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kTriBlockedMaxLds));
        if (e != hipSuccess) return e;
        attr = true;
      }
      hipLaunchKernelGGL((k_tri_chain_pre<T, K, true>), dim3(seg.nchain), dim3(kChainBlock),
                         lds, s, L.lrow, L.lbeg, L.lcol, lval, ldiag, bit, dbit, L.lev_ptr,
                         seg.chain_lev, L.chain_row0, vin, vout);
    } else {
      hipLaunchKernelGGL((k_tri_chain_pre<T, K, false>), dim3(seg.nchain),
                         dim3(kChainBlock), 0, s, L.lrow, L.lbeg, L.lcol, lval, ldiag, bit,
                         dbit, L.lev_ptr, seg.chain_lev, L.chain_row0, vin, vout);
    }
  } else {
    const int rows = seg.row_end - seg.row_begin;
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_tri_level_pre<T, K>), dim3((rows + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, L.lrow, L.lbeg, L.lcol, lval, ldiag, bit, dbit,
                       seg.row_begin, seg.row_end, vin, vout);
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

size_t tri_blocked_lds(int n, size_t elem, size_t acc) {
  size_t b = acc * n + 2 * elem * n;
  b = (b + 15) & ~static_cast<size_t>(15);
  return b + 8 * kTB + elem * kTB * kTB;
}

hipError_t launch_trsv_blocked(Elem e, TriKind kind, const DCsr& a, const int* bptr,
                               int nb, const void* in, void* out, hipStream_t s) {
  if (e == Elem::kI32) {
    if (kind == TriKind::kL) return tri_blocked<int32_t, TriKind::kL>(a, bptr, nb, in, out, s);
    if (kind == TriKind::kLD) return tri_blocked<int32_t, TriKind::kLD>(a, bptr, nb, in, out, s);
    return tri_blocked<int32_t, TriKind::kDU>(a, bptr, nb, in, out, s);
  }
  if (e == Elem::kF64) {
    if (kind == TriKind::kL) return tri_blocked<double, TriKind::kL>(a, bptr, nb, in, out, s);
    if (kind == TriKind::kLD) return tri_blocked<double, TriKind::kLD>(a, bptr, nb, in, out, s);
    return tri_blocked<double, TriKind::kDU>(a, bptr, nb, in, out, s);
  }
  if (kind == TriKind::kL) return tri_blocked<float, TriKind::kL>(a, bptr, nb, in, out, s);
  if (kind == TriKind::kLD) return tri_blocked<float, TriKind::kLD>(a, bptr, nb, in, out, s);
  return tri_blocked<float, TriKind::kDU>(a, bptr, nb, in, out, s);
}

hipError_t launch_trsv_layout(Elem e, TriKind kind, const TriLayout& L,
                              const TriSegment& seg, int bit, int dbit, const void* in,
                              void* out, hipStream_t s) {
  if (e == Elem::kI32) {
    if (kind == TriKind::kL) return tri_pre_seg<int32_t, TriKind::kL>(L, seg, bit, dbit, in, out, s);
    if (kind == TriKind::kLD) return tri_pre_seg<int32_t, TriKind::kLD>(L, seg, bit, dbit, in, out, s);
    return tri_pre_seg<int32_t, TriKind::kDU>(L, seg, bit, dbit, in, out, s);
  }
  if (e == Elem::kF64) {
    if (kind == TriKind::kL) return tri_pre_seg<double, TriKind::kL>(L, seg, bit, dbit, in, out, s);
    if (kind == TriKind::kLD) return tri_pre_seg<double, TriKind::kLD>(L, seg, bit, dbit, in, out, s);
    return tri_pre_seg<double, TriKind::kDU>(L, seg, bit, dbit, in, out, s);
  }
  if (kind == TriKind::kL) return tri_pre_seg<float, TriKind::kL>(L, seg, bit, dbit, in, out, s);
  if (kind == TriKind::kLD) return tri_pre_seg<float, TriKind::kLD>(L, seg, bit, dbit, in, out, s);
  return tri_pre_seg<float, TriKind::kDU>(L, seg, bit, dbit, in, out, s);
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
