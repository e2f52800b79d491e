// Sparse matrix-vector products and level-scheduled triangular solves.
//
// Integer kernels (CSR int32 values, int64 accumulation, arithmetic shift;
// Library/matrix/spmat.hpp:77-86,128-137) are order-free, so a row may be
// split across lanes.  FP kernels (CSR<T>::SpMV spmat.hpp:43-52 and the
// triangular solves) keep each row's sum strictly in column order, one thread
// per row, so that the device rounds exactly like the CPU.
#include <hip/hip_ext.h>

#include <cstdlib>
#include <type_traits>

#include "dev_common.hpp"
#include "launch.hpp"
#include "tri_common.hpp"

namespace sc23ad::dev {

namespace {
thread_local DispatchEvents t_armed;
}  // namespace

void arm_dispatch_events(hipEvent_t start, hipEvent_t stop) { t_armed = {start, stop}; }

DispatchEvents take_dispatch_events() {
  const DispatchEvents e = t_armed;
  t_armed = {};
  return e;
}

namespace {

constexpr int kBlock = 256;

// SpMV workgroups renumbered per XCD (xcd_block) when SC23AD_XCD_REMAP=1.
// Off by default: the fine-level SpMV ran 4 % slower with it.  The SpMV
// tuner also tries it per matrix on the row-sorted SELL copies (restriction,
// prolongation), whose scattered rows of one window then stay under one L2
// (at 128^3 their HBM fetches fell 30-45 %).
int xcd_remap() {
  static const int on = [] {
    const char* e = std::getenv("SC23AD_XCD_REMAP");
    return (e && e[0] == '1') ? 1 : 0;
  }();
  return on;
}

// A launch that honours armed dispatch events (arm_dispatch_events).
template <typename F, typename... A>
void launch_timed(F kern, dim3 grid, dim3 block, uint32_t lds, hipStream_t s, A... args) {
  const DispatchEvents ev = take_dispatch_events();
  if (ev.start)
    hipExtLaunchKernelGGL(kern, grid, block, lds, s, ev.start, ev.stop, 0u, args...);
  else
    hipLaunchKernelGGL(kern, grid, block, lds, s, args...);
}

// CSR integer SpMV, G lanes per row.  Each lane takes the row's entries
// lane, lane + G, ... eight at a time with every load unconditional (clamped
// to the chunk's first entry), so a chunk costs two memory latencies (entries,
// then the gathered x values) instead of two per entry.
template <int G, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_int(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const int32_t* __restrict__ val, int bit, const int32_t* __restrict__ x,
    const int32_t* z, int32_t* y, int remap) {
  const int blk = remap ? xcd_block() : static_cast<int>(blockIdx.x);
  const int gid = blk * kBlock + threadIdx.x;
  const int row = gid / G;
  const int lane = gid & (G - 1);
  uint64_t acc = 0;
  int32_t zv = 0;
  if (row < n) {
    if (MODE != 0 && lane == 0) zv = z[row];
    const int rb = rptr[row], re = rptr[row + 1];
    for (int p0 = rb + lane; p0 < re; p0 += 8 * G) {
      int cc[8];
      int32_t vv[8], xx[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int p = p0 + j * G < re ? p0 + j * G : p0;
        cc[j] = cind[p];
        vv[j] = val[p];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) xx[j] = x[cc[j]];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (p0 + j * G < re) acc += prod64(vv[j], xx[j]);
    }
  }
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) acc += shfl_xor_u64(acc, m);
  if (row < n && lane == 0) {
    const int32_t s = low32(as_i64(acc) >> bit);
    if (MODE == 0) y[row] = s;
    else if (MODE == 1) y[row] = sub32(zv, s);
    else y[row] = add32(zv, s);
  }
}

// SELL-64 SpMV, one thread per row.  Entries are read eight at a time with
// clamped indices (every load unconditional), padding contributes nothing.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_sell_int(
    int n, const long long* __restrict__ soff, const int* __restrict__ scol,
    const int32_t* __restrict__ sval, const int* __restrict__ srow, int bit,
    const int32_t* __restrict__ x, const int32_t* z, int32_t* y, int remap) {
  const int row = (remap ? xcd_block() : static_cast<int>(blockIdx.x)) * kBlock + threadIdx.x;
  if (row >= n) return;
  const int orow = srow ? srow[row] : row;
  const int32_t zv = MODE != 0 ? z[orow] : 0;  // issued before the entries
  const long long b = soff[row >> 6];
  const int w = static_cast<int>((soff[(row >> 6) + 1] - b) >> 6);
  const int* c = scol + b + (row & 63);
  const int32_t* v = sval + b + (row & 63);
  uint64_t acc = 0;
  for (int k0 = 0; k0 < w; k0 += 8) {
    int cc[8];
    int32_t vv[8], xx[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + j < w ? k0 + j : k0;
      cc[j] = c[64 * k];
      vv[j] = v[64 * k];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) xx[j] = x[cc[j] < 0 ? 0 : cc[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < w) acc += prod64(vv[j], xx[j]);
  }
  const int32_t s = low32(as_i64(acc) >> bit);
  if (MODE == 0) y[orow] = s;
  else if (MODE == 1) y[orow] = sub32(zv, s);
  else y[orow] = add32(zv, s);
}

// SELL-64 integer SpMV with 16-bit column offsets: entry k of row r holds
// col - r (square matrices whose entries all lie within 32767 of the
// diagonal; padding is offset 0 with value 0, an exact zero product).  Six
// bytes per entry instead of eight.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_sell16_int(
    int n, const long long* __restrict__ soff, const int16_t* __restrict__ scol,
    const int32_t* __restrict__ sval, int bit, const int32_t* __restrict__ x, const int32_t* z,
    int32_t* y, int remap) {
  const int row = (remap ? xcd_block() : static_cast<int>(blockIdx.x)) * kBlock + threadIdx.x;
  if (row >= n) return;
  const int32_t zv = MODE != 0 ? z[row] : 0;  // issued before the entries
  const long long b = soff[row >> 6];
  const int w = static_cast<int>((soff[(row >> 6) + 1] - b) >> 6);
  const int16_t* c = scol + b + (row & 63);
  const int32_t* v = sval + b + (row & 63);
  uint64_t acc = 0;
  for (int k0 = 0; k0 < w; k0 += 8) {
    int cc[8];
    int32_t vv[8], xx[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + j < w ? k0 + j : k0;
      cc[j] = row + c[64 * k];
      vv[j] = v[64 * k];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) xx[j] = x[cc[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < w) acc += prod64(vv[j], xx[j]);
  }
  const int32_t s = low32(as_i64(acc) >> bit);
  if (MODE == 0) y[row] = s;
  else if (MODE == 1) y[row] = sub32(zv, s);
  else y[row] = add32(zv, s);
}

// FP SELL-64 SpMV: the row's products are added in stored (column) order,
// as the CSR kernel adds them; padding is skipped, not added.
template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_sell_fp(
    int n, const long long* __restrict__ soff, const int* __restrict__ scol,
    const T* __restrict__ sval, const int* __restrict__ srow, const T* __restrict__ x,
    const T* z, T* y, int remap) {
  const int row = (remap ? xcd_block() : static_cast<int>(blockIdx.x)) * kBlock + threadIdx.x;
  if (row >= n) return;
  const int orow = srow ? srow[row] : row;
  const T zv = MODE != 0 ? z[orow] : T(0);  // issued before the entries
  const long long b = soff[row >> 6];
  const int w = static_cast<int>((soff[(row >> 6) + 1] - b) >> 6);
  const int* c = scol + b + (row & 63);
  const T* v = sval + b + (row & 63);
  T t = 0;
  for (int k0 = 0; k0 < w; k0 += 8) {
    int cc[8];
    T vv[8], xx[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + j < w ? k0 + j : k0;
      cc[j] = c[64 * k];
      vv[j] = v[64 * k];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) xx[j] = x[cc[j] < 0 ? 0 : cc[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (k0 + j < w && cc[j] >= 0) t += vv[j] * xx[j];
  }
  if (MODE == 0) y[orow] = t;
  else if (MODE == 1) y[orow] = zv - t;
  else y[orow] = zv + t;
}

template <typename T, int MODE>
__global__ __launch_bounds__(kBlock) void k_spmv_fp(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const T* __restrict__ val, const T* __restrict__ x, const T* z, T* y, int remap) {
  const int row = (remap ? xcd_block() : static_cast<int>(blockIdx.x)) * kBlock + threadIdx.x;
  if (row >= n) return;
  T t = 0;
  const int rb = rptr[row], re = rptr[row + 1];
  // Eight entries at a time, every load unconditional (clamped), so that a
  // chunk costs two memory latencies; the sum keeps the stored order.
  for (int p0 = rb; p0 < re; p0 += 8) {
    int cc[8];
    T vv[8], xx[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = p0 + j < re ? p0 + j : p0;
      cc[j] = cind[p];
      vv[j] = val[p];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) xx[j] = x[cc[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (p0 + j < re) t += vv[j] * xx[j];
  }
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
  launch_timed(&k_spmv_int<G, MODE>, dim3(grid), dim3(kBlock), 0, s, a.n, a.rptr, a.cind,
               static_cast<const int32_t*>(a.val), a.bit, static_cast<const int32_t*>(x),
               static_cast<const int32_t*>(z), static_cast<int32_t*>(y), xcd_remap());
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
  launch_timed(&k_spmv_fp<T, MODE>, dim3(grid), dim3(kBlock), 0, s, a.n, a.rptr, a.cind,
               static_cast<const T*>(a.val), static_cast<const T*>(x),
               static_cast<const T*>(z), static_cast<T*>(y), xcd_remap());
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

// Integer version of the blocked solve.  Integer row sums are order-free, so
// (i) the sweeping wave keeps its row's 64 tile values in registers and the
// 64-step sweep is fully unrolled, and (ii) the off-block update gives each
// row a whole wave: lane q takes the row's q-th entry in the block and the
// wave sums the products.
template <TriKind K>
__global__ __launch_bounds__(kBlkThreads) void k_tri_blocked_int(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const int32_t* __restrict__ val, int bit, int dbit, const int* __restrict__ bptr,
    int nb, const int32_t* __restrict__ in, int32_t* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* acc = reinterpret_cast<uint64_t*>(smem);
  int32_t* xs = reinterpret_cast<int32_t*>(acc + n);
  int32_t* dinv = xs + n;
  int32_t* tile = reinterpret_cast<int32_t*>(
      (reinterpret_cast<uintptr_t>(dinv + n) + 15) & ~static_cast<uintptr_t>(15));
  constexpr bool fwd = (K != TriKind::kDU);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  constexpr int nw = kBlkThreads / 64;
  const int stride_b = nb + 1;
  for (int i = tid; i < n; i += kBlkThreads) {
    acc[i] = static_cast<uint64_t>(static_cast<int64_t>(in[i])) << bit;
    if (K == TriKind::kLD) dinv[i] = val[rptr[i + 1] - 1];
    if (K == TriKind::kDU) dinv[i] = val[rptr[i]];
  }
  __syncthreads();
  for (int q = 0; q < nb; ++q) {
    const int b = fwd ? q : nb - 1 - q;
    const int r0 = b * kTB;
    const int r1 = min(r0 + kTB, n);
    const int rows = r1 - r0;
    // (a) the diagonal tile, column-major, zero where no entry
    for (int idx = tid; idx < kTB * kTB; idx += kBlkThreads) tile[idx] = 0;
    __syncthreads();
    for (int rr = wid; rr < rows; rr += nw) {
      const int i = r0 + rr;
      const int pb = bptr[i * stride_b + b], pe = bptr[i * stride_b + b + 1];
      const int p = pb + lane;
      if (p < pe) tile[(cind[p] - r0) * kTB + rr] = val[p];
    }
    __syncthreads();
    // (b) one wave sweeps the tile with its operands in registers
    if (wid == 0) {
      int32_t reg[kTB];
#pragma unroll
      for (int c = 0; c < kTB; ++c) reg[c] = tile[c * kTB + lane];
      uint64_t t = lane < rows ? acc[r0 + lane] : 0;
      const int32_t dv = lane < rows ? dinv[r0 + lane] : 0;
      int32_t xv = 0;
#pragma unroll
      for (int st = 0; st < kTB; ++st) {
        const int c = fwd ? st : kTB - 1 - st;
        if (c < rows) {
          const int32_t mine = acc_final<int32_t, K>(t, dv, bit, dbit);
          if (lane == c) xv = mine;
          const int32_t xc = __builtin_amdgcn_readlane(mine, c);
          t -= prod64(reg[c], xc);
        }
      }
      if (lane < rows) {
        xs[r0 + lane] = xv;
        out[r0 + lane] = xv;
      }
    }
    __syncthreads();
    // (c) one wave per remaining row
    const int lo = fwd ? r1 : 0;
    const int hi = fwd ? n : r0;
    for (int i = lo + wid; i < hi; i += nw) {
      const int pb = bptr[i * stride_b + b], pe = bptr[i * stride_b + b + 1];
      if (pb == pe) continue;
      const int p = pb + lane;
      uint64_t prod = 0;
      if (p < pe) prod = prod64(val[p], xs[cind[p]]);
      prod = wave_sum_u64(prod);
      if (lane == 0) acc[i] -= prod;
    }
    __syncthreads();
  }
}

// Tiled integer coarse solve, one workgroup per 64-row block.  Workgroup b
// owns rows 64b .. 64b+63.  Its factor entries are stored as dense 64 x 64
// tiles (zeros where the factor has no entry), column-major (entry (row r,
// column c) of a tile at c * 64 + r), in processing order: the off-diagonal
// tiles of the blocks it depends on (forward: 0 .. b-1; backward: nb-1 ..
// b+1), then its diagonal tile.  A solved row is published as one 64-bit
// word {epoch, value} (xp[row], an agent-scope atomic store), so a consumer
// learns that a value is ready and reads it in the same load.  For each
// dependency block, wave 0 polls the block's 64 words until all carry the
// current epoch, the workgroup subtracts the tile's product (four waves, 16
// columns each; integer sums are order-free), then wave 0 sweeps the
// diagonal tile: at step c lane c finishes its row and the value is broadcast
// to the other lanes.  Dependencies point to blocks earlier in processing
// order (forward: lower blocks; backward: higher ones), and workgroups take
// their block from a ticket counter in that order, so a workgroup only ever
// waits for workgroups that started before it, whatever order the hardware
// starts them in and whatever else shares the GPU.  A launch that runs out
// of its wait budget (tile_wait_over) records an error and gives up instead
// of hanging.

template <TriKind K>
__global__ __launch_bounds__(kTileThreads) void k_tri_tiles_int(
    int n, int nb, const int32_t* __restrict__ tiles, const long long* __restrict__ toff,
    const int32_t* __restrict__ dinv_g, int bit, int dbit, const int32_t* __restrict__ in,
    int32_t* out, unsigned long long* xp, int epoch, int* err, unsigned* tickets) {
  constexpr bool fwd = (K != TriKind::kDU);
  __shared__ uint64_t part[kTileThreads / 64][64];
  __shared__ int32_t xsh[64];
  __shared__ int ticket_sh;
  const long long t0 = wall_clock64();
  const int pb = tile_ticket(tickets, &ticket_sh);
  const int b = fwd ? pb : nb - 1 - pb;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = b * kTB;
  const int rows = min(kTB, n - r0);
  const int32_t* T = tiles + toff[b];
  const int ndep = fwd ? b : nb - 1 - b;
  uint64_t t = 0;
  if (w == 0 && lane < rows) t = static_cast<uint64_t>(static_cast<int64_t>(in[r0 + lane])) << bit;
  // the diagonal tile and the row scales depend on no other block: wave 0
  // loads them before its first wait, so their latency stays off the chain
  int32_t reg[kTB];
  int32_t dv = 0;
  if (w == 0) {
    const int32_t* Td = T + static_cast<size_t>(ndep) * (kTB * kTB);
#pragma unroll
    for (int c = 0; c < kTB; ++c) reg[c] = Td[c * kTB + lane];
    dv = (K != TriKind::kL && lane < rows) ? dinv_g[r0 + lane] : 0;
  }
  for (int k = 0; k < ndep; ++k) {
    const int bp = fwd ? k : nb - 1 - k;
    const int32_t* Tk = T + static_cast<size_t>(k) * (kTB * kTB);
    int32_t tv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) tv[j] = Tk[(16 * w + j) * kTB + lane];
    if (w == 0) {
      // two polls in flight: the next load is issued before the previous
      // one is checked, so a block's words are seen about half a round trip
      // sooner than with one poll at a time
      const int c = bp * kTB + lane;
      const unsigned long long none =
          static_cast<unsigned long long>(static_cast<unsigned>(epoch)) << 32;
      // every lane loads (rows past the end read row 0's word and ignore
      // it): a load under a divergent branch would make the compiler wait
      // for both polls where the paths join
      const int cl = c < n ? c : 0;
      auto poll = [&]() {
        const unsigned long long u =
            __hip_atomic_load(&xp[cl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return c < n ? u : none;
      };
      auto ready = [&](unsigned long long u) {
        return __all(static_cast<int>(u >> 32) == epoch);
      };
      unsigned long long v = 0, va = poll();
      int spins = 0;
      while (true) {
        __builtin_amdgcn_s_sleep(1);
        const unsigned long long vb = poll();
        if (ready(va)) {
          v = va;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        va = poll();
        if (ready(vb)) {
          v = vb;
          break;
        }
        if ((++spins & 15) == 0 && tile_wait_over(t0, err)) {
          if (lane == 0) atomicAdd(err, 1);
          v = vb;
          break;
        }
      }
      xsh[lane] = c < n ? static_cast<int32_t>(static_cast<uint32_t>(v)) : 0;
    }
    __syncthreads();
    uint64_t ps = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) ps += prod64(tv[j], xsh[16 * w + j]);
    part[w][lane] = ps;
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int q = 0; q < kTileThreads / 64; ++q) t -= part[q][lane];
    }
  }
  if (w == 0) {
    // Diagonal sweep: at step c every lane finishes its row, lane c's value
    // is broadcast and subtracted from the later rows.  (Chaining the row
    // finish on the scalar unit instead, the next row's accumulator read
    // from its lane, measured 45-55 % slower: the dependent scalar
    // multiplies and the lane reads cost more than the vector finish.)
    int32_t xv = 0;
#pragma unroll
    for (int st = 0; st < kTB; ++st) {
      const int c = fwd ? st : kTB - 1 - st;
      if (c < rows) {
        const int32_t mine = acc_final<int32_t, K>(t, dv, bit, dbit);
        if (lane == c) xv = mine;
        const int32_t xc = __builtin_amdgcn_readlane(mine, c);
        t -= prod64(reg[c], xc);
      }
    }
    if (lane < rows) {
      out[r0 + lane] = xv;
      const unsigned long long word =
          (static_cast<unsigned long long>(static_cast<unsigned>(epoch)) << 32) |
          static_cast<uint32_t>(xv);
      __hip_atomic_store(&xp[r0 + lane], word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  tile_done(tickets, nb);
}

// Tiled FP coarse solve (solvers 0 and 1), one wave per 64-row block, with
// each row's terms subtracted in the sequential solve's order: the
// dependency blocks in processing order (forward 0, 1, ...; backward nb-1,
// nb-2, ...) and inside a tile the columns in the same direction, so a row
// receives its terms in ascending column order (L) or descending (DU),
// entries absent from the factor (clear bits of the row's 64-bit mask per
// tile) skipped, exactly as the sequential CSR solve does.  Lane r owns row
// 64 b + r.  A finished block publishes its values in `out`, then raises its
// flag to the call's epoch (agent-scope release); a consumer waits for the
// flag (acquire) and reads the block's values.  Blocks are taken from a
// ticket counter in processing order, so every wait is on a workgroup that
// started earlier; a launch out of its wait budget (tile_wait_over) records
// an error and gives up instead of hanging.
template <typename T>
__device__ __forceinline__ T readlane_fp(T v, int src);
template <>
__device__ __forceinline__ double readlane_fp<double>(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned long long>(b)), src);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<unsigned long long>(b) >> 32), src);
  return __longlong_as_double(static_cast<long long>(
      (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) | static_cast<unsigned>(lo)));
}
template <>
__device__ __forceinline__ float readlane_fp<float>(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

template <typename T, TriKind K>
__global__ __launch_bounds__(64) void k_tri_tiles_fp(
    int n, int nb, const T* __restrict__ tiles, const long long* __restrict__ toff,
    const uint64_t* __restrict__ tmask, const T* __restrict__ dinv_g,
    const T* __restrict__ in, T* out, unsigned long long* flags, int epoch, int* err,
    unsigned* tickets) {
  constexpr bool fwd = (K != TriKind::kDU);
  __shared__ T dl[kTB][kTB];  // the diagonal tile, [column][row]: lane r reads only its row
  __shared__ int ticket_sh;
  const long long t0 = wall_clock64();
  const int pb = tile_ticket(tickets, &ticket_sh);
  const int b = fwd ? pb : nb - 1 - pb;
  const int lane = threadIdx.x;
  const int r0 = b * kTB;
  const int rows = min(kTB, n - r0);
  const T* Tb = tiles + toff[b];
  const uint64_t* Mb = tmask + toff[b] / kTB;  // one mask per tile row
  const int ndep = fwd ? b : nb - 1 - b;
  // the diagonal tile, its masks and the row scales depend on no other
  // block: loaded before the first wait, so their latency stays off the chain
  const T* Td = Tb + static_cast<size_t>(ndep) * (kTB * kTB);
#pragma unroll
  for (int c = 0; c < kTB; ++c) dl[c][lane] = Td[c * kTB + lane];
  const uint64_t md = Mb[static_cast<size_t>(ndep) * kTB + lane];
  const T dv = (K != TriKind::kL && lane < rows) ? dinv_g[r0 + lane] : T(0);
  T t = lane < rows ? in[r0 + lane] : T(0);
  const unsigned long long want = static_cast<unsigned long long>(static_cast<unsigned>(epoch));
  for (int k = 0; k < ndep; ++k) {
    const int bp = fwd ? k : nb - 1 - k;
    // the dependency tile is loaded before the wait for its block
    const T* Tk = Tb + static_cast<size_t>(k) * (kTB * kTB);
    T vk[kTB];
#pragma unroll
    for (int c = 0; c < kTB; ++c) vk[c] = Tk[c * kTB + lane];
    const uint64_t mk = Mb[static_cast<size_t>(k) * kTB + lane];
    int spins = 0;
    while (__hip_atomic_load(&flags[bp], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != want) {
      __builtin_amdgcn_s_sleep(1);
      if ((++spins & 31) == 0 && tile_wait_over(t0, err)) {
        if (lane == 0) atomicAdd(err, 1);
        break;
      }
    }
    const int cmax = min(kTB, n - bp * kTB);
    const T xmine = lane < cmax ? out[bp * kTB + lane] : T(0);
#pragma unroll
    for (int st = 0; st < kTB; ++st) {
      const int c = fwd ? st : kTB - 1 - st;
      const T xc = readlane_fp<T>(xmine, c);
      if ((mk >> c) & 1) t = t - vk[c] * xc;
    }
  }
  T xv = T(0);
#pragma unroll
  for (int st = 0; st < kTB; ++st) {
    const int c = fwd ? st : kTB - 1 - st;
    if (c < rows) {
      const T mine = acc_final<T, K>(t, dv, 0, 0);
      if (lane == c) xv = mine;
      const T xc = readlane_fp<T>(mine, c);
      if ((md >> c) & 1) t = t - dl[c][lane] * xc;
    }
  }
  if (lane < rows) out[r0 + lane] = xv;
  __threadfence();
  if (lane == 0) __hip_atomic_store(&flags[b], want, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  tile_done(tickets, nb);
}

template <typename T, TriKind K>
hipError_t tri_blocked(const DCsr& a, const int* bptr, int nb, const void* in, void* out,
                       hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  if constexpr (std::is_same_v<T, int32_t>) {
    const size_t lds = tri_blocked_lds(a.n, 4, 8);
    static bool attr_i = false;
    if (!attr_i) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tri_blocked_int<K>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         static_cast<int>(kTriBlockedMaxLds));
      if (e != hipSuccess) return e;
      attr_i = true;
    }
    hipLaunchKernelGGL((k_tri_blocked_int<K>), dim3(1), dim3(kBlkThreads), lds, s, a.n,
                       a.rptr, a.cind, static_cast<const int32_t*>(a.val), a.bit, a.dbit,
                       bptr, nb, static_cast<const int32_t*>(in), static_cast<int32_t*>(out));
    return hipGetLastError();
  }
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

// The first kPre dependency entries of position idx are also stored with a
// fixed stride (ecol / eval[idx * kPre + q]), so every load of a row's data
// depends only on idx; entries beyond kPre are read from lcol / lval.
template <typename T>
struct PreRow {
  int row = -1, beg = 0, end = 0;
  T in = 0, dinv = 0;
  int col[kPre];
  T val[kPre];
};

template <typename T, TriKind K>
__device__ __forceinline__ void pre_meta(PreRow<T>& r, int idx, int lim,
                                         const int* __restrict__ lrow,
                                         const int* __restrict__ lbeg,
                                         const int* __restrict__ ecol,
                                         const T* __restrict__ eval,
                                         const T* __restrict__ ldiag) {
  if (idx >= lim) {
    r.row = -1;
    return;
  }
  r.row = lrow[idx];
  r.beg = lbeg[idx];
  r.end = lbeg[idx + 1];
  if (K != TriKind::kL) r.dinv = ldiag[idx];
#pragma unroll
  for (int q = 0; q < kPre; ++q) {
    r.col[q] = ecol[static_cast<size_t>(idx) * kPre + q];
    r.val[q] = eval[static_cast<size_t>(idx) * kPre + q];
  }
}

template <typename T>
__device__ __forceinline__ void pre_in(PreRow<T>& r, const T* __restrict__ in) {
  if (r.row >= 0) r.in = in[r.row];
}

template <typename T, TriKind K, bool kRing>
__device__ __forceinline__ T pre_eval(const PreRow<T>& r, int bit, int dbit,
                                      const int* __restrict__ lcol,
                                      const T* __restrict__ lval, const T* out,
                                      const T* ring, int rmask) {
  using Acc = typename AccOf<T>::type;
  Acc t = acc_init<T>(r.in, bit);
  // kRing: entries name dependency positions, whose values sit in the LDS ring
  auto X = [&](int c) -> T { return kRing ? ring[c & rmask] : out[c]; };
  const int cnt = r.end - r.beg;
#pragma unroll
  for (int q = 0; q < kPre; ++q)
    if (q < cnt) t = acc_sub<T>(t, r.val[q], X(r.col[q]));
  for (int p = r.beg + kPre; p < r.end; ++p) t = acc_sub<T>(t, lval[p], X(lcol[p]));
  return acc_final<T, K>(t, r.dinv, bit, dbit);
}


// kRing: dependency entries hold level-ordered positions; the values of the
// last rmask+1 positions of the chain live in an LDS ring (the host checks
// that no dependency reaches further back than that).  Otherwise entries hold
// row indices and values are read back from `out` after a full barrier.
template <typename T, TriKind K, bool kRing>
__global__ __launch_bounds__(kChainBlock) void k_tri_chain_pre(
    const int* __restrict__ lrow, const int* __restrict__ lbeg,
    const int* __restrict__ lcol, const T* __restrict__ lval,
    const int* __restrict__ ecol, const T* __restrict__ eval,
    const T* __restrict__ ldiag, int bit, int dbit, const int* __restrict__ lev_ptr,
    const int* __restrict__ chain_lev, int rmask, const T* __restrict__ in, T* out) {
  extern __shared__ __align__(16) unsigned char smem_c[];
  T* ring = reinterpret_cast<T*>(smem_c);
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int tid = threadIdx.x;
  // pipeline: nxt = level L+1 (metadata and input), far = level L+2 (metadata)
  PreRow<T> nxt, far;
  if (lb < le) {
    pre_meta<T, K>(nxt, lev_ptr[lb] + tid, lev_ptr[lb + 1], lrow, lbeg, ecol, eval, ldiag);
    pre_in<T>(nxt, in);
  }
  if (lb + 1 < le)
    pre_meta<T, K>(far, lev_ptr[lb + 1] + tid, lev_ptr[lb + 2], lrow, lbeg, ecol, eval,
                   ldiag);
  for (int L = lb; L < le; ++L) {
    const PreRow<T> cur = nxt;
    const int s = lev_ptr[L], e = lev_ptr[L + 1];
    if (L + 1 < le) {
      nxt = far;
      pre_in<T>(nxt, in);
    }
    if (L + 2 < le)
      pre_meta<T, K>(far, lev_ptr[L + 2] + tid, lev_ptr[L + 3], lrow, lbeg, ecol, eval,
                     ldiag);
    if (cur.row >= 0) {
      const T v = pre_eval<T, K, kRing>(cur, bit, dbit, lcol, lval, out, ring, rmask);
      out[cur.row] = v;
      if (kRing) ring[(s + tid) & rmask] = v;
    }
    // rows beyond one per thread (large levels)
    for (int idx = s + kChainBlock + tid; idx < e; idx += kChainBlock) {
      PreRow<T> r;
      pre_meta<T, K>(r, idx, e, lrow, lbeg, ecol, eval, ldiag);
      pre_in<T>(r, in);
      const T v = pre_eval<T, K, kRing>(r, bit, dbit, lcol, lval, out, ring, rmask);
      out[r.row] = v;
      if (kRing) ring[idx & rmask] = v;
    }
    if (kRing)
      lds_barrier();
    else
      __syncthreads();
  }
}

// ---- entry-parallel chain solve (integer factors, LDS ring) -------------------
//
// For integer factors a row's sum is order-free, so the dependency entries of
// a whole level (contiguous in the level-ordered layout) are spread over the
// lanes: each lane multiplies its entries by solution values and adds the
// products into the row's LDS accumulator with 64-bit LDS atomics; then one
// lane per row finishes the row.  The factor is numbered in level order (row
// = position), so a level's right-hand sides, inverse diagonals and results
// are contiguous.
//
// Entries are packed as int4 {dep, val, slot, 0}: dep >= 0 is the position of
// a dependency whose value sits in the LDS ring; dep < 0 (kFar kernels only)
// names position -dep-1, a dependency at least kEpFarLevels levels back that
// the ring cannot hold, read from `out` in global memory.  A full workgroup
// fence every kEpFence levels makes such values visible before any load of
// them is issued.  The chain's level offsets are staged in LDS, so the loads
// of a level depend only on LDS data: entries, right-hand sides and inverse
// diagonals of level L + kD are loaded once level L is done (into the
// registers it has just freed), far values of level L + kEpFarAhead (whose
// addresses need the entries) at the start of level L.  Every load is issued
// unconditionally (lanes past a level's end read element 0 and ignore it):
// loads under divergent branches make the compiler wait for all outstanding
// loads at the join.

// Prefetch distance kD in levels: 4 for the wide register layout (8 entries,
// 2 rows per lane), 8 for the narrow one (4 entries, 1 row), whose buffers
// are half the size.  A level costs at least the global-load latency / kD.
constexpr int kEpFarAhead = 2;  // far-value distance in levels
constexpr int kEpFence = 32;    // full fence period in levels (kFar kernels)
static_assert(kEpFarAhead < 4, "far values need their entries loaded first");
static_assert(kEpFarLevels >= kEpFarAhead + kEpFence,
              "far dependencies must be fenced before their values are loaded");

template <int kE, int kR>
struct EpPre {
  int2 lo, hi;  // the level's {position, entry} bounds
  int dep[kE];
  int32_t val[kE];
  int slot[kE];
  int32_t xfar[kE];
  int32_t dinv[kR];
  int32_t in[kR];
};

template <int kThreads, int kE, int kR>
__device__ __forceinline__ void ep_load(EpPre<kE, kR>& r, int2 lo, int2 hi,
                                        const int4* __restrict__ lent,
                                        const int32_t* __restrict__ dinv,
                                        const int32_t* __restrict__ in) {
  const int t = threadIdx.x;
  r.lo = lo;
  r.hi = hi;
#pragma unroll
  for (int q = 0; q < kE; ++q) {
    const int p = lo.y + t + q * kThreads;
    const int4 w = lent[p < hi.y ? p : 0];
    r.dep[q] = w.x;
    r.val[q] = w.y;
    r.slot[q] = w.z;
  }
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int idx = lo.x + t + q * kThreads;
    const int id = idx < hi.x ? idx : 0;
    r.dinv[q] = dinv[id];
    r.in[q] = in[id];
  }
}

template <int kThreads, int kE, int kR>
__device__ __forceinline__ void ep_far(EpPre<kE, kR>& r, const int32_t* out) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kE; ++q) {
    const int p = r.lo.y + t + q * kThreads;
    const int d = r.dep[q];
    r.xfar[q] = out[(p < r.hi.y && d < 0) ? -d - 1 : 0];
  }
}


struct EpShared {
  const int2* lvo;  // level offsets {position, entry}, LDS
  uint64_t* acc;    // accumulators, then kEpDummy slots for lanes past a level's end
  int32_t* ring;
  int rmask;
  int dummy;        // index of the first dummy slot
};

// One level of the chain; `cur` holds the level's prefetched data, `ahead`
// the buffer of level L + kEpFarAhead.  Level indices past the chain's end
// are clamped to an empty level.  kC accumulator copies per row: lane t adds
// into copy t % kC of its entry's row; the row's thread sums the copies.
template <TriKind K, int kWaves, bool kFar, int kE, int kR, int kC, int kD>
__device__ __forceinline__ void ep_level(EpPre<kE, kR>& cur, EpPre<kE, kR>& ahead, int L,
                                         int nl, const EpShared& sh,
                                         const int4* __restrict__ lent,
                                         const int32_t* __restrict__ dinv,
                                         const int32_t* __restrict__ in, int32_t* out, int bit,
                                         int dbit) {
  constexpr int kThreads = 64 * kWaves;
  const int t = threadIdx.x;
  const int s = cur.lo.x, e = cur.hi.x, e0 = cur.lo.y, e1 = cur.hi.y;
  uint64_t* acc = sh.acc;
  int32_t* ring = sh.ring;
  const int rmask = sh.rmask;
  const int cp = t & (kC - 1);
  // Offsets of the level prefetched at the end of this one, read now so
  // their LDS latency overlaps this level's work.
  const int Lp = min(L + kD, nl);
  const int2 plo = sh.lvo[Lp], phi = sh.lvo[min(Lp + 1, nl)];
  if (kFar) ep_far<kThreads>(ahead, out);
  // All ring reads first (lanes past the level's end read a valid, unused
  // slot), then the atomics: one LDS round trip for the whole level.
  int32_t xv[kE];
#pragma unroll
  for (int q = 0; q < kE; ++q) xv[q] = ring[cur.dep[q] & rmask];
  if (kFar) {
    // Far values are rare; waiting for their loads only where a lane of
    // the wave has one keeps the other levels off the global-load latency.
#pragma unroll
    for (int q = 0; q < kE; ++q)
      if (__any(cur.dep[q] < 0)) xv[q] = cur.dep[q] < 0 ? cur.xfar[q] : xv[q];
  }
  // Unconditional atomics: lanes past the level's end add into a dummy slot
  // of their own, which keeps the LDS traffic free of branches and waits.
#pragma unroll
  for (int q = 0; q < kE; ++q) {
    const int p = e0 + t + q * kThreads;
    const int a = p < e1 ? cur.slot[q] * kC + cp : sh.dummy + (t & (kEpDummy - 1));
    atomicAdd(reinterpret_cast<unsigned long long*>(&acc[a]),
              static_cast<unsigned long long>(prod64(cur.val[q], xv[q])));
  }
  for (int p = e0 + t + kE * kThreads; p < e1; p += kThreads) {
    const int4 w = lent[p];
    const int32_t x = (kFar && w.x < 0) ? out[-w.x - 1] : ring[w.x & rmask];
    atomicAdd(reinterpret_cast<unsigned long long*>(&acc[w.z * kC + cp]),
              static_cast<unsigned long long>(prod64(w.y, x)));
  }
  ep_sync<kWaves>();
  uint64_t av[kR];
  if (kC == 1) {
#pragma unroll
    for (int q = 0; q < kR; ++q) av[q] = acc[t + q * kThreads];
  } else {
#pragma unroll
    for (int q = 0; q < kR; ++q) {
      const int k = t + q * kThreads;
      uint64_t sum = 0;
      if (s + k < e) {
#pragma unroll
        for (int c = 0; c < kC; ++c) {
          sum += acc[k * kC + c];
          acc[k * kC + c] = 0;
        }
      }
      av[q] = sum;
    }
  }
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int k = t + q * kThreads;
    if (s + k < e) {
      const uint64_t tt =
          (static_cast<uint64_t>(static_cast<int64_t>(cur.in[q])) << bit) - av[q];
      if (kC == 1) acc[k] = 0;
      const int32_t v = acc_final<int32_t, K>(tt, cur.dinv[q], bit, dbit);
      ring[(s + k) & rmask] = v;
      out[s + k] = v;
    }
  }
  for (int k = t + kR * kThreads; s + k < e; k += kThreads) {
    uint64_t sum = 0;
    for (int c = 0; c < kC; ++c) {
      sum += acc[k * kC + c];
      acc[k * kC + c] = 0;
    }
    const uint64_t tt = (static_cast<uint64_t>(static_cast<int64_t>(in[s + k])) << bit) - sum;
    const int32_t v = acc_final<int32_t, K>(tt, dinv[s + k], bit, dbit);
    ring[(s + k) & rmask] = v;
    out[s + k] = v;
  }
  if (kFar && (L % kEpFence) == kEpFence - 1)
    __syncthreads();  // full fence: global results visible to later loads
  else
    ep_sync<kWaves>();
  ep_load<kThreads>(cur, plo, phi, lent, dinv, in);
}

// LDS: acc_slots uint64 accumulators (kC per row of the widest level), the
// ring, then the chain's level offsets (max_levels + 1 int2).
template <TriKind K, int kWaves, bool kFar, int kE, int kR, int kC, int kD>
__global__ __launch_bounds__(64 * kWaves) void k_tri_chain_ep(
    const int4* __restrict__ lent, const int32_t* __restrict__ dinv,
    const int* __restrict__ lev_ptr, const int* __restrict__ lev_eptr,
    const int* __restrict__ chain_lev, int rmask, int acc_slots, int bit, int dbit,
    const int32_t* __restrict__ in, int32_t* out) {
  constexpr int kThreads = 64 * kWaves;
  extern __shared__ __align__(16) unsigned char smem_e[];
  EpShared sh;
  sh.acc = reinterpret_cast<uint64_t*>(smem_e);
  sh.dummy = acc_slots;
  sh.ring = reinterpret_cast<int32_t*>(sh.acc + acc_slots + kEpDummy);
  sh.rmask = rmask;
  int2* lvo = reinterpret_cast<int2*>(sh.ring + rmask + 1);
  sh.lvo = lvo;
  const int t = threadIdx.x;
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int nl = le - lb;
  for (int k = t; k < acc_slots; k += kThreads) sh.acc[k] = 0;
  for (int k = t; k <= nl; k += kThreads) lvo[k] = make_int2(lev_ptr[lb + k], lev_eptr[lb + k]);
  __syncthreads();
  EpPre<kE, kR> buf[kD];
#pragma unroll
  for (int d = 0; d < kD; ++d) {
    const int Lc = min(d, nl);
    ep_load<kThreads>(buf[d], lvo[Lc], lvo[min(Lc + 1, nl)], lent, dinv, in);
  }
  // Dependencies stay inside a chain, so its first kEpFarLevels levels have
  // no far entries: the first far loads are issued from the loop.
  int L0 = 0;
  for (; L0 + kD <= nl; L0 += kD) {
#pragma unroll
    for (int d = 0; d < kD; ++d)
      ep_level<K, kWaves, kFar, kE, kR, kC, kD>(buf[d], buf[(d + kEpFarAhead) % kD], L0 + d, nl,
                                                sh, lent, dinv, in, out, bit, dbit);
  }
#pragma unroll
  for (int d = 0; d < kD; ++d)
    if (L0 + d < nl)
      ep_level<K, kWaves, kFar, kE, kR, kC, kD>(buf[d], buf[(d + kEpFarAhead) % kD], L0 + d, nl,
                                                sh, lent, dinv, in, out, bit, dbit);
}

template <TriKind K, int kWaves, bool kFar, int kE, int kR, int kC, int kD>
hipError_t tri_ep_go(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                     const int32_t* in, int32_t* out, hipStream_t s) {
  const size_t lds = ep_lds_bytes(L.ring, L.ep_maxlev, L.ep_acc);
  if (lds > static_cast<size_t>(kTriBlockedMaxLds) || L.ep_acc < kC) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_tri_chain_ep<K, kWaves, kFar, kE, kR, kC, kD>),
        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kTriBlockedMaxLds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  launch_timed(&k_tri_chain_ep<K, kWaves, kFar, kE, kR, kC, kD>, dim3(seg.nchain),
               dim3(64 * kWaves), static_cast<uint32_t>(lds), s,
               reinterpret_cast<const int4*>(L.lent), static_cast<const int32_t*>(L.ldiag),
               L.lev_ptr, L.lev_eptr, seg.chain_lev, L.ring - 1, L.ep_acc, bit, dbit, in, out);
  return hipGetLastError();
}

template <TriKind K, int kWaves, int kC>
hipError_t tri_ep_launch_c(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                           const int32_t* in, int32_t* out, hipStream_t s) {
  if (L.ep_far) return tri_ep_go<K, kWaves, true, 8, 2, kC, 4>(L, seg, bit, dbit, in, out, s);
  if (L.ep_narrow)
    return tri_ep_go<K, kWaves, false, 4, 1, kC, 8>(L, seg, bit, dbit, in, out, s);
  return tri_ep_go<K, kWaves, false, 8, 2, kC, 4>(L, seg, bit, dbit, in, out, s);
}

// Accumulator copies only for the smoother's LD factors (rows with many
// entries on coarse levels).
template <TriKind K, int kWaves>
hipError_t tri_ep_launch(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                         const int32_t* in, int32_t* out, hipStream_t s) {
  if constexpr (K == TriKind::kLD) {
    if (L.ep_copies == kEpCopies)
      return tri_ep_launch_c<K, kWaves, kEpCopies>(L, seg, bit, dbit, in, out, s);
  }
  return tri_ep_launch_c<K, kWaves, 1>(L, seg, bit, dbit, in, out, s);
}

// ---------------------------------------------------------------------------
// Row-parallel integer LD solve: the chains, levels, LDS ring and far
// dependencies of k_tri_chain_ep, for factors whose rows hold at most kG * kE
// dependency entries.  kG lanes serve one row (kG = 1: each lane serves kR
// rows t, t + kThreads, ...); each lane keeps kE entries of its row, padded
// with zero values, in registers.  A level costs one round of ring reads, the
// products in registers, a sum over the row's kG lanes (DPP moves),
// the ring writes and one barrier: no accumulators, no atomics.  The integer
// sum is order-free (wrap-around), so the result equals the sequential one
// bit for bit.  Entries are stored by position as kG * kE int2 {dep, val}
// pairs, lane j's kE contiguous (a row's entry m goes to lane m % kG); dep is
// encoded as in k_tri_chain_ep and padding is {0, 0}.
template <int kE, int kR>
struct RpPre {
  int lo, hi;             // the level's position bounds
  int4 ent[kR][kE / 2];   // {dep, val, dep, val}
  int32_t xfar[kR][kE];
  int32_t dinv[kR];
  int32_t in[kR];
};

template <int kThreads, int kG, int kE, int kR>
__device__ __forceinline__ void rp_load(RpPre<kE, kR>& r, int lo, int hi,
                                        const int4* __restrict__ rent,
                                        const int32_t* __restrict__ dinv,
                                        const int32_t* __restrict__ in) {
  const int t = threadIdx.x;
  r.lo = lo;
  r.hi = hi;
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int idx = lo + t / kG + q * (kThreads / kG);
    const int id = idx < hi ? idx : 0;  // lanes past the level's end load position 0
    const size_t base = (static_cast<size_t>(id) * kG + t % kG) * (kE / 2);
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) r.ent[q][h] = rent[base + h];
    r.dinv[q] = dinv[id];
    r.in[q] = in[id];
  }
}

// Far values of the first kFq entries of each row (the plan puts a row's far
// entries first; kFq bounds their number).
template <int kThreads, int kG, int kE, int kR, int kFq = kE>
__device__ __forceinline__ void rp_far(RpPre<kE, kR>& r, const int32_t* out) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const bool ok = r.lo + t / kG + q * (kThreads / kG) < r.hi;
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) {
      const int d0 = r.ent[q][h].x, d1 = r.ent[q][h].z;
      if (2 * h < kFq) r.xfar[q][2 * h] = out[(ok && d0 < 0) ? -d0 - 1 : 0];
      if (2 * h + 1 < kFq) r.xfar[q][2 * h + 1] = out[(ok && d1 < 0) ? -d1 - 1 : 0];
    }
  }
}


template <TriKind K, int kWaves, bool kFar, int kG, int kE, int kR, int kD, int kA, int kFq>
__device__ __forceinline__ void rp_level(RpPre<kE, kR>& cur, RpPre<kE, kR>& ahead, int L, int nl,
                                         const int* lvo, int32_t* ring, int rmask,
                                         const int4* __restrict__ rent,
                                         const int32_t* __restrict__ dinv,
                                         const int32_t* __restrict__ in, int32_t* out, int bit,
                                         int dbit) {
  constexpr int kThreads = 64 * kWaves;
  const int t = threadIdx.x;
  const int s = cur.lo, e = cur.hi;
  // bounds of the level prefetched at the end of this one
  const int Lp = min(L + kD, nl);
  const int plo = lvo[Lp], phi = lvo[min(Lp + 1, nl)];
  if (kFar) rp_far<kThreads, kG, kE, kR, kFq>(ahead, out);
  int32_t xv[kR][kE];
#pragma unroll
  for (int q = 0; q < kR; ++q)
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) {
      xv[q][2 * h] = ring[cur.ent[q][h].x & rmask];
      xv[q][2 * h + 1] = ring[cur.ent[q][h].z & rmask];
    }
  // rows of 8-12 entries: every ring read is issued before the first
  // product, so the level pays one LDS round trip (left to itself the
  // scheduler splits them in two: 4-5 % slower; with 4 entries the barrier
  // measured 2 % slower)
  if constexpr (kE >= 8) __builtin_amdgcn_sched_barrier(0);
  if (kFar) {
#pragma unroll
    for (int q = 0; q < kR; ++q)
#pragma unroll
      for (int h = 0; h < kE / 2; ++h) {
        const int d0 = cur.ent[q][h].x, d1 = cur.ent[q][h].z;
        if (2 * h < kFq && __any(d0 < 0))
          xv[q][2 * h] = d0 < 0 ? cur.xfar[q][2 * h] : xv[q][2 * h];
        if (2 * h + 1 < kFq && __any(d1 < 0))
          xv[q][2 * h + 1] = d1 < 0 ? cur.xfar[q][2 * h + 1] : xv[q][2 * h + 1];
      }
  }
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    uint64_t acc = 0, acc2 = 0;  // two chains of multiply-adds
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) {
      acc += prod64(cur.ent[q][h].y, xv[q][2 * h]);
      acc2 += prod64(cur.ent[q][h].w, xv[q][2 * h + 1]);
    }
    // opaque ends keep the two chains apart (the compiler would otherwise
    // fold them into one chain of dependent multiply-adds)
    asm volatile("" : "+v"(acc));
    asm volatile("" : "+v"(acc2));
    acc = group_sum<kG>(acc + acc2);
    const int k = s + t / kG + q * (kThreads / kG);
    if (k < e && t % kG == 0) {
      const uint64_t tt = (static_cast<uint64_t>(static_cast<int64_t>(cur.in[q])) << bit) - acc;
      const int32_t v = acc_final<int32_t, K>(tt, cur.dinv[q], bit, dbit);
      ring[k & rmask] = v;
      out[k] = v;
    }
  }
  if (kFar && (L % kEpFence) == kEpFence - 1)
    __syncthreads();  // full fence: global results visible to later loads
  else
    ep_sync<kWaves>();
  rp_load<kThreads, kG>(cur, plo, phi, rent, dinv, in);
}

// LDS: the ring, then the chain's level offsets (max_levels + 1 ints).
// Far values of level L + kA are loaded at the start of level L, from entries
// loaded kD - kA levels before that.
template <TriKind K, int kWaves, bool kFar, int kG, int kE, int kR, int kD, int kA, int kFq = kE>
__global__ __launch_bounds__(64 * kWaves) void k_tri_chain_rp(
    const int4* __restrict__ rent, const int32_t* __restrict__ dinv,
    const int* __restrict__ lev_ptr, const int* __restrict__ chain_lev, int rmask, int bit,
    int dbit, const int32_t* __restrict__ in, int32_t* out) {
  constexpr int kThreads = 64 * kWaves;
  static_assert(kG == 1 || kR == 1, "several lanes per row serve one row each");
  static_assert(kA >= 1 && kA < kD, "far values need their entries loaded first");
  static_assert(kEpFarLevels >= kA + kEpFence,
                "far dependencies must be fenced before their values are loaded");
  extern __shared__ __align__(16) unsigned char smem_r[];
  int32_t* ring = reinterpret_cast<int32_t*>(smem_r);
  int* lvo = ring + rmask + 1;
  const int t = threadIdx.x;
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int nl = le - lb;
  for (int k = t; k <= nl; k += kThreads) lvo[k] = lev_ptr[lb + k];
  __syncthreads();
  RpPre<kE, kR> buf[kD];
#pragma unroll
  for (int d = 0; d < kD; ++d) {
    const int Lc = min(d, nl);
    rp_load<kThreads, kG>(buf[d], lvo[Lc], lvo[min(Lc + 1, nl)], rent, dinv, in);
  }
  // Whole rounds of kD levels, then the tail in a copy of the body.  (One
  // body for all levels measured slower: with a test around each level the
  // compiler waits for every outstanding load where the paths join, and
  // without one the last round runs up to kD - 1 empty levels.)
  int L0 = 0;
  for (; L0 + kD <= nl; L0 += kD) {
#pragma unroll
    for (int d = 0; d < kD; ++d)
      rp_level<K, kWaves, kFar, kG, kE, kR, kD, kA, kFq>(buf[d], buf[(d + kA) % kD], L0 + d, nl,
                                                         lvo, ring, rmask, rent, dinv, in, out,
                                                         bit, dbit);
  }
#pragma unroll
  for (int d = 0; d < kD; ++d)
    if (L0 + d < nl)
      rp_level<K, kWaves, kFar, kG, kE, kR, kD, kA, kFq>(buf[d], buf[(d + kA) % kD], L0 + d, nl,
                                                         lvo, ring, rmask, rent, dinv, in, out,
                                                         bit, dbit);
}

template <int kWaves, bool kFar, int kG, int kE, int kR, int kD, int kA, int kFq = kE>
hipError_t tri_rp_go(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                     const int32_t* in, int32_t* out, hipStream_t s) {
  constexpr TriKind K = TriKind::kLD;
  const size_t lds = rp_lds_bytes(L.ring, L.ep_maxlev);
  if (lds > static_cast<size_t>(kTriBlockedMaxLds)) return hipErrorInvalidValue;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_tri_chain_rp<K, kWaves, kFar, kG, kE, kR, kD, kA, kFq>),
        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kTriBlockedMaxLds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  launch_timed(&k_tri_chain_rp<K, kWaves, kFar, kG, kE, kR, kD, kA, kFq>, dim3(seg.nchain),
               dim3(64 * kWaves), static_cast<uint32_t>(lds), s,
               reinterpret_cast<const int4*>(L.rent), static_cast<const int32_t*>(L.ldiag),
               L.lev_ptr, seg.chain_lev, L.ring - 1, bit, dbit, in, out);
  return hipGetLastError();
}

// Register layouts.  4 entries and 1 row per lane: prefetch distance 8
// levels and far values 4 levels ahead (a level takes less than one
// global-load latency, and a far value waits for two: at 128^3 the finest
// level's solve took 116 us, against 157 us at distance 4); 2 rows per lane:
// distance 6, far values 3 ahead (the registers of 8 levels spill).  One lane with 8,
// 10 or 12 entries, or 2, 4, 8 or 16 lanes of 8, 10 or 12 entries per row (factors
// without far dependencies): distance 4, far values 2 ahead (distance 8
// measured no faster there).
template <int kWaves, bool kFar>
hipError_t tri_rp_layout(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                         const int32_t* in, int32_t* out, hipStream_t s) {
  if (L.rp_lanes <= 1) {
    if (L.rp_e == 4 && L.rp_rows == 1) {
      if (kFar && L.rp_farq == 1)  // one far entry at most, stored first
        return tri_rp_go<kWaves, kFar, 1, 4, 1, 8, 4, 1>(L, seg, bit, dbit, in, out, s);
      return tri_rp_go<kWaves, kFar, 1, 4, 1, 8, 4>(L, seg, bit, dbit, in, out, s);
    }
    if (L.rp_e == 4 && L.rp_rows == 2)
      return tri_rp_go<kWaves, kFar, 1, 4, 2, 6, 3>(L, seg, bit, dbit, in, out, s);
    if (L.rp_e == 8 && L.rp_rows == 1)
      return tri_rp_go<kWaves, kFar, 1, 8, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
    if (L.rp_e == 10 && L.rp_rows == 1)
      return tri_rp_go<kWaves, kFar, 1, 10, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
    if (L.rp_e == 12 && L.rp_rows == 1)
      return tri_rp_go<kWaves, kFar, 1, 12, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
    return hipErrorInvalidValue;
  }
  if constexpr (!kFar) {
    if (L.rp_e == 10 && L.rp_rows == 1) {
      if (L.rp_lanes == 2) return tri_rp_go<kWaves, false, 2, 10, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 4) return tri_rp_go<kWaves, false, 4, 10, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 8) return tri_rp_go<kWaves, false, 8, 10, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 16)
        return tri_rp_go<kWaves, false, 16, 10, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
    }
    if (L.rp_e == 12 && L.rp_rows == 1) {
      if (L.rp_lanes == 2) return tri_rp_go<kWaves, false, 2, 12, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 4) return tri_rp_go<kWaves, false, 4, 12, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 8) return tri_rp_go<kWaves, false, 8, 12, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 16)
        return tri_rp_go<kWaves, false, 16, 12, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
    }
    if (L.rp_e == 8 && L.rp_rows == 1) {
      if (L.rp_lanes == 2) return tri_rp_go<kWaves, false, 2, 8, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 4) return tri_rp_go<kWaves, false, 4, 8, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 8) return tri_rp_go<kWaves, false, 8, 8, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
      if (L.rp_lanes == 16)
        return tri_rp_go<kWaves, false, 16, 8, 1, 4, 2>(L, seg, bit, dbit, in, out, s);
    }
  }
  return hipErrorInvalidValue;
}

template <int kWaves>
hipError_t tri_rp_far(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                      const int32_t* in, int32_t* out, hipStream_t s) {
  if (L.ep_far) return tri_rp_layout<kWaves, true>(L, seg, bit, dbit, in, out, s);
  return tri_rp_layout<kWaves, false>(L, seg, bit, dbit, in, out, s);
}

// Row-parallel FP LD solve (solvers 0 and 1): the chains, levels, LDS ring and
// far rule of k_tri_chain_rp with one lane per row.  A row's products are
// subtracted in stored order, one at a time, as the sequential solve does
// (entries past the row's count are skipped, not added as zeros), so the
// result is the reference's bit for bit.  Entries are stored by position:
// kE dependency codes (as in k_tri_chain_ep) and kE values, and the count.
template <typename T, int kE, int kR>
struct RpfPre {
  int lo, hi;
  int dep[kR][kE];
  T val[kR][kE];
  T xfar[kR][kE];
  T dinv[kR], in[kR];
  int cnt[kR];
};

// kR rows per lane: lane t serves rows t, t + kThreads, ... of the level.
template <int kThreads, typename T, int kE, int kR>
__device__ __forceinline__ void rpf_load(RpfPre<T, kE, kR>& r, int lo, int hi,
                                         const int* __restrict__ rdep, const T* __restrict__ rval,
                                         const int* __restrict__ rcnt,
                                         const T* __restrict__ dinv,
                                         const T* __restrict__ in) {
  r.lo = lo;
  r.hi = hi;
#pragma unroll
  for (int u = 0; u < kR; ++u) {
    const int idx = lo + static_cast<int>(threadIdx.x) + u * kThreads;
    const int id = idx < hi ? idx : 0;  // lanes past the level's end load position 0
#pragma unroll
    for (int q = 0; q < kE; ++q) {
      r.dep[u][q] = rdep[static_cast<size_t>(id) * kE + q];
      r.val[u][q] = rval[static_cast<size_t>(id) * kE + q];
    }
    r.cnt[u] = rcnt[id];
    r.dinv[u] = dinv[id];
    r.in[u] = in[id];
  }
}

template <int kThreads, typename T, int kE, int kR>
__device__ __forceinline__ void rpf_far(RpfPre<T, kE, kR>& r, const T* out) {
#pragma unroll
  for (int u = 0; u < kR; ++u) {
    const bool ok = r.lo + static_cast<int>(threadIdx.x) + u * kThreads < r.hi;
#pragma unroll
    for (int q = 0; q < kE; ++q) {
      const int d = r.dep[u][q];
      r.xfar[u][q] = out[(ok && d < 0) ? -d - 1 : 0];
    }
  }
}

template <typename T, TriKind K, int kWaves, bool kFar, int kE, int kR, int kD, int kA>
__device__ __forceinline__ void rpf_level(RpfPre<T, kE, kR>& cur, RpfPre<T, kE, kR>& ahead,
                                          int L, int nl, const int* lvo, T* ring, int rmask,
                                          const int* __restrict__ rdep,
                                          const T* __restrict__ rval,
                                          const int* __restrict__ rcnt,
                                          const T* __restrict__ dinv,
                                          const T* __restrict__ in, T* out) {
  constexpr int kThreads = 64 * kWaves;
  const int Lp = min(L + kD, nl);
  const int plo = lvo[Lp], phi = lvo[min(Lp + 1, nl)];
  if (kFar) rpf_far<kThreads>(ahead, out);
  T xv[kR][kE];
#pragma unroll
  for (int u = 0; u < kR; ++u)
#pragma unroll
    for (int q = 0; q < kE; ++q) xv[u][q] = ring[cur.dep[u][q] & rmask];
  if (kFar) {
#pragma unroll
    for (int u = 0; u < kR; ++u)
#pragma unroll
      for (int q = 0; q < kE; ++q)
        if (__any(cur.dep[u][q] < 0)) xv[u][q] = cur.dep[u][q] < 0 ? cur.xfar[u][q] : xv[u][q];
  }
#pragma unroll
  for (int u = 0; u < kR; ++u) {
    T t = cur.in[u];
#pragma unroll
    for (int q = 0; q < kE; ++q) t = q < cur.cnt[u] ? acc_sub<T>(t, cur.val[u][q], xv[u][q]) : t;
    const int k = cur.lo + static_cast<int>(threadIdx.x) + u * kThreads;
    if (k < cur.hi) {
      const T v = acc_final<T, K>(t, cur.dinv[u], 0, 0);
      ring[k & rmask] = v;
      out[k] = v;
    }
  }
  if (kFar && (L % kEpFence) == kEpFence - 1)
    __syncthreads();
  else
    ep_sync<kWaves>();
  rpf_load<kThreads>(cur, plo, phi, rdep, rval, rcnt, dinv, in);
}

template <typename T, TriKind K, int kWaves, bool kFar, int kE, int kR, int kD, int kA>
__global__ __launch_bounds__(64 * kWaves) void k_tri_chain_rpf(
    const int* __restrict__ rdep, const T* __restrict__ rval, const int* __restrict__ rcnt,
    const T* __restrict__ dinv, const int* __restrict__ lev_ptr,
    const int* __restrict__ chain_lev, int rmask, const T* __restrict__ in, T* out) {
  constexpr int kThreads = 64 * kWaves;
  static_assert(kA >= 1 && kA < kD, "far values need their entries loaded first");
  static_assert(kEpFarLevels >= kA + kEpFence,
                "far dependencies must be fenced before their values are loaded");
  extern __shared__ __align__(16) unsigned char smem_f[];
  T* ring = reinterpret_cast<T*>(smem_f);
  int* lvo = reinterpret_cast<int*>(ring + rmask + 1);
  const int t = threadIdx.x;
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int nl = le - lb;
  // The ring starts zeroed: a skipped entry's value is then never a
  // leftover bit pattern.
  for (int k = t; k <= rmask; k += kThreads) ring[k] = T(0);
  for (int k = t; k <= nl; k += kThreads) lvo[k] = lev_ptr[lb + k];
  __syncthreads();
  RpfPre<T, kE, kR> buf[kD];
#pragma unroll
  for (int d = 0; d < kD; ++d) {
    const int Lc = min(d, nl);
    rpf_load<kThreads>(buf[d], lvo[Lc], lvo[min(Lc + 1, nl)], rdep, rval, rcnt, dinv, in);
  }
  int L0 = 0;
  for (; L0 + kD <= nl; L0 += kD) {
#pragma unroll
    for (int d = 0; d < kD; ++d)
      rpf_level<T, K, kWaves, kFar, kE, kR, kD, kA>(buf[d], buf[(d + kA) % kD], L0 + d, nl, lvo,
                                                    ring, rmask, rdep, rval, rcnt, dinv, in,
                                                    out);
  }
#pragma unroll
  for (int d = 0; d < kD; ++d)
    if (L0 + d < nl)
      rpf_level<T, K, kWaves, kFar, kE, kR, kD, kA>(buf[d], buf[(d + kA) % kD], L0 + d, nl, lvo,
                                                    ring, rmask, rdep, rval, rcnt, dinv, in,
                                                    out);
}

template <typename T, int kWaves, bool kFar, int kE, int kR>
hipError_t tri_rpf_go(const TriLayout& L, const TriSegment& seg, const T* in, T* out,
                      hipStream_t s) {
  constexpr TriKind K = TriKind::kLD;
  const size_t lds = sizeof(T) * static_cast<size_t>(L.ring) +
                     sizeof(int) * (static_cast<size_t>(L.ep_maxlev) + 1);
  if (lds > static_cast<size_t>(kTriBlockedMaxLds)) return hipErrorInvalidValue;
  // prefetch 4 levels ahead; 3 with two rows per lane, 2 for two FP64 rows
  // of 8 entries (whose registers spill otherwise); far values 2 ahead (1)
  constexpr int kD = kR == 1 ? 4 : (sizeof(T) == 8 && kE == 8) ? 2 : 3;
  constexpr int kA = kD > 2 ? 2 : 1;
  auto kern = &k_tri_chain_rpf<T, K, kWaves, kFar, kE, kR, kD, kA>;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(kTriBlockedMaxLds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(kern, dim3(seg.nchain), dim3(64 * kWaves), static_cast<uint32_t>(lds), s,
                     L.rpf_dep, static_cast<const T*>(L.rpf_val), L.rpf_cnt,
                     static_cast<const T*>(L.ldiag), L.lev_ptr, seg.chain_lev, L.ring - 1, in,
                     out);
  return hipGetLastError();
}

template <typename T, int kWaves, int kR>
hipError_t tri_rpf_layout(const TriLayout& L, const TriSegment& seg, const T* in, T* out,
                          hipStream_t s) {
  if (L.ep_far) {
    if (L.rp_e == 4) return tri_rpf_go<T, kWaves, true, 4, kR>(L, seg, in, out, s);
    return tri_rpf_go<T, kWaves, true, 8, kR>(L, seg, in, out, s);
  }
  if (L.rp_e == 4) return tri_rpf_go<T, kWaves, false, 4, kR>(L, seg, in, out, s);
  return tri_rpf_go<T, kWaves, false, 8, kR>(L, seg, in, out, s);
}

template <typename T, int kR>
hipError_t tri_rpf_rows(const TriLayout& L, const TriSegment& seg, const T* in, T* out,
                        hipStream_t s) {
  const int w = L.rp_waves;
  if (w <= 1) return tri_rpf_layout<T, 1, kR>(L, seg, in, out, s);
  if (w <= 2) return tri_rpf_layout<T, 2, kR>(L, seg, in, out, s);
  if (w <= 4) return tri_rpf_layout<T, 4, kR>(L, seg, in, out, s);
  return tri_rpf_layout<T, 8, kR>(L, seg, in, out, s);
}

template <typename T>
hipError_t tri_rpf_waves(const TriLayout& L, const TriSegment& seg, const T* in, T* out,
                         hipStream_t s) {
  // one or two rows per lane (levels of more than 512 rows)
  if (L.rp_rows == 2) return tri_rpf_rows<T, 2>(L, seg, in, out, s);
  return tri_rpf_rows<T, 1>(L, seg, in, out, s);
}

// FP row-parallel LD solve with several lanes per row (solvers 0 and 1;
// factors of up to kG * kE dependencies per row, no far entries): a row's
// entries are split into kG contiguous chunks of kE, lane g holding entries
// g kE .. g kE + kE - 1.  Every lane forms its products val * x in parallel;
// the row's first lane then takes them in stored order (lane by lane, slot
// by slot, through cross-lane moves) and subtracts them one at a time, as
// the sequential solve does, so the result is the reference's bit for bit.
// Chains, levels, LDS ring and prefetch as in k_tri_chain_rpf; waves per
// workgroup are a launch parameter.
template <typename T, int kG, int kE>
struct RpgPre {
  int lo, hi;
  int dep[kE];
  T val[kE];
  T xfar[kE];  // far values (far layouts only)
  T dinv, in;
  int cnt;
};

// How a row's lanes combine their products: shuffled to every lane of the
// row; staged in LDS and read back by the row's first lane; shuffled, with
// far dependencies (dep < 0: position -dep-1, at least kEpFarLevels levels
// back) read from `out` as in k_tri_chain_rpf.
constexpr int kRpgShuffle = 0, kRpgStage = 1, kRpgFar = 2;

template <typename T, int kG, int kE>
__device__ __forceinline__ void rpg_load(RpgPre<T, kG, kE>& r, int lo, int hi, int threads,
                                         const int* __restrict__ rdep, const T* __restrict__ rval,
                                         const int* __restrict__ rcnt,
                                         const T* __restrict__ dinv, const T* __restrict__ in) {
  const int t = threadIdx.x;
  const int idx = lo + t / kG;
  const int id = idx < hi ? idx : 0;  // lanes past the level's end load position 0
  r.lo = lo;
  r.hi = hi;
  const size_t base = (static_cast<size_t>(id) * kG + t % kG) * kE;
#pragma unroll
  for (int q = 0; q < kE; ++q) {
    r.dep[q] = rdep[base + q];
    r.val[q] = rval[base + q];
  }
  r.cnt = rcnt[id];
  r.dinv = dinv[id];
  r.in = in[id];
  (void)threads;
}

template <typename T, TriKind K, int kG, int kE, int kD, int kMode>
__device__ __forceinline__ void rpg_level(RpgPre<T, kG, kE>& cur, RpgPre<T, kG, kE>& ahead, int L,
                                          int nl, int threads, const int* lvo, T* ring, T* stage,
                                          int rmask, const int* __restrict__ rdep,
                                          const T* __restrict__ rval,
                                          const int* __restrict__ rcnt,
                                          const T* __restrict__ dinv,
                                          const T* __restrict__ in, T* out) {
  constexpr bool kFar = kMode == kRpgFar;
  const int Lp = min(L + kD, nl);
  const int plo = lvo[Lp], phi = lvo[min(Lp + 1, nl)];
  const int t = threadIdx.x;
  const int g = t % kG;
  if constexpr (kFar) {
    // far values of level L + kA (its entries are loaded already)
    const bool ok = ahead.lo + t / kG < ahead.hi;
#pragma unroll
    for (int q = 0; q < kE; ++q) {
      const int d = ahead.dep[q];
      ahead.xfar[q] = out[(ok && d < 0) ? -d - 1 : 0];
    }
  }
  if constexpr (kMode == kRpgStage) {
    // each lane forms the products of its chunk and stages them in the
    // row's LDS slots (entry m of the row at stage[row * kG * kE + m]); the
    // row's lanes are in one wave, whose LDS accesses stay in order, so its
    // first lane then reads them back and subtracts them in stored order
    T* st = stage + static_cast<size_t>(t / kG) * (kG * kE);
#pragma unroll
    for (int q = 0; q < kE; ++q) st[g * kE + q] = cur.val[q] * ring[cur.dep[q] & rmask];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    const int k = cur.lo + t / kG;
    if (k < cur.hi && g == 0) {
      T tt = cur.in;
#pragma unroll 16
      for (int m = 0; m < kG * kE; ++m)
        if (m < cur.cnt) tt = tt - st[m];
      const T v = acc_final<T, K>(tt, cur.dinv, 0, 0);
      ring[k & rmask] = v;
      out[k] = v;
    }
  } else {
    // no room for the staging area: the products reach every lane of the
    // row by shuffles and each subtracts them in stored order
    const int lane0 = (t & 63) - g;  // the row's first lane inside the wave
    T prod[kE];
#pragma unroll
    for (int q = 0; q < kE; ++q) {
      T x = ring[cur.dep[q] & rmask];
      if constexpr (kFar) x = cur.dep[q] < 0 ? cur.xfar[q] : x;
      prod[q] = cur.val[q] * x;
    }
    T tt = cur.in;
#pragma unroll
    for (int h = 0; h < kG; ++h)
#pragma unroll
      for (int q = 0; q < kE; ++q) {
        const T v = __shfl(prod[q], lane0 + h, 64);
        tt = h * kE + q < cur.cnt ? tt - v : tt;
      }
    const int k = cur.lo + t / kG;
    if (k < cur.hi && g == 0) {
      const T v = acc_final<T, K>(tt, cur.dinv, 0, 0);
      ring[k & rmask] = v;
      out[k] = v;
    }
  }
  // far layouts: a full workgroup fence every kEpFence levels makes the
  // values in `out` visible before any far load of them is issued
  if (kFar && (L % kEpFence) == kEpFence - 1)
    __syncthreads();
  else
    lds_barrier();
  rpg_load<T, kG, kE>(cur, plo, phi, threads, rdep, rval, rcnt, dinv, in);
}

template <typename T, TriKind K, int kG, int kE, int kD, int kMode>
__global__ __launch_bounds__(512) void k_tri_chain_rpg(
    const int* __restrict__ rdep, const T* __restrict__ rval, const int* __restrict__ rcnt,
    const T* __restrict__ dinv, const int* __restrict__ lev_ptr,
    const int* __restrict__ chain_lev, int rmask, const T* __restrict__ in, T* out) {
  constexpr bool kStage = kMode == kRpgStage;
  constexpr int kA = kD > 2 ? 2 : 1;  // far values are loaded kA levels ahead
  static_assert(kMode != kRpgFar || (kA < kD && kEpFarLevels >= kA + kEpFence),
                "far values need their entries loaded and their writes fenced first");
  extern __shared__ __align__(16) unsigned char smem_g[];
  T* ring = reinterpret_cast<T*>(smem_g);
  T* stage = ring + rmask + 1;  // kE products per thread (staged variant)
  const int threads = static_cast<int>(blockDim.x);
  int* lvo = reinterpret_cast<int*>(stage + (kStage ? static_cast<size_t>(threads) * kE : 0));
  const int t = threadIdx.x;
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int nl = le - lb;
  // the ring starts zeroed: a skipped entry's slot never holds a leftover
  // bit pattern (its product is formed but not subtracted)
  for (int k = t; k <= rmask; k += threads) ring[k] = T(0);
  for (int k = t; k <= nl; k += threads) lvo[k] = lev_ptr[lb + k];
  __syncthreads();
  RpgPre<T, kG, kE> buf[kD];
#pragma unroll
  for (int d = 0; d < kD; ++d) {
    const int Lc = min(d, nl);
    rpg_load<T, kG, kE>(buf[d], lvo[Lc], lvo[min(Lc + 1, nl)], threads, rdep, rval, rcnt, dinv,
                        in);
  }
  int L0 = 0;
  for (; L0 + kD <= nl; L0 += kD) {
#pragma unroll
    for (int d = 0; d < kD; ++d)
      rpg_level<T, K, kG, kE, kD, kMode>(buf[d], buf[(d + kA) % kD], L0 + d, nl, threads, lvo,
                                         ring, stage, rmask, rdep, rval, rcnt, dinv, in, out);
  }
#pragma unroll
  for (int d = 0; d < kD; ++d)
    if (L0 + d < nl)
      rpg_level<T, K, kG, kE, kD, kMode>(buf[d], buf[(d + kA) % kD], L0 + d, nl, threads, lvo,
                                         ring, stage, rmask, rdep, rval, rcnt, dinv, in, out);
}

template <typename T, int kG, int kE, int kMode>
hipError_t tri_rpg_go(const TriLayout& L, const TriSegment& seg, const T* in, T* out,
                      hipStream_t s) {
  constexpr TriKind K = TriKind::kLD;
  constexpr bool kStage = kMode == kRpgStage;
  const size_t lds = rpg_lds_bytes(sizeof(T), L.ring, L.rp_waves, kE, L.ep_maxlev, kStage);
  if (lds > static_cast<size_t>(kTriBlockedMaxLds) || L.rp_waves < 1 || L.rp_waves > 8)
    return hipErrorInvalidValue;
  // prefetch 4 levels ahead; fewer where the registers spill otherwise:
  // 3 for staged FP64 rows of 8-entry chunks, 2 for 16-entry chunks
  constexpr int kD = kE == 16 ? 2 : (kStage && sizeof(T) == 8 && kE == 8) ? 3 : 4;
  auto kern = &k_tri_chain_rpg<T, K, kG, kE, kD, kMode>;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(kTriBlockedMaxLds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(kern, dim3(seg.nchain), dim3(64 * L.rp_waves), static_cast<uint32_t>(lds),
                     s, L.rpf_dep, static_cast<const T*>(L.rpf_val), L.rpf_cnt,
                     static_cast<const T*>(L.ldiag), L.lev_ptr, seg.chain_lev, L.ring - 1, in,
                     out);
  return hipGetLastError();
}

template <typename T>
hipError_t tri_rpg(const TriLayout& L, const TriSegment& seg, const T* in, T* out,
                   hipStream_t s) {
  auto lanes = [&](auto e_tag) -> hipError_t {
    constexpr int E = decltype(e_tag)::value;
    // the layout builder makes: staged products for FP64 rows over 16
    // lanes only, far dependencies for rows over 2 or 4 lanes only, no
    // shuffled FP64 rows of 16 chunks of 8 (whose registers spill), and
    // chunks of 16 for FP32 rows over at most 8 lanes only
    if constexpr (sizeof(T) == 8 && E == 16) return hipErrorInvalidValue;
    else {
      if (L.rp_stage && L.ep_far) return hipErrorInvalidValue;
      if constexpr (sizeof(T) == 8) {
        if (L.rp_stage)
          return L.rp_lanes == 16 ? tri_rpg_go<T, 16, E, kRpgStage>(L, seg, in, out, s)
                                  : hipErrorInvalidValue;
      } else {
        if (L.rp_stage) return hipErrorInvalidValue;
      }
      if (L.ep_far) {
        switch (L.rp_lanes) {
          case 2: return tri_rpg_go<T, 2, E, kRpgFar>(L, seg, in, out, s);
          case 4: return tri_rpg_go<T, 4, E, kRpgFar>(L, seg, in, out, s);
          default: return hipErrorInvalidValue;
        }
      }
      switch (L.rp_lanes) {
        case 2: return tri_rpg_go<T, 2, E, kRpgShuffle>(L, seg, in, out, s);
        case 4: return tri_rpg_go<T, 4, E, kRpgShuffle>(L, seg, in, out, s);
        case 8: return tri_rpg_go<T, 8, E, kRpgShuffle>(L, seg, in, out, s);
        case 16:
          if constexpr ((sizeof(T) == 8 && E == 8) || E == 16) return hipErrorInvalidValue;
          else return tri_rpg_go<T, 16, E, kRpgShuffle>(L, seg, in, out, s);
        default: return hipErrorInvalidValue;
      }
    }
  };
  if (L.rp_e == 4) return lanes(std::integral_constant<int, 4>{});
  if (L.rp_e == 8) return lanes(std::integral_constant<int, 8>{});
  if (L.rp_e == 16) return lanes(std::integral_constant<int, 16>{});
  return hipErrorInvalidValue;
}

// One level, many workgroups, level-ordered layout.
template <typename T, TriKind K>
__global__ __launch_bounds__(kBlock) void k_tri_level_pre(
    const int* __restrict__ lrow, const int* __restrict__ lbeg,
    const int* __restrict__ lcol, const T* __restrict__ lval,
    const int* __restrict__ ecol, const T* __restrict__ eval,
    const T* __restrict__ ldiag, int bit, int dbit, int row_begin, int row_end,
    const T* __restrict__ in, T* out) {
  const int idx = row_begin + blockIdx.x * kBlock + threadIdx.x;
  if (idx >= row_end) return;
  PreRow<T> r;
  pre_meta<T, K>(r, idx, row_end, lrow, lbeg, ecol, eval, ldiag);
  pre_in<T>(r, in);
  out[r.row] = pre_eval<T, K, false>(r, bit, dbit, lcol, lval, out, nullptr, 0);
}

// One level of an integer factor, `G` lanes per row: each lane sums a
// strided share of the row's entries (integer sums are order-free), the
// group reduces with shuffles and its first lane finishes the row.
template <TriKind K, int G>
__global__ __launch_bounds__(kBlock) void k_tri_level_g(
    const int* __restrict__ lrow, const int* __restrict__ lbeg, const int* __restrict__ lcol,
    const int32_t* __restrict__ lval, const int32_t* __restrict__ ldiag, int bit, int dbit,
    int row_begin, int row_end, const int32_t* __restrict__ in, int32_t* out) {
  const int gid = blockIdx.x * kBlock + threadIdx.x;
  const int idx = row_begin + gid / G;
  const int lane = gid & (G - 1);
  uint64_t acc = 0;
  if (idx < row_end) {
    const int pb = lbeg[idx], pe = lbeg[idx + 1];
    for (int p = pb + lane; p < pe; p += G) acc += prod64(lval[p], out[lcol[p]]);
  }
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) acc += shfl_xor_u64(acc, m);
  if (idx < row_end && lane == 0) {
    const int row = lrow[idx];
    const uint64_t tt = (static_cast<uint64_t>(static_cast<int64_t>(in[row])) << bit) - acc;
    out[row] = acc_final<int32_t, K>(tt, ldiag[idx], bit, dbit);
  }
}

template <TriKind K>
hipError_t tri_level_group(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                           const int32_t* in, int32_t* out, hipStream_t s) {
  const long long threads = static_cast<long long>(seg.row_end - seg.row_begin) * seg.group;
  const int grid = static_cast<int>((threads + kBlock - 1) / kBlock);
  if (grid <= 0) return hipSuccess;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, s, L.lrow, L.lbeg, L.lcol,
                       static_cast<const int32_t*>(L.lval),
                       static_cast<const int32_t*>(L.ldiag), bit, dbit, seg.row_begin,
                       seg.row_end, in, out);
    return hipGetLastError();
  };
  switch (seg.group) {
    case 2: return go(&k_tri_level_g<K, 2>);
    case 4: return go(&k_tri_level_g<K, 4>);
    case 8: return go(&k_tri_level_g<K, 8>);
    case 16: return go(&k_tri_level_g<K, 16>);
    case 32: return go(&k_tri_level_g<K, 32>);
    case 64: return go(&k_tri_level_g<K, 64>);
    default: return hipErrorInvalidValue;
  }
}

template <typename T, TriKind K>
hipError_t tri_pre_seg(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                       const void* in, void* out, hipStream_t s) {
  if constexpr (std::is_same_v<T, int32_t>) {
    if (!seg.chain_kernel && seg.group > 1)
      return tri_level_group<K>(L, seg, bit, dbit, static_cast<const int32_t*>(in),
                                static_cast<int32_t*>(out), s);
  }
  const T* vin = static_cast<const T*>(in);
  T* vout = static_cast<T*>(out);
  const T* lval = static_cast<const T*>(L.lval);
  const T* eval = static_cast<const T*>(L.eval);
  const T* ldiag = static_cast<const T*>(L.ldiag);
  if (seg.chain_kernel) {
    if (seg.nchain <= 0) return hipSuccess;
    if (L.ring > 0) {
      const size_t lds = sizeof(T) * static_cast<size_t>(L.ring);
      static bool attr = false;
      if (!attr) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&k_tri_chain_pre<T, K, true>),
            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kTriBlockedMaxLds));
        if (e != hipSuccess) return e;
        attr = true;
      }
      hipLaunchKernelGGL((k_tri_chain_pre<T, K, true>), dim3(seg.nchain), dim3(kChainBlock),
                         lds, s, L.lrow, L.lbeg, L.lpos, lval, L.epos, eval, ldiag, bit, dbit,
                         L.lev_ptr, seg.chain_lev, L.ring - 1, vin, vout);
    } else {
      hipLaunchKernelGGL((k_tri_chain_pre<T, K, false>), dim3(seg.nchain),
                         dim3(kChainBlock), 0, s, L.lrow, L.lbeg, L.lcol, lval, L.ecol, eval,
                         ldiag, bit, dbit, L.lev_ptr, seg.chain_lev, 0, vin, vout);
    }
  } else {
    const int rows = seg.row_end - seg.row_begin;
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_tri_level_pre<T, K>), dim3((rows + kBlock - 1) / kBlock),
                       dim3(kBlock), 0, s, L.lrow, L.lbeg, L.lcol, lval, L.ecol, eval, ldiag,
                       bit, dbit, seg.row_begin, seg.row_end, vin, vout);
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

template <typename T, int MODE>
hipError_t spmv_sell(const DCsr& a, const void* x, const void* z, void* y, hipStream_t s) {
  const int grid = (a.n + kBlock - 1) / kBlock;
  if (grid == 0) return hipSuccess;
  const int remap = a.remap >= 0 ? a.remap : xcd_remap();
  if constexpr (std::is_same_v<T, int32_t>) {
    if (a.scol16) {
      launch_timed(&k_spmv_sell16_int<MODE>, dim3(grid), dim3(kBlock), 0, s, a.n, a.soff,
                   a.scol16, static_cast<const int32_t*>(a.sval), a.bit,
                   static_cast<const int32_t*>(x), static_cast<const int32_t*>(z),
                   static_cast<int32_t*>(y), remap);
      return hipGetLastError();
    }
    launch_timed(&k_spmv_sell_int<MODE>, dim3(grid), dim3(kBlock), 0, s, a.n, a.soff, a.scol,
                 static_cast<const int32_t*>(a.sval), a.srow, a.bit,
                 static_cast<const int32_t*>(x),
                 static_cast<const int32_t*>(z), static_cast<int32_t*>(y), remap);
  } else {
    launch_timed(&k_spmv_sell_fp<T, MODE>, dim3(grid), dim3(kBlock), 0, s, a.n, a.soff, a.scol,
                 static_cast<const T*>(a.sval), a.srow, static_cast<const T*>(x),
                 static_cast<const T*>(z), static_cast<T*>(y), remap);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t spmv_sell_mode(const DCsr& a, int md, const void* x, const void* z, void* y,
                          hipStream_t s) {
  if (md == 0) return spmv_sell<T, 0>(a, x, z, y, s);
  if (md == 1) return spmv_sell<T, 1>(a, x, z, y, s);
  return spmv_sell<T, 2>(a, x, z, y, s);
}

hipError_t launch_spmv(Elem e, const DCsr& a, SpmvMode mode, const void* x,
                       const void* z, void* y, hipStream_t s) {
  const int md = static_cast<int>(mode);
  if (a.soff) {
    if (e == Elem::kI32) return spmv_sell_mode<int32_t>(a, md, x, z, y, s);
    if (e == Elem::kF64) return spmv_sell_mode<double>(a, md, x, z, y, s);
    return spmv_sell_mode<float>(a, md, x, z, y, s);
  }
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

hipError_t launch_trsv_tiles(TriKind kind, int n, int nb, const int32_t* tiles,
                             const long long* toff, const int32_t* dinv, int bit, int dbit,
                             const void* in, void* out, unsigned long long* xp, int epoch,
                             int* err, unsigned* tickets, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  auto go = [&](auto kern) -> hipError_t {
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kTileThreads), 0, s, n, nb, tiles, toff, dinv, bit,
                       dbit, static_cast<const int32_t*>(in), static_cast<int32_t*>(out), xp,
                       epoch, err, tickets);
    return hipGetLastError();
  };
  if (kind == TriKind::kL) return go(&k_tri_tiles_int<TriKind::kL>);
  if (kind == TriKind::kDU) return go(&k_tri_tiles_int<TriKind::kDU>);
  return go(&k_tri_tiles_int<TriKind::kLD>);
}

hipError_t launch_trsv_tiles_fp(Elem e, TriKind kind, int n, int nb, const void* tiles,
                                const long long* toff, const uint64_t* tmask, const void* dinv,
                                const void* in, void* out, unsigned long long* flags, int epoch,
                                int* err, unsigned* tickets, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (nb > kTileMaxBlocks) return hipErrorInvalidValue;
  auto go = [&](auto tag, auto kern) -> hipError_t {
    using T = decltype(tag);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(64), 0, s, n, nb, static_cast<const T*>(tiles), toff,
                       tmask, static_cast<const T*>(dinv), static_cast<const T*>(in),
                       static_cast<T*>(out), flags, epoch, err, tickets);
    return hipGetLastError();
  };
  if (e == Elem::kF64) {
    if (kind == TriKind::kL) return go(double{}, &k_tri_tiles_fp<double, TriKind::kL>);
    if (kind == TriKind::kDU) return go(double{}, &k_tri_tiles_fp<double, TriKind::kDU>);
    return go(double{}, &k_tri_tiles_fp<double, TriKind::kLD>);
  }
  if (e == Elem::kF32) {
    if (kind == TriKind::kL) return go(float{}, &k_tri_tiles_fp<float, TriKind::kL>);
    if (kind == TriKind::kDU) return go(float{}, &k_tri_tiles_fp<float, TriKind::kDU>);
    return go(float{}, &k_tri_tiles_fp<float, TriKind::kLD>);
  }
  return hipErrorInvalidValue;
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

hipError_t launch_trsv_ep(TriKind kind, const TriLayout& L, const TriSegment& seg, int bit,
                          int dbit, const void* in_v, void* out_v, hipStream_t s) {
  const int32_t* in = static_cast<const int32_t*>(in_v);
  int32_t* o = static_cast<int32_t*>(out_v);
  const int w = L.ep_waves;
  if (kind == TriKind::kLD) {
    if (w <= 1) return tri_ep_launch<TriKind::kLD, 1>(L, seg, bit, dbit, in, o, s);
    if (w <= 2) return tri_ep_launch<TriKind::kLD, 2>(L, seg, bit, dbit, in, o, s);
    if (w <= 4) return tri_ep_launch<TriKind::kLD, 4>(L, seg, bit, dbit, in, o, s);
    return tri_ep_launch<TriKind::kLD, 8>(L, seg, bit, dbit, in, o, s);
  }
  if (kind == TriKind::kL) {
    if (w <= 1) return tri_ep_launch<TriKind::kL, 1>(L, seg, bit, dbit, in, o, s);
    if (w <= 2) return tri_ep_launch<TriKind::kL, 2>(L, seg, bit, dbit, in, o, s);
    if (w <= 4) return tri_ep_launch<TriKind::kL, 4>(L, seg, bit, dbit, in, o, s);
    return tri_ep_launch<TriKind::kL, 8>(L, seg, bit, dbit, in, o, s);
  }
  if (w <= 1) return tri_ep_launch<TriKind::kDU, 1>(L, seg, bit, dbit, in, o, s);
  if (w <= 2) return tri_ep_launch<TriKind::kDU, 2>(L, seg, bit, dbit, in, o, s);
  if (w <= 4) return tri_ep_launch<TriKind::kDU, 4>(L, seg, bit, dbit, in, o, s);
  return tri_ep_launch<TriKind::kDU, 8>(L, seg, bit, dbit, in, o, s);
}

hipError_t launch_trsv_rp(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                          const void* in_v, void* out_v, hipStream_t s) {
  const int32_t* in = static_cast<const int32_t*>(in_v);
  int32_t* o = static_cast<int32_t*>(out_v);
  const int w = L.rp_waves;
  if (w <= 1) return tri_rp_far<1>(L, seg, bit, dbit, in, o, s);
  if (w <= 2) return tri_rp_far<2>(L, seg, bit, dbit, in, o, s);
  if (w <= 3) return tri_rp_far<3>(L, seg, bit, dbit, in, o, s);
  if (w <= 4) return tri_rp_far<4>(L, seg, bit, dbit, in, o, s);
  if (w <= 5) return tri_rp_far<5>(L, seg, bit, dbit, in, o, s);
  if (w <= 6) return tri_rp_far<6>(L, seg, bit, dbit, in, o, s);
  if (w <= 7) return tri_rp_far<7>(L, seg, bit, dbit, in, o, s);
  return tri_rp_far<8>(L, seg, bit, dbit, in, o, s);
}

hipError_t launch_trsv_rpf(Elem e, const TriLayout& L, const TriSegment& seg, const void* in,
                           void* out, hipStream_t s) {
  if (L.rp_lanes > 1) {
    if (e == Elem::kF64)
      return tri_rpg<double>(L, seg, static_cast<const double*>(in), static_cast<double*>(out),
                             s);
    if (e == Elem::kF32)
      return tri_rpg<float>(L, seg, static_cast<const float*>(in), static_cast<float*>(out), s);
    return hipErrorInvalidValue;
  }
  if (e == Elem::kF64)
    return tri_rpf_waves<double>(L, seg, static_cast<const double*>(in),
                                 static_cast<double*>(out), s);
  if (e == Elem::kF32)
    return tri_rpf_waves<float>(L, seg, static_cast<const float*>(in), static_cast<float*>(out),
                                s);
  return hipErrorInvalidValue;
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
