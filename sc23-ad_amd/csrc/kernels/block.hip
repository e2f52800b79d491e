// Block kernels: several right-hand sides of one hierarchy in one launch.
//
// k independent integer solves on the same hierarchy (INTAMG.hpp:269-293
// applied to k vectors) share every matrix, factor and level schedule, so a
// kernel loads each entry once and applies it to all of them.  Vectors of a
// block are stored in groups of four right-hand sides: element (row i,
// right-hand side r) sits at ((r / 4) * n + i) * 4 + r % 4, so the four
// values of a row form one 16-byte word.  Every operation is integer and
// per element (the sums of a row are order-free), so each right-hand side's
// values are those its own single solve computes, bit for bit.
#include <hip/hip_ext.h>

#include "dev_common.hpp"
#include "launch.hpp"
#include "tri_common.hpp"

namespace sc23ad::dev {

namespace {

constexpr int kBk = 256;
constexpr int kEpFenceBlk = 8;  // full fence period in levels (kernels with far entries)

__device__ __forceinline__ int32_t lane_of(const int4& v, int c) {
  return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

__device__ __forceinline__ int4 fin4(const uint64_t* acc, int bit, int mode, int4 z) {
  int4 y;
  int32_t s[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) s[c] = low32(as_i64(acc[c]) >> bit);
  if (mode == 0) {
    y = make_int4(s[0], s[1], s[2], s[3]);
  } else if (mode == 1) {
    y = make_int4(sub32(z.x, s[0]), sub32(z.y, s[1]), sub32(z.z, s[2]), sub32(z.w, s[3]));
  } else {
    y = make_int4(add32(z.x, s[0]), add32(z.y, s[1]), add32(z.z, s[2]), add32(z.w, s[3]));
  }
  return y;
}

// v = take ? f : v, component by component
__device__ __forceinline__ void sel4(int4& v, bool take, const int4& f) {
  v.x = take ? f.x : v.x;
  v.y = take ? f.y : v.y;
  v.z = take ? f.z : v.z;
  v.w = take ? f.w : v.w;
}

__device__ __forceinline__ void mac4(uint64_t* acc, int32_t v, const int4& x) {
  acc[0] += prod64(v, x.x);
  acc[1] += prod64(v, x.y);
  acc[2] += prod64(v, x.z);
  acc[3] += prod64(v, x.w);
}

// ---- permutation into / out of a level's smoother order -------------------
template <bool kGather>
__global__ __launch_bounds__(kBk) void k_permute_blk(int n, int groups, const int* __restrict__ perm,
                                                     const int4* __restrict__ src,
                                                     int4* __restrict__ dst) {
  const long long t = static_cast<long long>(blockIdx.x) * kBk + threadIdx.x;
  if (t >= static_cast<long long>(n) * groups) return;
  const int g = static_cast<int>(t / n), k = static_cast<int>(t % n);
  const size_t base = static_cast<size_t>(g) * n;
  if (kGather) dst[base + k] = src[base + perm[k]];
  else dst[base + perm[k]] = src[base + k];
}

__global__ __launch_bounds__(kBk) void k_copy_blk(size_t n4, const int4* __restrict__ src,
                                                  int4* __restrict__ dst) {
  for (size_t i = blockIdx.x * static_cast<size_t>(kBk) + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * kBk)
    dst[i] = src[i];
}

// One right-hand side of a block <-> a plain vector.
template <bool kGet>
__global__ __launch_bounds__(kBk) void k_blk_column(int n, int r, int32_t* __restrict__ blk,
                                                   int32_t* __restrict__ col) {
  const int i = blockIdx.x * kBk + threadIdx.x;
  if (i >= n) return;
  const size_t at = (static_cast<size_t>(r >> 2) * n + i) * 4 + (r & 3);
  if (kGet) col[i] = blk[at];
  else blk[at] = col[i];
}

// ---- SpMV: y_r = (int32)((A x_r) >> bit) for every right-hand side r -----
// One thread per row; the row's entries are read eight at a time (clamped,
// every load unconditional) and applied to the `groups` four-wide words of
// each gathered column.  MODE 1 / 2 fuse y = z - A x / y = z + A x.
template <int MODE, int kGroups>
__global__ __launch_bounds__(kBk) void k_spmv_blk_csr(int n, int m, const int* __restrict__ rptr,
                                                      const int* __restrict__ cind,
                                                      const int32_t* __restrict__ val, int bit,
                                                      const int4* __restrict__ x,
                                                      const int4* z, int4* y) {
  const int row = blockIdx.x * kBk + threadIdx.x;
  if (row >= n) return;
  uint64_t acc[kGroups][4];
#pragma unroll
  for (int g = 0; g < kGroups; ++g)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[g][c] = 0;
  const int rb = rptr[row], re = rptr[row + 1];
  for (int p0 = rb; p0 < re; p0 += 8) {
    int cc[8];
    int32_t vv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = p0 + j < re ? p0 + j : p0;
      cc[j] = cind[p];
      vv[j] = p0 + j < re ? val[p] : 0;  // padding adds an exact zero
    }
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
      int4 xx[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) xx[j] = x[static_cast<size_t>(g) * m + cc[j]];
#pragma unroll
      for (int j = 0; j < 8; ++j) mac4(acc[g], vv[j], xx[j]);
    }
  }
#pragma unroll
  for (int g = 0; g < kGroups; ++g) {
    const size_t at = static_cast<size_t>(g) * n + row;
    const int4 zv = MODE != 0 ? z[at] : make_int4(0, 0, 0, 0);
    y[at] = fin4(acc[g], bit, MODE, zv);
  }
}

template <int MODE, int kGroups>
__global__ __launch_bounds__(kBk) void k_spmv_blk_sell(int n, int m,
                                                       const long long* __restrict__ soff,
                                                       const int* __restrict__ scol,
                                                       const int32_t* __restrict__ sval,
                                                       const int* __restrict__ srow, int bit,
                                                       const int4* __restrict__ x,
                                                       const int4* z, int4* y) {
  const int row = blockIdx.x * kBk + threadIdx.x;
  if (row >= n) return;
  const int orow = srow ? srow[row] : row;
  const long long b = soff[row >> 6];
  const int w = static_cast<int>((soff[(row >> 6) + 1] - b) >> 6);
  const int* c = scol + b + (row & 63);
  const int32_t* v = sval + b + (row & 63);
  uint64_t acc[kGroups][4];
#pragma unroll
  for (int g = 0; g < kGroups; ++g)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[g][q] = 0;
  for (int k0 = 0; k0 < w; k0 += 8) {
    int cc[8];
    int32_t vv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + j < w ? k0 + j : k0;
      const int col = c[64 * k];
      cc[j] = col < 0 ? 0 : col;
      vv[j] = (k0 + j < w && col >= 0) ? v[64 * k] : 0;
    }
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
      int4 xx[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) xx[j] = x[static_cast<size_t>(g) * m + cc[j]];
#pragma unroll
      for (int j = 0; j < 8; ++j) mac4(acc[g], vv[j], xx[j]);
    }
  }
#pragma unroll
  for (int g = 0; g < kGroups; ++g) {
    const size_t at = static_cast<size_t>(g) * n + orow;
    const int4 zv = MODE != 0 ? z[at] : make_int4(0, 0, 0, 0);
    y[at] = fin4(acc[g], bit, MODE, zv);
  }
}

// ---- row-parallel LD solve, four right-hand sides per workgroup -----------
// The layout, chains, levels, LDS ring and far rule of k_tri_chain_rp (the
// single-vector kernel): workgroup (chain, group) walks the levels of its
// chain for the four right-hand sides of its group.  A level's entries and
// inverse diagonals are loaded once; every ring slot holds the four values
// of one position.  Waves per workgroup are a launch parameter.
template <int kE, int kR>
struct RpbPre {
  int lo, hi;
  int4 ent[kR][kE / 2];  // {dep, val, dep, val}
  int4 xfar[kR][kE];
  int32_t dinv[kR];
  int4 in[kR];
};

template <int kG, int kE, int kR>
__device__ __forceinline__ void rpb_load(RpbPre<kE, kR>& r, int lo, int hi, int threads,
                                         const int4* __restrict__ rent,
                                         const int32_t* __restrict__ dinv,
                                         const int4* __restrict__ in) {
  const int t = threadIdx.x;
  r.lo = lo;
  r.hi = hi;
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int idx = lo + t / kG + q * (threads / kG);
    const int id = idx < hi ? idx : 0;
    const size_t base = (static_cast<size_t>(id) * kG + t % kG) * (kE / 2);
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) r.ent[q][h] = rent[base + h];
    r.dinv[q] = dinv[id];
    r.in[q] = in[id];
  }
}

template <int kG, int kE, int kR, int kFq>
__device__ __forceinline__ void rpb_far(RpbPre<kE, kR>& r, int threads, const int4* out) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const bool ok = r.lo + t / kG + q * (threads / kG) < r.hi;
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) {
      const int d0 = r.ent[q][h].x, d1 = r.ent[q][h].z;
      if (2 * h < kFq) r.xfar[q][2 * h] = out[(ok && d0 < 0) ? -d0 - 1 : 0];
      if (2 * h + 1 < kFq) r.xfar[q][2 * h + 1] = out[(ok && d1 < 0) ? -d1 - 1 : 0];
    }
  }
}

template <bool kFar, int kG, int kE, int kR, int kD, int kA, int kFq>
__device__ __forceinline__ void rpb_level(RpbPre<kE, kR>& cur, RpbPre<kE, kR>& ahead, int L,
                                          int nl, int threads, const int* lvo, int4* ring,
                                          int rmask, const int4* __restrict__ rent,
                                          const int32_t* __restrict__ dinv,
                                          const int4* __restrict__ in, int4* out, int bit,
                                          int dbit) {
  const int t = threadIdx.x;
  const int s = cur.lo, e = cur.hi;
  const int Lp = min(L + kD, nl);
  const int plo = lvo[Lp], phi = lvo[min(Lp + 1, nl)];
  if (kFar) rpb_far<kG, kE, kR, kFq>(ahead, threads, out);
  int4 xv[kR][kE];
#pragma unroll
  for (int q = 0; q < kR; ++q)
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) {
      xv[q][2 * h] = ring[cur.ent[q][h].x & rmask];
      xv[q][2 * h + 1] = ring[cur.ent[q][h].z & rmask];
    }
  // rows of 8 or more entries: every ring read issued before the first
  // product (one LDS round trip per level, as in k_tri_chain_rp)
  if constexpr (kE >= 8) __builtin_amdgcn_sched_barrier(0);
  if (kFar) {
#pragma unroll
    for (int q = 0; q < kR; ++q)
#pragma unroll
      for (int h = 0; h < kE / 2; ++h) {
        const int d0 = cur.ent[q][h].x, d1 = cur.ent[q][h].z;
        if (2 * h < kFq && __any(d0 < 0)) sel4(xv[q][2 * h], d0 < 0, cur.xfar[q][2 * h]);
        if (2 * h + 1 < kFq && __any(d1 < 0))
          sel4(xv[q][2 * h + 1], d1 < 0, cur.xfar[q][2 * h + 1]);
      }
  }
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    uint64_t acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < kE / 2; ++h) {
      mac4(acc, cur.ent[q][h].y, xv[q][2 * h]);
      mac4(acc, cur.ent[q][h].w, xv[q][2 * h + 1]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = group_sum<kG>(acc[c]);
    const int k = s + t / kG + q * (threads / kG);
    if (k < e && t % kG == 0) {
      const int4 iv = cur.in[q];
      const int32_t dv = cur.dinv[q];
      int4 v;
      v.x = acc_final<int32_t, TriKind::kLD>(
          (static_cast<uint64_t>(static_cast<int64_t>(iv.x)) << bit) - acc[0], dv, bit, dbit);
      v.y = acc_final<int32_t, TriKind::kLD>(
          (static_cast<uint64_t>(static_cast<int64_t>(iv.y)) << bit) - acc[1], dv, bit, dbit);
      v.z = acc_final<int32_t, TriKind::kLD>(
          (static_cast<uint64_t>(static_cast<int64_t>(iv.z)) << bit) - acc[2], dv, bit, dbit);
      v.w = acc_final<int32_t, TriKind::kLD>(
          (static_cast<uint64_t>(static_cast<int64_t>(iv.w)) << bit) - acc[3], dv, bit, dbit);
      ring[k & rmask] = v;
      out[k] = v;
    }
  }
  if (kFar && (L % kEpFenceBlk) == kEpFenceBlk - 1)
    __syncthreads();  // full fence: global results visible to later loads
  else
    lds_barrier();
  rpb_load<kG>(cur, plo, phi, threads, rent, dinv, in);
}

// LDS: the ring (int4 per position), then the chain's level offsets.
template <bool kFar, int kG, int kE, int kR, int kD, int kA, int kFq>
__global__ __launch_bounds__(512) void k_tri_chain_rpb(const int4* __restrict__ rent,
                                                       const int32_t* __restrict__ dinv,
                                                       const int* __restrict__ lev_ptr,
                                                       const int* __restrict__ chain_lev,
                                                       int rmask, int bit, int dbit, int n,
                                                       const int4* __restrict__ in_all,
                                                       int4* out_all) {
  static_assert(kA >= 1 && kA < kD, "far values need their entries loaded first");
  static_assert(kRpbFarLevels >= kA + kEpFenceBlk,
                "far dependencies must be fenced before their values are loaded");
  extern __shared__ __align__(16) unsigned char smem_b[];
  int4* ring = reinterpret_cast<int4*>(smem_b);
  int* lvo = reinterpret_cast<int*>(ring + rmask + 1);
  const int threads = static_cast<int>(blockDim.x);
  const int t = threadIdx.x;
  const size_t goff = static_cast<size_t>(blockIdx.y) * n;
  const int4* in = in_all + goff;
  int4* out = out_all + goff;
  const int lb = chain_lev[2 * blockIdx.x], le = chain_lev[2 * blockIdx.x + 1];
  const int nl = le - lb;
  for (int k = t; k <= nl; k += threads) lvo[k] = lev_ptr[lb + k];
  __syncthreads();
  RpbPre<kE, kR> buf[kD];
#pragma unroll
  for (int d = 0; d < kD; ++d) {
    const int Lc = min(d, nl);
    rpb_load<kG>(buf[d], lvo[Lc], lvo[min(Lc + 1, nl)], threads, rent, dinv, in);
  }
  int L0 = 0;
  for (; L0 + kD <= nl; L0 += kD) {
#pragma unroll
    for (int d = 0; d < kD; ++d)
      rpb_level<kFar, kG, kE, kR, kD, kA, kFq>(buf[d], buf[(d + kA) % kD], L0 + d, nl, threads,
                                               lvo, ring, rmask, rent, dinv, in, out, bit, dbit);
  }
#pragma unroll
  for (int d = 0; d < kD; ++d)
    if (L0 + d < nl)
      rpb_level<kFar, kG, kE, kR, kD, kA, kFq>(buf[d], buf[(d + kA) % kD], L0 + d, nl, threads,
                                               lvo, ring, rmask, rent, dinv, in, out, bit, dbit);
}

template <bool kFar, int kG, int kE, int kR, int kD, int kA, int kFq>
hipError_t rpb_go(const TriLayout& L, const TriSegment& seg, int groups, int bit, int dbit,
                  const int4* in, int4* out, hipStream_t s) {
  const size_t lds = rpb_lds_bytes(L.ring, L.ep_maxlev);
  if (lds > static_cast<size_t>(kTriBlockedMaxLds)) return hipErrorInvalidValue;
  auto kern = &k_tri_chain_rpb<kFar, kG, kE, kR, kD, kA, kFq>;
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(kTriBlockedMaxLds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(kern, dim3(seg.nchain, groups), dim3(64 * L.rp_waves),
                     static_cast<uint32_t>(lds), s, static_cast<const int4*>(L.rent),
                     static_cast<const int32_t*>(L.ldiag), L.lev_ptr, seg.chain_lev, L.ring - 1,
                     bit, dbit, L.n, in, out);
  return hipGetLastError();
}

// Prefetch distances by layout (the registers of deeper buffers spill):
// one row per lane: 6 levels for 4 entries, 4 for 8, 3 for 10 and 12; two
// rows per lane: 4 for 4 entries, 2 otherwise.  Far values are loaded for
// the slots that can hold one (a row's far entries come first): 1 with 4
// entries, 2 with more, and with one row per lane also 4 or every slot for
// factors whose rows have more far entries; with several lanes per row (one
// row per lane) the far entries are dealt over the lanes, slot 0 of each
// lane, then slot 1, up to four slots per lane (eight with 10 or 12
// entries).
template <bool kFar, int kE, int kR>
hipError_t rpb_one_lane(const TriLayout& L, const TriSegment& seg, int groups, int bit, int dbit,
                        const int4* in, int4* out, hipStream_t s) {
  constexpr int kFq = kE == 4 ? 1 : 2;
  constexpr int kD = kR == 1 ? (kE == 4 ? 6 : kE == 8 ? 4 : 3) : (kE == 4 ? 4 : 2);
  constexpr int kA = kR == 1 ? (kE == 4 ? 3 : kE == 8 ? 2 : 1) : (kE == 4 ? 2 : 1);
  if constexpr (kFar && kR == 1) {
    // rows with more far entries (a ring of four values per slot holds a
    // quarter of the single kernel's span): 4 far slots, or every slot
    if (L.rp_farq > kFq) {
      if (kE > 4 && L.rp_farq <= 4)
        return rpb_go<kFar, 1, kE, kR, kD, kA, 4>(L, seg, groups, bit, dbit, in, out, s);
      // every slot far: one level less of prefetch for 8 and 12 entries,
      // whose registers spill otherwise
      constexpr int kDa = (kE == 8 || kE == 12) ? kD - 1 : kD;
      return rpb_go<kFar, 1, kE, kR, kDa, kA, kE>(L, seg, groups, bit, dbit, in, out, s);
    }
  }
  return rpb_go<kFar, 1, kE, kR, kD, kA, kFq>(L, seg, groups, bit, dbit, in, out, s);
}

template <bool kFar, int kG, int kE, int kR>
hipError_t rpb_lanes(const TriLayout& L, const TriSegment& seg, int groups, int bit, int dbit,
                     const int4* in, int4* out, hipStream_t s) {
  constexpr int kD = kR == 1 ? (kE == 8 ? 4 : 3) : 2;
  constexpr int kA = kR == 1 ? (kE == 8 ? 2 : 1) : 1;
  if constexpr (kFar && kR == 1) {
    // a row's far entries come first, dealt over its lanes: slot 0 of each
    // lane, then slot 1
    if (L.rp_farq <= kG)
      return rpb_go<true, kG, kE, kR, kD, kA, 1>(L, seg, groups, bit, dbit, in, out, s);
    // two or four far slots per lane: less prefetch where the registers
    // would spill otherwise
    constexpr int kD2 = (kE == 8 || kE == 12) ? kD - 1 : kD;
    if (L.rp_farq <= 2 * kG)
      return rpb_go<true, kG, kE, kR, kD2, kA, 2>(L, seg, groups, bit, dbit, in, out, s);
    constexpr int kD4 = kA + 1;
    if (kE < 10 || L.rp_farq <= 4 * kG)
      return rpb_go<true, kG, kE, kR, kD4, kA, 4>(L, seg, groups, bit, dbit, in, out, s);
    if constexpr (kE >= 10)  // eight far slots per lane (10 or 12 entries)
      return rpb_go<true, kG, kE, kR, kD4, kA, 8>(L, seg, groups, bit, dbit, in, out, s);
    return hipErrorInvalidValue;
  } else {
    return rpb_go<false, kG, kE, kR, kD, kA, 0>(L, seg, groups, bit, dbit, in, out, s);
  }
}

template <bool kFar, int kR>
hipError_t rpb_rows(const TriLayout& L, const TriSegment& seg, int groups, int bit, int dbit,
                    const int4* in, int4* out, hipStream_t s) {
  if (L.rp_lanes <= 1) {
    switch (L.rp_e) {
      case 4: return rpb_one_lane<kFar, 4, kR>(L, seg, groups, bit, dbit, in, out, s);
      case 8: return rpb_one_lane<kFar, 8, kR>(L, seg, groups, bit, dbit, in, out, s);
      case 10: return rpb_one_lane<kFar, 10, kR>(L, seg, groups, bit, dbit, in, out, s);
      case 12: return rpb_one_lane<kFar, 12, kR>(L, seg, groups, bit, dbit, in, out, s);
      default: return hipErrorInvalidValue;
    }
  }
  if constexpr (!kFar || kR == 1) {  // far entries with several lanes: one row per lane
    auto lanes = [&](auto e_tag) -> hipError_t {
      constexpr int E = decltype(e_tag)::value;
      switch (L.rp_lanes) {
        case 2: return rpb_lanes<kFar, 2, E, kR>(L, seg, groups, bit, dbit, in, out, s);
        case 4: return rpb_lanes<kFar, 4, E, kR>(L, seg, groups, bit, dbit, in, out, s);
        case 8: return rpb_lanes<kFar, 8, E, kR>(L, seg, groups, bit, dbit, in, out, s);
        case 16: return rpb_lanes<kFar, 16, E, kR>(L, seg, groups, bit, dbit, in, out, s);
        default: return hipErrorInvalidValue;
      }
    };
    if (L.rp_e == 8) return lanes(std::integral_constant<int, 8>{});
    if constexpr (kR == 1) {  // two rows of 10 or 12 entries per lane spill
      if (L.rp_e == 10) return lanes(std::integral_constant<int, 10>{});
      if (L.rp_e == 12) return lanes(std::integral_constant<int, 12>{});
    }
  }
  return hipErrorInvalidValue;
}

template <bool kFar>
hipError_t rpb_layout(const TriLayout& L, const TriSegment& seg, int groups, int bit, int dbit,
                      const int4* in, int4* out, hipStream_t s) {
  if (L.rp_rows == 2) return rpb_rows<kFar, 2>(L, seg, groups, bit, dbit, in, out, s);
  return rpb_rows<kFar, 1>(L, seg, groups, bit, dbit, in, out, s);
}

// ---- tiled coarse solve, four right-hand sides per workgroup --------------
// k_tri_tiles_int (the single-vector kernel) for a group of four right-hand
// sides: workgroup (block b, group g) owns rows 64b .. 64b+63 of group g;
// the four waves share the dependency products, then wave q sweeps the
// diagonal tile for right-hand side q.  A
// solved row publishes one {epoch, value} word per right-hand side
// (xp[((g * n) + row) * 4 + c]); a consumer polls the 4 x 64 words of a
// dependency block until all carry the current epoch.  Workgroups take
// their (block, group) from a ticket counter, in processing order, so a
// workgroup only waits for tickets handed out before its own, whatever
// order the hardware starts workgroups in; the last workgroup to finish
// resets the counters for the next launch.  A launch out of its wait budget
// (tile_wait_over) records an error and gives up instead of hanging.
template <TriKind K>
__global__ __launch_bounds__(kTileThreads) void k_tri_tiles_blk(
    int n, int nb, int groups, const int32_t* __restrict__ tiles,
    const long long* __restrict__ toff, const int32_t* __restrict__ dinv_g, int bit, int dbit,
    const int4* __restrict__ in_all, int4* out_all, unsigned long long* xp_all, int epoch,
    int* err, unsigned* tickets) {
  static_assert(kTileThreads == 256, "one wave per right-hand side of a group of four");
  constexpr bool fwd = (K != TriKind::kDU);
  __shared__ uint64_t part[kTileThreads / 64][4][64];
  __shared__ int32_t xsh[4][64];
  __shared__ int ticket_sh;
  const long long t0 = wall_clock64();
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ticket = tile_ticket(tickets, &ticket_sh);
  // tickets in processing order (forward: blocks 0, 1, ...; backward: nb-1,
  // nb-2, ...), block-major, so every block's dependency blocks (all
  // groups) hold smaller tickets
  const int pb = ticket / groups, g = ticket % groups;
  const int b = fwd ? pb : nb - 1 - pb;
  const size_t goff = static_cast<size_t>(g) * n;
  const int4* in = in_all + goff;
  int4* out = out_all + goff;
  unsigned long long* xp = xp_all + goff * 4;
  const int r0 = b * kTB;
  const int rows = min(kTB, n - r0);
  const int32_t* T = tiles + toff[b];
  const int ndep = fwd ? b : nb - 1 - b;
  // wave w owns right-hand side w of the group: its accumulators, its
  // reduction of the dependency products and its diagonal sweep, so the
  // four sweeps run side by side on the CU's four SIMDs
  const int32_t* in32 = reinterpret_cast<const int32_t*>(in);
  uint64_t t = 0;
  if (lane < rows)
    t = static_cast<uint64_t>(static_cast<int64_t>(in32[static_cast<size_t>(r0 + lane) * 4 + w]))
        << bit;
  // the diagonal tile and the row scales depend on no other block: loaded
  // before the first wait, so their latency stays off the chain
  int32_t reg[kTB];
  const int32_t* Td = T + static_cast<size_t>(ndep) * (kTB * kTB);
#pragma unroll
  for (int c = 0; c < kTB; ++c) reg[c] = Td[c * kTB + lane];
  const int32_t dv = (K != TriKind::kL && lane < rows) ? dinv_g[r0 + lane] : 0;
  for (int k = 0; k < ndep; ++k) {
    const int bp = fwd ? k : nb - 1 - k;
    const int32_t* Tk = T + static_cast<size_t>(k) * (kTB * kTB);
    int32_t tv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) tv[j] = Tk[(16 * w + j) * kTB + lane];
    if (w == 0) {
      const int c = bp * kTB + lane;
      unsigned long long v[4] = {0, 0, 0, 0};
      int spins = 0;
      while (true) {
        bool ready = true;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          v[q] = c < n ? __hip_atomic_load(&xp[static_cast<size_t>(c) * 4 + q], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT)
                       : static_cast<unsigned long long>(static_cast<unsigned>(epoch)) << 32;
          ready = ready && static_cast<int>(v[q] >> 32) == epoch;
        }
        if (__all(ready)) break;
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 31) == 0 && tile_wait_over(t0, err)) {
          if (lane == 0) atomicAdd(err, 1);
          break;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        xsh[q][lane] = c < n ? static_cast<int32_t>(static_cast<uint32_t>(v[q])) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint64_t ps = 0;
#pragma unroll
      for (int j = 0; j < 16; ++j) ps += prod64(tv[j], xsh[q][16 * w + j]);
      part[w][q][lane] = ps;
    }
    __syncthreads();
#pragma unroll
    for (int ww = 0; ww < kTileThreads / 64; ++ww) t -= part[ww][w][lane];
  }
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
    reinterpret_cast<int32_t*>(out)[static_cast<size_t>(r0 + lane) * 4 + w] = xv;
    const unsigned long long word =
        (static_cast<unsigned long long>(static_cast<unsigned>(epoch)) << 32) |
        static_cast<uint32_t>(xv);
    __hip_atomic_store(&xp[static_cast<size_t>(r0 + lane) * 4 + w], word, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  tile_done(tickets, nb * groups);
}

}  // namespace

hipError_t launch_permute_blk(int n, int groups, const int* perm, const void* src, void* dst,
                              bool gather, hipStream_t s) {
  const long long tot = static_cast<long long>(n) * groups;
  if (tot <= 0) return hipSuccess;
  const int grid = static_cast<int>((tot + kBk - 1) / kBk);
  if (gather)
    hipLaunchKernelGGL(k_permute_blk<true>, dim3(grid), dim3(kBk), 0, s, n, groups, perm,
                       static_cast<const int4*>(src), static_cast<int4*>(dst));
  else
    hipLaunchKernelGGL(k_permute_blk<false>, dim3(grid), dim3(kBk), 0, s, n, groups, perm,
                       static_cast<const int4*>(src), static_cast<int4*>(dst));
  return hipGetLastError();
}

hipError_t launch_copy_blk(int n, int groups, const void* src, void* dst, hipStream_t s) {
  const size_t n4 = static_cast<size_t>(n) * groups;
  if (n4 == 0) return hipSuccess;
  const int grid = static_cast<int>(std::min<size_t>((n4 + kBk - 1) / kBk, 8192));
  hipLaunchKernelGGL(k_copy_blk, dim3(grid), dim3(kBk), 0, s, n4, static_cast<const int4*>(src),
                     static_cast<int4*>(dst));
  return hipGetLastError();
}

hipError_t launch_blk_column(int n, int r, void* blk, void* col, bool get, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (n + kBk - 1) / kBk;
  if (get)
    hipLaunchKernelGGL(k_blk_column<true>, dim3(grid), dim3(kBk), 0, s, n, r,
                       static_cast<int32_t*>(blk), static_cast<int32_t*>(col));
  else
    hipLaunchKernelGGL(k_blk_column<false>, dim3(grid), dim3(kBk), 0, s, n, r,
                       static_cast<int32_t*>(blk), static_cast<int32_t*>(col));
  return hipGetLastError();
}

template <int MODE, int kGroups>
static hipError_t spmv_blk_go(const DCsr& a, const void* x, const void* z, void* y,
                              hipStream_t s) {
  const int grid = (a.n + kBk - 1) / kBk;
  if (grid == 0) return hipSuccess;
  const int4* xv = static_cast<const int4*>(x);
  const int4* zv = static_cast<const int4*>(z);
  int4* yv = static_cast<int4*>(y);
  if (a.soff && a.scol)
    hipLaunchKernelGGL((k_spmv_blk_sell<MODE, kGroups>), dim3(grid), dim3(kBk), 0, s, a.n, a.m,
                       a.soff, a.scol, static_cast<const int32_t*>(a.sval), a.srow, a.bit, xv,
                       zv, yv);
  else
    hipLaunchKernelGGL((k_spmv_blk_csr<MODE, kGroups>), dim3(grid), dim3(kBk), 0, s, a.n, a.m,
                       a.rptr, a.cind, static_cast<const int32_t*>(a.val), a.bit, xv, zv, yv);
  return hipGetLastError();
}

template <int MODE>
static hipError_t spmv_blk_mode(const DCsr& a, int groups, const void* x, const void* z,
                                void* y, hipStream_t s) {
  // up to four groups per launch (16 right-hand sides); more in slices
  for (int g0 = 0; g0 < groups; g0 += 4) {
    const int gs = std::min(4, groups - g0);
    const size_t xo = static_cast<size_t>(g0) * a.m, yo = static_cast<size_t>(g0) * a.n;
    const int4* xx = static_cast<const int4*>(x) + xo;
    const int4* zz = z ? static_cast<const int4*>(z) + yo : nullptr;
    int4* yy = static_cast<int4*>(y) + yo;
    hipError_t e;
    switch (gs) {
      case 1: e = spmv_blk_go<MODE, 1>(a, xx, zz, yy, s); break;
      case 2: e = spmv_blk_go<MODE, 2>(a, xx, zz, yy, s); break;
      case 3: e = spmv_blk_go<MODE, 3>(a, xx, zz, yy, s); break;
      default: e = spmv_blk_go<MODE, 4>(a, xx, zz, yy, s); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_spmv_blk(const DCsr& a, int groups, SpmvMode mode, const void* x,
                           const void* z, void* y, hipStream_t s) {
  if (mode == SpmvMode::kPlain) return spmv_blk_mode<0>(a, groups, x, z, y, s);
  if (mode == SpmvMode::kSub) return spmv_blk_mode<1>(a, groups, x, z, y, s);
  return spmv_blk_mode<2>(a, groups, x, z, y, s);
}

// The layout the block kernel reads (the plan's block fields).
static TriLayout block_layout(const TriLayout& L) {
  TriLayout b = L;
  b.rent = L.rent_blk;
  b.ring = L.ring_blk;
  b.ep_far = L.far_blk;
  b.rp_farq = L.farq_blk;
  b.rp_e = L.e_blk;
  b.rp_lanes = L.lanes_blk;
  b.rp_rows = L.rows_blk;
  b.rp_waves = L.waves_blk;
  return b;
}

bool trsv_rp_blk_supported(const TriLayout& Lin) {
  const TriLayout L = block_layout(Lin);
  if (!L.rent || L.rp_waves < 1 || L.rp_waves > 8) return false;
  if (rpb_lds_bytes(L.ring, L.ep_maxlev) > kTriBlockedMaxLds) return false;
  if (L.rp_rows != 1 && L.rp_rows != 2) return false;
  if (L.rp_lanes <= 1)
    return (L.rp_e == 4 || L.rp_e == 8 || L.rp_e == 10 || L.rp_e == 12) &&
           (!L.ep_far || L.rp_farq <= (L.rp_rows == 1 ? L.rp_e : L.rp_e == 4 ? 1 : 2));
  return (!L.ep_far ||
          (L.rp_rows == 1 && L.rp_farq <= (L.rp_e >= 10 ? 8 : 4) * L.rp_lanes)) &&
         (L.rp_e == 8 || ((L.rp_e == 10 || L.rp_e == 12) && L.rp_rows == 1)) &&
         (L.rp_lanes == 2 || L.rp_lanes == 4 || L.rp_lanes == 8 || L.rp_lanes == 16);
}

hipError_t launch_trsv_rp_blk(const TriLayout& Lin, const TriSegment& seg, int groups, int bit,
                              int dbit, const void* in, void* out, hipStream_t s) {
  if (!trsv_rp_blk_supported(Lin)) return hipErrorInvalidValue;
  const TriLayout L = block_layout(Lin);
  const int4* vin = static_cast<const int4*>(in);
  int4* vout = static_cast<int4*>(out);
  if (L.ep_far) return rpb_layout<true>(L, seg, groups, bit, dbit, vin, vout, s);
  return rpb_layout<false>(L, seg, groups, bit, dbit, vin, vout, s);
}

hipError_t launch_trsv_tiles_blk(TriKind kind, int n, int nb, int groups, const int32_t* tiles,
                                 const long long* toff, const int32_t* dinv, int bit, int dbit,
                                 const void* in, void* out, unsigned long long* xp, int epoch,
                                 int* err, unsigned* tickets, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  auto go = [&](auto kern) -> hipError_t {
    hipLaunchKernelGGL(kern, dim3(nb * groups), dim3(kTileThreads), 0, s, n, nb, groups, tiles,
                       toff, dinv, bit, dbit, static_cast<const int4*>(in),
                       static_cast<int4*>(out), xp, epoch, err, tickets);
    return hipGetLastError();
  };
  if (kind == TriKind::kL) return go(&k_tri_tiles_blk<TriKind::kL>);
  if (kind == TriKind::kDU) return go(&k_tri_tiles_blk<TriKind::kDU>);
  return go(&k_tri_tiles_blk<TriKind::kLD>);
}

}  // namespace sc23ad::dev
