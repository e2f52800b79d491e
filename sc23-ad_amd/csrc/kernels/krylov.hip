// Krylov kernels of Restarted(FGMRES(m)) for the three precisions.
//
// Integer solver: FGMRESflex::Solve (Library/solver/gmres.hpp:225-271) with
// blas1::int32 (Library/blas/blas1.hpp:106-119,230-307) and
// blas2::int32::Trsv (Library/blas/blas2.hpp:51-61).
// FP solvers: FGMRES<double|float>::Solve (gmres.hpp:48-88,130-170).
// Outer loop: Restarted::Solve (Library/solver/irsolver.hpp:27-43).
//
// Modified Gram-Schmidt keeps the reference's order: each pass updates w with
// V_k and then accumulates the next inner product <V_{k+1}, w> (or the norm
// of w) from the updated values, so one pass per k replaces the reference's
// separate Axpy and Dot.  Scalar work (lsqrt, H_inv, Givens, e, the
// convergence test) runs in the last workgroup of the reduction, so the
// Arnoldi step never waits for the host except for the convergence flag.
#include "dev_common.hpp"
#include "launch.hpp"

namespace sc23ad::dev {

namespace {

constexpr int B = kRedThreads;

// lsqrt (blas1.hpp:106-119): floor(sqrt(x)) for x > 0, 0 otherwise.  The
// reference's Newton iteration from a power-of-two overestimate ends at
// exactly that floor; here the FP64 square root gives an estimate within a
// few units and the two loops correct it to the floor (squares in unsigned
// 64 bits: s + 1 < 2^32 for x < 2^63), without the iteration's chain of
// 64-bit divisions.
__device__ __forceinline__ int64_t lsqrt64(int64_t x) {
  if (x <= 0) return 0;
  const uint64_t ux = static_cast<uint64_t>(x);
  uint64_t s = static_cast<uint64_t>(sqrt(static_cast<double>(x)));
  while (s * s > ux) --s;
  while ((s + 1) * (s + 1) <= ux) ++s;
  return static_cast<int64_t>(s);
}

// Tree over the kRedBlocks partials in the last block (256 threads, 4 values
// each): identical to the halving tree over 1024 values.
template <typename T>
__device__ __forceinline__ T final_tree(const T* part, T* sh) {
  const int t = threadIdx.x;
  const T a = part[t] + part[t + 512];
  const T b = part[t + 256] + part[t + 768];
  return block_tree<B, T>(a + b, sh);
}

// Thread t of virtual block vb sums f(i) over i = vb*B + t + q*kRedTotal in
// order; the block's halving tree gives part[vb].
template <typename T, typename F>
__device__ __forceinline__ void dro_partials(int n, T* part, T* sh, F&& f) {
  for (int vb = blockIdx.x; vb < kRedBlocks; vb += gridDim.x) {
    if (vb * B >= n) {
      if (threadIdx.x == 0) part[vb] = 0;
      continue;
    }
    T acc = 0;
    for (int i = vb * B + threadIdx.x; i < n; i += kRedTotal) acc += f(i);
    const T p = block_tree<B, T>(acc, sh);
    if (threadIdx.x == 0) part[vb] = p;
  }
}

__device__ __forceinline__ void start_scalars(const KState& k, Elem e,
                                              double nrm_r) {
  const double prev = k.scal[1];
  k.flags[2] = (fabs(nrm_r - prev) / nrm_r < 1.0e-8) ? 1 : 0;
  k.scal[0] = nrm_r;
  k.scal[1] = nrm_r;
  k.flags[0] = 0;
  k.flags[1] = 0;
  k.flags[3] = 0;
  if (e == Elem::kI32) {
    const double e0 = nrm_r;
    const double e_inv = 1.0 / e0;
    const double fact = static_cast<double>(static_cast<int64_t>(1) << k.bit);
    k.scal[2] = e0;
    k.scal[3] = e_inv;
    k.iscal[0] = trunc_i64(k.nrm_b * k.eps * e_inv * fact);
    static_cast<int64_t*>(k.e)[0] = static_cast<int64_t>(1) << k.bit;
  } else if (e == Elem::kF64) {
    static_cast<double*>(k.e)[0] = nrm_r;
    k.scal[2] = nrm_r;
    k.scal[3] = 1.0 / nrm_r;
  } else {
    const float e0 = static_cast<float>(nrm_r);
    static_cast<float*>(k.e)[0] = e0;
    k.scal[2] = static_cast<double>(e0);
    k.scal[3] = 1.0 / static_cast<double>(e0);
  }
}

// r = b - A x with each row summed in column order, plus the blocked-order
// partial sums of r_i^2.
__global__ __launch_bounds__(B) void k_outer_residual(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const double* __restrict__ val, const double* __restrict__ b,
    const double* __restrict__ x, double* __restrict__ r, double* __restrict__ part) {
  __shared__ double sh[B];
  dro_partials<double>(n, part, sh, [&](int i) {
    double t = 0;
    const int re = rptr[i + 1];
    for (int p = rptr[i]; p < re; ++p) t += val[p] * x[cind[p]];
    const double ri = b[i] - t;
    r[i] = ri;
    return ri * ri;
  });
}

// norm = sqrt(tree over the partials); optionally the Restarted / FGMRES
// start scalars.
__global__ __launch_bounds__(B) void k_norm_finalize(const double* __restrict__ part,
                                                     KState k, Elem e, int with_start,
                                                     double* out_norm) {
  __shared__ double sh[B];
  const double tot = final_tree<double>(part, sh);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(tot);
    if (with_start) start_scalars(k, e, nrm);
    if (out_norm) *out_norm = nrm;
  }
}

// Residual only (no reduction), for the strictly sequential norm mode.
__global__ __launch_bounds__(B) void k_residual_only(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const double* __restrict__ val, const double* __restrict__ b,
    const double* __restrict__ x, double* __restrict__ r) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  double t = 0;
  const int re = rptr[i + 1];
  for (int p = rptr[i]; p < re; ++p) t += val[p] * x[cind[p]];
  r[i] = b[i] - t;
}

// Left-to-right sum of squares by one thread (reference order; small n only).
__global__ void k_seq_norm(int n, const double* __restrict__ r, KState k, Elem e,
                           int with_start, double* out_norm) {
  double s = 0;
  for (int i = 0; i < n; ++i) s += r[i] * r[i];
  const double nrm = sqrt(s);
  if (with_start) start_scalars(k, e, nrm);
  if (out_norm) *out_norm = nrm;
}

__global__ __launch_bounds__(B) void k_v0(int n, const double* __restrict__ r,
                                          void* v0, KState k, Elem e) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  const double e_inv = k.scal[3];
  if (e == Elem::kI32) {
    const double fact = static_cast<double>(static_cast<int64_t>(1) << k.bit);
    static_cast<int32_t*>(v0)[i] = trunc_i32(r[i] * e_inv * fact);
  } else if (e == Elem::kF64) {
    static_cast<double*>(v0)[i] = r[i] * e_inv;
  } else {
    static_cast<float*>(v0)[i] = static_cast<float>(r[i] * e_inv);
  }
}

// ---- partial sums -----------------------------------------------------------
//
// Reductions write one partial per (virtual) block and never use atomics: the
// next kernel in the stream re-derives the scalar from the partials in every
// workgroup (the same tree in each, so the same value), which avoids
// device-wide fences and contended atomics across the XCDs.
//
// Integer partials: G physical blocks, grid-stride, order-free sum.
// FP partials: the fixed 1024 x 256 layout; physical block pb computes the
// virtual blocks pb, pb + G, ... exactly as that layout defines them.

__device__ __forceinline__ uint64_t sum_parts(const int64_t* part, int G, uint64_t* sh) {
  uint64_t s = 0;
  for (int b = threadIdx.x; b < G; b += B) s += static_cast<uint64_t>(part[b]);
  s = block_sum_u64<B>(s, sh);
  return s;  // valid in thread 0
}

// ---- integer MGS ---------------------------------------------------------

// Element loops of the integer MGS kernels: four consecutive elements per
// step with 16-byte loads when every vector starts on a 16-byte boundary
// (n % 4 == 0; the Krylov block is one allocation), else one at a time.
// Integer sums are order-free, so the split does not change a bit.
template <typename F>
__device__ __forceinline__ void int_elems(int n, F&& f) {
  const int stride = gridDim.x * B;
  const int t = blockIdx.x * B + threadIdx.x;
  if ((n & 3) == 0) {
    // two steps at a time, both steps' loads issued before either store
    // (the stores may alias later loads as far as the compiler knows, so
    // one step at a time leaves a single step's loads in flight)
    const int nq = n >> 2;
    int q = t;
    for (; q + stride < nq; q += 2 * stride) f.quad2(q, q + stride);
    if (q < nq) f.quad(q);
  } else {
    for (int i = t; i < n; i += stride) f.one(i);
  }
}

__global__ __launch_bounds__(B) void k_dot_first_int(int n, const int32_t* __restrict__ v,
                                                     const int32_t* __restrict__ w,
                                                     int64_t* __restrict__ part) {
  __shared__ uint64_t sh[B / 64];
  uint64_t acc = 0;
  struct {
    const int32_t* v;
    const int32_t* w;
    uint64_t* acc;
    __device__ void one(int i) { *acc += prod64(v[i], w[i]); }
    __device__ void quad(int q) {
      const int4 a = reinterpret_cast<const int4*>(v)[q];
      const int4 b = reinterpret_cast<const int4*>(w)[q];
      *acc += prod64(a.x, b.x) + prod64(a.y, b.y) + prod64(a.z, b.z) + prod64(a.w, b.w);
    }
    __device__ void quad2(int q0, int q1) {
      const int4 a0 = reinterpret_cast<const int4*>(v)[q0];
      const int4 b0 = reinterpret_cast<const int4*>(w)[q0];
      const int4 a1 = reinterpret_cast<const int4*>(v)[q1];
      const int4 b1 = reinterpret_cast<const int4*>(w)[q1];
      *acc += prod64(a0.x, b0.x) + prod64(a0.y, b0.y) + prod64(a0.z, b0.z) + prod64(a0.w, b0.w);
      *acc += prod64(a1.x, b1.x) + prod64(a1.y, b1.y) + prod64(a1.z, b1.z) + prod64(a1.w, b1.w);
    }
  } body{v, w, &acc};
  int_elems(n, body);
  const uint64_t p = block_sum_u64<B>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = static_cast<int64_t>(p);
}

__device__ void givens_int(const KState& k, int j, int64_t hn) {
  int64_t* H = static_cast<int64_t*>(k.H) + j * (k.m + 1);
  int64_t* c = static_cast<int64_t*>(k.c);
  int64_t* s = static_cast<int64_t*>(k.s);
  int64_t* e = static_cast<int64_t*>(k.e);
  const int bit = k.bit;
  H[j + 1] = hn;
  k.iscal[2] = hn;
  k.flags[3] = j + 1;
  if (hn == 0) {  // the reference divides by zero here
    k.flags[1] = 1;
    return;
  }
  for (int q = 0; q < j; ++q) {  // Grot
    const int64_t t = H[q];
    H[q] = as_i64(static_cast<uint64_t>(mul_wrap(c[q], t)) -
                  static_cast<uint64_t>(mul_wrap(s[q], H[q + 1]))) >> bit;
    H[q + 1] = as_i64(static_cast<uint64_t>(mul_wrap(s[q], t)) +
                      static_cast<uint64_t>(mul_wrap(c[q], H[q + 1]))) >> bit;
  }
  {  // Ggen
    const int64_t a = H[j], b = H[j + 1];
    int64_t r = lsqrt64(as_i64(static_cast<uint64_t>(mul_wrap(a, a)) +
                               static_cast<uint64_t>(mul_wrap(b, b))));
    if (r == 0) {
      k.flags[1] = 1;
      return;
    }
    const int64_t inv = ((static_cast<int64_t>(1) << 62) / r) >> (62 - bit - bit);
    c[j] = mul_wrap(a, inv) >> bit;
    s[j] = as_i64(0 - static_cast<uint64_t>(mul_wrap(b, inv))) >> bit;
    H[j] = r;
  }
  H[j + 1] = 0;
  e[j + 1] = mul_wrap(s[j], e[j]) >> bit;
  e[j] = mul_wrap(c[j], e[j]) >> bit;
  const int64_t ej = e[j + 1];
  const int64_t ae = ej < 0 ? as_i64(0 - static_cast<uint64_t>(ej)) : ej;
  k.flags[0] = (ae <= k.iscal[0]) ? 1 : 0;
  const double fact = static_cast<double>(static_cast<int64_t>(1) << bit);
  k.scal[5] = fabs(k.scal[2] * static_cast<double>(ej) / fact) / k.nrm_b;
}

// w += (int32)((-H(kk,j) * V_kk) >> F), then the partials of <V_{kk+1}, w>
// (kk < j) or of <w, w> (kk == j).  H(kk,j) is finished here from the
// previous pass's partials.
__global__ __launch_bounds__(B) void k_axpy_dot_int(int n, const int32_t* __restrict__ vk,
                                                    const int32_t* __restrict__ vk1,
                                                    int32_t* __restrict__ w, KState k,
                                                    int kk, int j,
                                                    const int64_t* __restrict__ pin,
                                                    int64_t* __restrict__ pout) {
  __shared__ uint64_t sh[B / 64];
  __shared__ int64_t hsh;
  const bool norm = (kk == j);
  // the operands of the thread's first two steps do not depend on H(kk,j):
  // their loads are issued before the previous pass's partial sums are
  // read (indices clamped into the vectors; a step past the end is skipped)
  const int nq = n >> 2, qstride = gridDim.x * B, qt = blockIdx.x * B + threadIdx.x;
  const bool vec = (n & 3) == 0 && nq > 0;
  int4 pv[2], px[2], pu[2];
  if (vec) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int q = min(qt + r * qstride, nq - 1);
      pv[r] = reinterpret_cast<const int4*>(vk)[q];
      px[r] = reinterpret_cast<const int4*>(w)[q];
      pu[r] = norm ? make_int4(0, 0, 0, 0) : reinterpret_cast<const int4*>(vk1)[q];
    }
  }
  const uint64_t tot = sum_parts(pin, gridDim.x, sh);
  if (threadIdx.x == 0) {
    const int64_t h = as_i64(tot) >> k.bit;
    hsh = h;
    if (blockIdx.x == 0) static_cast<int64_t*>(k.H)[j * (k.m + 1) + kk] = h;
  }
  __syncthreads();
  const int64_t a = as_i64(0 - static_cast<uint64_t>(hsh));
  const int bit = k.bit;
  uint64_t acc = 0;
  struct {
    const int32_t* vk;
    const int32_t* vk1;
    int32_t* w;
    int64_t a;
    int bit;
    bool norm;
    uint64_t* acc;
    __device__ int32_t upd(int32_t wi, int32_t v) const {
      return add32(wi, low32(mul_wrap(a, v) >> bit));
    }
    __device__ void one(int i) {
      const int32_t wi = upd(w[i], vk[i]);
      w[i] = wi;
      *acc += norm ? prod64(wi, wi) : prod64(vk1[i], wi);
    }
    __device__ void quad(int q) {
      const int4 v = reinterpret_cast<const int4*>(vk)[q];
      int4 x = reinterpret_cast<const int4*>(w)[q];
      x.x = upd(x.x, v.x);
      x.y = upd(x.y, v.y);
      x.z = upd(x.z, v.z);
      x.w = upd(x.w, v.w);
      reinterpret_cast<int4*>(w)[q] = x;
      if (norm) {
        *acc += prod64(x.x, x.x) + prod64(x.y, x.y) + prod64(x.z, x.z) + prod64(x.w, x.w);
      } else {
        const int4 u = reinterpret_cast<const int4*>(vk1)[q];
        *acc += prod64(u.x, x.x) + prod64(u.y, x.y) + prod64(u.z, x.z) + prod64(u.w, x.w);
      }
    }
    __device__ void quad2(int q0, int q1) {
      const int4 v0 = reinterpret_cast<const int4*>(vk)[q0];
      const int4 v1 = reinterpret_cast<const int4*>(vk)[q1];
      int4 x0 = reinterpret_cast<const int4*>(w)[q0];
      int4 x1 = reinterpret_cast<const int4*>(w)[q1];
      int4 u0 = make_int4(0, 0, 0, 0), u1 = u0;
      if (!norm) {
        u0 = reinterpret_cast<const int4*>(vk1)[q0];
        u1 = reinterpret_cast<const int4*>(vk1)[q1];
      }
      x0.x = upd(x0.x, v0.x);
      x0.y = upd(x0.y, v0.y);
      x0.z = upd(x0.z, v0.z);
      x0.w = upd(x0.w, v0.w);
      x1.x = upd(x1.x, v1.x);
      x1.y = upd(x1.y, v1.y);
      x1.z = upd(x1.z, v1.z);
      x1.w = upd(x1.w, v1.w);
      reinterpret_cast<int4*>(w)[q0] = x0;
      reinterpret_cast<int4*>(w)[q1] = x1;
      const int4 p0 = norm ? x0 : u0, p1 = norm ? x1 : u1;
      *acc += prod64(p0.x, x0.x) + prod64(p0.y, x0.y) + prod64(p0.z, x0.z) + prod64(p0.w, x0.w);
      *acc += prod64(p1.x, x1.x) + prod64(p1.y, x1.y) + prod64(p1.z, x1.z) + prod64(p1.w, x1.w);
    }
    // one step on operands loaded beforehand
    __device__ void step(int q, int4 v, int4 x, int4 u) {
      x.x = upd(x.x, v.x);
      x.y = upd(x.y, v.y);
      x.z = upd(x.z, v.z);
      x.w = upd(x.w, v.w);
      reinterpret_cast<int4*>(w)[q] = x;
      const int4 p = norm ? x : u;
      *acc += prod64(p.x, x.x) + prod64(p.y, x.y) + prod64(p.z, x.z) + prod64(p.w, x.w);
    }
  } body{vk, vk1, w, a, bit, norm, &acc};
  if (vec) {
#pragma unroll
    for (int r = 0; r < 2; ++r)
      if (qt + r * qstride < nq) body.step(qt + r * qstride, pv[r], px[r], pu[r]);
    int q = qt + 2 * qstride;
    for (; q + qstride < nq; q += 2 * qstride) body.quad2(q, q + qstride);
    if (q < nq) body.quad(q);
  } else {
    int_elems(n, body);
  }
  const uint64_t p = block_sum_u64<B>(acc, sh);
  if (threadIdx.x == 0) pout[blockIdx.x] = static_cast<int64_t>(p);
}

// w = (int32)((H_inv w) >> F); block 0 also finishes the Arnoldi step:
// h = lsqrt(<w,w>), Givens rotations, e, convergence flag.
__global__ __launch_bounds__(B) void k_scal_givens_int(int n, int32_t* __restrict__ w,
                                                       KState k, int j,
                                                       const int64_t* __restrict__ pin) {
  __shared__ uint64_t sh[B / 64];
  __shared__ int64_t hinv_sh;
  // the thread's first two words of w are loaded before the partial sums
  // are read and the Givens step is done (indices clamped into the vector)
  const int nq = n >> 2, qstride = gridDim.x * B, qt = blockIdx.x * B + threadIdx.x;
  const bool vec = (n & 3) == 0 && nq > 0;
  int4 px[2];
  if (vec) {
#pragma unroll
    for (int r = 0; r < 2; ++r) px[r] = reinterpret_cast<const int4*>(w)[min(qt + r * qstride, nq - 1)];
  }
  const uint64_t tot = sum_parts(pin, gridDim.x, sh);
  if (threadIdx.x == 0) {
    const int64_t hn = lsqrt64(as_i64(tot));
    const int bit = k.bit;
    hinv_sh = hn == 0 ? 0 : ((static_cast<int64_t>(1) << 62) / hn) >> (62 - bit - bit);
    if (blockIdx.x == 0) {
      k.iscal[1] = hinv_sh;
      givens_int(k, j, hn);
    }
  }
  __syncthreads();
  struct {
    int32_t* w;
    int64_t hinv;
    int bit;
    __device__ void one(int i) { w[i] = low32(mul_wrap(hinv, w[i]) >> bit); }
    __device__ void quad(int q) {
      int4 x = reinterpret_cast<int4*>(w)[q];
      x.x = low32(mul_wrap(hinv, x.x) >> bit);
      x.y = low32(mul_wrap(hinv, x.y) >> bit);
      x.z = low32(mul_wrap(hinv, x.z) >> bit);
      x.w = low32(mul_wrap(hinv, x.w) >> bit);
      reinterpret_cast<int4*>(w)[q] = x;
    }
    __device__ void quad2(int q0, int q1) {
      int4 x0 = reinterpret_cast<int4*>(w)[q0];
      int4 x1 = reinterpret_cast<int4*>(w)[q1];
      x0.x = low32(mul_wrap(hinv, x0.x) >> bit);
      x0.y = low32(mul_wrap(hinv, x0.y) >> bit);
      x0.z = low32(mul_wrap(hinv, x0.z) >> bit);
      x0.w = low32(mul_wrap(hinv, x0.w) >> bit);
      x1.x = low32(mul_wrap(hinv, x1.x) >> bit);
      x1.y = low32(mul_wrap(hinv, x1.y) >> bit);
      x1.z = low32(mul_wrap(hinv, x1.z) >> bit);
      x1.w = low32(mul_wrap(hinv, x1.w) >> bit);
      reinterpret_cast<int4*>(w)[q0] = x0;
      reinterpret_cast<int4*>(w)[q1] = x1;
    }
    __device__ void step(int q, int4 x) {
      x.x = low32(mul_wrap(hinv, x.x) >> bit);
      x.y = low32(mul_wrap(hinv, x.y) >> bit);
      x.z = low32(mul_wrap(hinv, x.z) >> bit);
      x.w = low32(mul_wrap(hinv, x.w) >> bit);
      reinterpret_cast<int4*>(w)[q] = x;
    }
  } body{w, hinv_sh, k.bit};
  if (vec) {
#pragma unroll
    for (int r = 0; r < 2; ++r)
      if (qt + r * qstride < nq) body.step(qt + r * qstride, px[r]);
    int q = qt + 2 * qstride;
    for (; q + qstride < nq; q += 2 * qstride) body.quad2(q, q + qstride);
    if (q < nq) body.quad(q);
  } else {
    int_elems(n, body);
  }
}

// ---- FP MGS (fixed blocked order) --------------------------------------------

template <typename T>
__global__ __launch_bounds__(B) void k_dot_first_fp(int n, const T* __restrict__ v,
                                                    const T* __restrict__ w,
                                                    T* __restrict__ part) {
  __shared__ T sh[B];
  dro_partials<T>(n, part, sh, [&](int i) { return v[i] * w[i]; });
}

template <typename T>
__device__ __forceinline__ T fp_sqrt(T v);
template <>
__device__ __forceinline__ double fp_sqrt<double>(double v) { return sqrt(v); }
template <>
__device__ __forceinline__ float fp_sqrt<float>(float v) { return sqrtf(v); }

template <typename T>
__device__ void givens_fp(const KState& k, int j, T hn) {
  T* H = static_cast<T*>(k.H) + j * (k.m + 1);
  T* c = static_cast<T*>(k.c);
  T* s = static_cast<T*>(k.s);
  T* e = static_cast<T*>(k.e);
  H[j + 1] = hn;
  k.flags[3] = j + 1;
  for (int q = 0; q < j; ++q) {
    const T t = H[q];
    H[q] = c[q] * t - s[q] * H[q + 1];
    H[q + 1] = s[q] * t + c[q] * H[q + 1];
  }
  {
    const T a = H[j], b = H[j + 1];
    const T r = fp_sqrt<T>(a * a + b * b);
    c[j] = a / r;
    s[j] = -b / r;
    H[j] = r;
  }
  H[j + 1] = 0;
  e[j + 1] = s[j] * e[j];
  e[j] = c[j] * e[j];
  const double ae = fabs(static_cast<double>(e[j + 1]));
  k.flags[0] = (ae <= k.nrm_b * k.eps) ? 1 : 0;
  k.scal[5] = ae / k.nrm_b;
}

template <typename T>
__global__ __launch_bounds__(B) void k_axpy_dot_fp(int n, const T* __restrict__ vk,
                                                   const T* __restrict__ vk1,
                                                   T* __restrict__ w, KState k, int kk,
                                                   int j, const T* __restrict__ pin,
                                                   T* __restrict__ pout) {
  __shared__ T sh[B];
  __shared__ T hsh;
  const T h = final_tree<T>(pin, sh);
  if (threadIdx.x == 0) {
    hsh = h;
    if (blockIdx.x == 0) static_cast<T*>(k.H)[j * (k.m + 1) + kk] = h;
  }
  __syncthreads();
  const T a = -hsh;
  const bool norm = (kk == j);
  // every element is updated exactly once, by the thread that also adds its
  // term to the blocked partial sums
  dro_partials<T>(n, pout, sh, [&](int i) {
    const T wi = w[i] + a * vk[i];
    w[i] = wi;
    return norm ? wi * wi : vk1[i] * wi;
  });
}

template <typename T>
__global__ __launch_bounds__(B) void k_scal_givens_fp(int n, T* __restrict__ w, KState k,
                                                      int j, const T* __restrict__ pin) {
  __shared__ T sh[B];
  __shared__ T sc_sh;
  const T tot = final_tree<T>(pin, sh);
  if (threadIdx.x == 0) {
    const T hn = fp_sqrt<T>(tot);
    sc_sh = static_cast<T>(1) / hn;
    if (blockIdx.x == 0) {
      k.scal[4] = static_cast<double>(sc_sh);
      givens_fp<T>(k, j, hn);
    }
  }
  __syncthreads();
  const T sc = sc_sh;
  const int stride = gridDim.x * B;
  for (int i = blockIdx.x * B + threadIdx.x; i < n; i += stride) w[i] *= sc;
}

// Sequential-order variants: one thread computes each reduction left to right.
template <typename T>
__global__ void k_seq_dot_fp(int n, const T* __restrict__ v, const T* __restrict__ w,
                             KState k, int slot_j, int slot_k) {
  T acc = 0;
  for (int i = 0; i < n; ++i) acc += v[i] * w[i];
  static_cast<T*>(k.H)[slot_j * (k.m + 1) + slot_k] = acc;
}
template <typename T>
__global__ __launch_bounds__(B) void k_axpy_fp(int n, const T* __restrict__ vk,
                                               T* __restrict__ w, KState k, int kk,
                                               int j) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  const T a = -static_cast<T*>(k.H)[j * (k.m + 1) + kk];
  w[i] += a * vk[i];
}
template <typename T>
__global__ void k_seq_norm_givens_fp(int n, const T* __restrict__ w, KState k, int j) {
  T acc = 0;
  for (int i = 0; i < n; ++i) acc += w[i] * w[i];
  const T hn = fp_sqrt<T>(acc);
  k.scal[4] = static_cast<double>(static_cast<T>(1) / hn);
  givens_fp<T>(k, j, hn);
}
template <typename T>
__global__ __launch_bounds__(B) void k_scal_fp(int n, T* __restrict__ w, KState k) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  w[i] *= static_cast<T>(k.scal[4]);
}

// ---- solution update -----------------------------------------------------

__device__ void trsv_int_body(const KState& k, int steps) {
  const int64_t* H = static_cast<const int64_t*>(k.H);
  const int64_t* e = static_cast<const int64_t*>(k.e);
  const int nn = k.m + 1;
  const double fact = static_cast<double>(static_cast<int64_t>(1) << k.bit);
  const double e0 = k.scal[2];
  for (int i = steps - 1; i >= 0; --i) {
    double temp = e0 * static_cast<double>(e[i]) / fact;
    for (int q = steps - 1; q > i; --q)
      temp -= static_cast<double>(H[q * nn + i]) / fact * k.y[q];
    k.y[i] = temp / (static_cast<double>(H[i * nn + i]) / fact);
  }
}

__global__ void k_trsv_int(KState k, int steps) { trsv_int_body(k, steps); }

template <typename T>
__global__ void k_trsv_fp(KState k, int steps, T* y) {
  const T* H = static_cast<const T*>(k.H);
  const T* e = static_cast<const T*>(k.e);
  const int nn = k.m + 1;
  for (int i = steps - 1; i >= 0; --i) {
    T temp = e[i];
    for (int q = steps - 1; q > i; --q) temp -= H[q * nn + i] * y[q];
    y[i] = temp / H[i * nn + i];
  }
}

__global__ __launch_bounds__(B) void k_update_x_int(int n, const int32_t* __restrict__ Z,
                                                    double* __restrict__ x, KState k,
                                                    int steps) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  const double fact_inv = 1.0 / static_cast<double>(static_cast<int64_t>(1) << k.bit);
  double xi = x[i];
  for (int q = 0; q < steps; ++q)
    xi += k.y[q] * static_cast<double>(Z[static_cast<size_t>(q) * n + i]) * fact_inv;
  x[i] = xi;
}

__global__ __launch_bounds__(B) void k_update_x_f64(int n, const double* __restrict__ Z,
                                                    double* __restrict__ x, KState k,
                                                    int steps) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  double xi = x[i];
  for (int q = 0; q < steps; ++q) xi += k.y[q] * Z[static_cast<size_t>(q) * n + i];
  x[i] = xi;
}

__global__ __launch_bounds__(B) void k_update_x_f32(int n, const float* __restrict__ Z,
                                                    double* __restrict__ x, KState k,
                                                    int steps) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  double xi = x[i];
  for (int q = 0; q < steps; ++q)
    xi += static_cast<double>(k.yf[q] * Z[static_cast<size_t>(q) * n + i]);
  x[i] = xi;
}

// ---- checksum --------------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(B) void k_checksum_part(int n, const int32_t* __restrict__ v,
                                                     int64_t* __restrict__ part) {
  __shared__ uint64_t sh[B / 64];
  uint64_t acc = 0;
  const int stride = gridDim.x * B;
  for (int i = blockIdx.x * B + threadIdx.x; i < n; i += stride)
    acc += mix64(static_cast<uint64_t>(i) * 0x9E3779B97F4A7C15ULL ^
                 static_cast<uint64_t>(static_cast<uint32_t>(v[i])));
  const uint64_t p = block_sum_u64<B>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = static_cast<int64_t>(p);
}

__global__ __launch_bounds__(B) void k_checksum_finalize(const int64_t* __restrict__ part,
                                                         int G, KState k) {
  __shared__ uint64_t sh[B / 64];
  const uint64_t s = sum_parts(part, G, sh);
  if (threadIdx.x == 0) {
    const int len = *k.trace_len;
    if (len < k.trace_cap) k.trace[len] = s;
    *k.trace_len = len + 1;
  }
}

__global__ __launch_bounds__(B) void k_copy_bytes(size_t n4, const uint32_t* __restrict__ a,
                                                  uint32_t* __restrict__ b) {
  for (size_t i = blockIdx.x * static_cast<size_t>(B) + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * B)
    b[i] = a[i];
}

__global__ __launch_bounds__(B) void k_zero_bytes(size_t n4, uint32_t* __restrict__ b) {
  for (size_t i = blockIdx.x * static_cast<size_t>(B) + threadIdx.x; i < n4;
       i += static_cast<size_t>(gridDim.x) * B)
    b[i] = 0;
}

// ---- block Krylov kernels -----------------------------------------------------
//
// k integer solves on one hierarchy advance in lockstep (every solve that has
// not finished is at the same Arnoldi step j, since all start together and
// restart after exactly m steps); solve r keeps its own KState ks[r] (its F,
// H, Givens pairs, e, scalars and partial sums), and its Krylov vectors are
// right-hand side r of the block vectors V_j, Z_j (four solves per 16-byte
// word, see block.hip).  live[r] marks the solves of the current cycle; a
// solve drops out of the cycle at its converged / overflow flag.  Each
// solve's arithmetic is its single solve's, element by element and in the
// same order, so its results are identical.

__device__ __forceinline__ bool blk_live(const KState* ks, const int* live, int r) {
  return live[r] != 0 && ks[r].flags[0] == 0 && ks[r].flags[1] == 0;
}

// r_r = b_r - A x_r and the blocked-order partials of r_r^2 (blockIdx.y = r).
__global__ __launch_bounds__(B) void k_outer_residual_blk(
    int n, const int* __restrict__ rptr, const int* __restrict__ cind,
    const double* __restrict__ val, const double* __restrict__ b_all,
    const double* __restrict__ x_all, double* __restrict__ r_all, const KState* ks,
    const int* live) {
  __shared__ double sh[B];
  const int r = blockIdx.y;
  if (!live[r]) return;
  const size_t off = static_cast<size_t>(r) * n;
  const double* b = b_all + off;
  const double* x = x_all + off;
  double* rr = r_all + off;
  dro_partials<double>(n, ks[r].dpart, sh, [&](int i) {
    double t = 0;
    const int re = rptr[i + 1];
    for (int p = rptr[i]; p < re; ++p) t += val[p] * x[cind[p]];
    const double ri = b[i] - t;
    rr[i] = ri;
    return ri * ri;
  });
}

__global__ __launch_bounds__(B) void k_norm_start_blk(const KState* ks, const int* live) {
  __shared__ double sh[B];
  const int r = blockIdx.x;
  if (!live[r]) return;
  const KState k = ks[r];
  const double tot = final_tree<double>(k.dpart, sh);
  if (threadIdx.x == 0) start_scalars(k, Elem::kI32, sqrt(tot));
}

// V_0 of every live solve: trunc(r * e_inv * 2^F) (zero for the others).
__global__ __launch_bounds__(B) void k_v0_blk(int n, const double* __restrict__ r_all,
                                              int4* __restrict__ v0, const KState* ks,
                                              const int* live) {
  const int i = blockIdx.x * B + threadIdx.x;
  const int g = blockIdx.y;
  if (i >= n) return;
  int32_t q[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int r = 4 * g + c;
    q[c] = 0;
    if (live[r]) {
      const double e_inv = ks[r].scal[3];
      const double fact = static_cast<double>(static_cast<int64_t>(1) << ks[r].bit);
      q[c] = trunc_i32(r_all[static_cast<size_t>(r) * n + i] * e_inv * fact);
    }
  }
  v0[static_cast<size_t>(g) * n + i] = make_int4(q[0], q[1], q[2], q[3]);
}

// Sum of four order-free values over the block; thread 0 gets them.
__device__ __forceinline__ void block_sum4(uint64_t* acc, uint64_t (*sh)[B / 64]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = wave_sum_u64(acc[c]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) sh[c][wid] = acc[c];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint64_t t = 0;
      for (int w = 0; w < B / 64; ++w) t += sh[c][w];
      acc[c] = t;
    }
  __syncthreads();
}

__global__ __launch_bounds__(B) void k_dot_first_blk(int n, const int4* __restrict__ v_all,
                                                     const int4* __restrict__ w_all,
                                                     const KState* ks) {
  __shared__ uint64_t sh[4][B / 64];
  const int g = blockIdx.y;
  const int4* v = v_all + static_cast<size_t>(g) * n;
  const int4* w = w_all + static_cast<size_t>(g) * n;
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
    const int4 a = v[i], b = w[i];
    acc[0] += prod64(a.x, b.x);
    acc[1] += prod64(a.y, b.y);
    acc[2] += prod64(a.z, b.z);
    acc[3] += prod64(a.w, b.w);
  }
  block_sum4(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) ks[4 * g + c].ipart[blockIdx.x] = static_cast<int64_t>(acc[c]);
}

// w += (int32)((-H(kk,j) V_kk) >> F) for each solve of the group, then the
// partials of <V_kk+1, w> (kk < j) or <w, w> (kk == j); H(kk,j) is finished
// from the previous pass's partials (half kk & 1 of ipart, written to half
// (kk + 1) & 1).
__global__ __launch_bounds__(B) void k_axpy_dot_blk(int n, const int4* __restrict__ vk_all,
                                                    const int4* __restrict__ vk1_all,
                                                    int4* __restrict__ w_all, const KState* ks,
                                                    const int* live, int kk, int j) {
  __shared__ uint64_t sh[4][B / 64];
  __shared__ int64_t hsh[4];
  const int g = blockIdx.y;
  const int G = gridDim.x;
  const int pin = (kk & 1) * kRedBlocks, pout = ((kk + 1) & 1) * kRedBlocks;
  {
    uint64_t tot[4] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < G; b += B)
#pragma unroll
      for (int c = 0; c < 4; ++c) tot[c] += static_cast<uint64_t>(ks[4 * g + c].ipart[pin + b]);
    block_sum4(tot, sh);
    if (threadIdx.x == 0)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int r = 4 * g + c;
        const int64_t h = as_i64(tot[c]) >> ks[r].bit;
        hsh[c] = h;
        if (blockIdx.x == 0 && blk_live(ks, live, r))
          static_cast<int64_t*>(ks[r].H)[j * (ks[r].m + 1) + kk] = h;
      }
  }
  __syncthreads();
  int64_t a[4];
  int bit[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    a[c] = as_i64(0 - static_cast<uint64_t>(hsh[c]));
    bit[c] = ks[4 * g + c].bit;
  }
  const bool norm = (kk == j);
  const size_t off = static_cast<size_t>(g) * n;
  const int4* vk = vk_all + off;
  const int4* vk1 = vk1_all + off;
  int4* w = w_all + off;
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int i = blockIdx.x * B + threadIdx.x; i < n; i += G * B) {
    const int4 v = vk[i];
    int4 x = w[i];
    x.x = add32(x.x, low32(mul_wrap(a[0], v.x) >> bit[0]));
    x.y = add32(x.y, low32(mul_wrap(a[1], v.y) >> bit[1]));
    x.z = add32(x.z, low32(mul_wrap(a[2], v.z) >> bit[2]));
    x.w = add32(x.w, low32(mul_wrap(a[3], v.w) >> bit[3]));
    w[i] = x;
    const int4 u = norm ? x : vk1[i];
    acc[0] += prod64(u.x, x.x);
    acc[1] += prod64(u.y, x.y);
    acc[2] += prod64(u.z, x.z);
    acc[3] += prod64(u.w, x.w);
  }
  block_sum4(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      ks[4 * g + c].ipart[pout + blockIdx.x] = static_cast<int64_t>(acc[c]);
}

// w = (int32)((H_inv w) >> F) per solve; block 0 finishes the Arnoldi step of
// each live solve (lsqrt, H_inv, Givens, e, convergence flag).
__global__ __launch_bounds__(B) void k_scal_givens_blk(int n, int4* __restrict__ w_all,
                                                       const KState* ks, const int* live,
                                                       int j) {
  __shared__ uint64_t sh[4][B / 64];
  __shared__ int64_t hinv_sh[4];
  const int g = blockIdx.y;
  const int G = gridDim.x;
  const int pin = ((j + 1) & 1) * kRedBlocks;
  uint64_t tot[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < G; b += B)
#pragma unroll
    for (int c = 0; c < 4; ++c) tot[c] += static_cast<uint64_t>(ks[4 * g + c].ipart[pin + b]);
  block_sum4(tot, sh);
  // lanes 0-3 of the first wave: one right-hand side each, its total from
  // the per-wave sums block_sum4 leaves in sh
  if (threadIdx.x < 4) {
    const int c = threadIdx.x;
    uint64_t tc = 0;
    for (int w = 0; w < B / 64; ++w) tc += sh[c][w];
    const int r = 4 * g + c;
    const KState k = ks[r];
    const int64_t hn = lsqrt64(as_i64(tc));
    const int bit = k.bit;
    hinv_sh[c] = hn == 0 ? 0 : ((static_cast<int64_t>(1) << 62) / hn) >> (62 - bit - bit);
    if (blockIdx.x == 0 && blk_live(ks, live, r)) {
      k.iscal[1] = hinv_sh[c];
      givens_int(k, j, hn);
    }
  }
  __syncthreads();
  int64_t hinv[4];
  int bit[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    hinv[c] = hinv_sh[c];
    bit[c] = ks[4 * g + c].bit;
  }
  int4* w = w_all + static_cast<size_t>(g) * n;
  for (int i = blockIdx.x * B + threadIdx.x; i < n; i += G * B) {
    int4 x = w[i];
    x.x = low32(mul_wrap(hinv[0], x.x) >> bit[0]);
    x.y = low32(mul_wrap(hinv[1], x.y) >> bit[1]);
    x.z = low32(mul_wrap(hinv[2], x.z) >> bit[2]);
    x.w = low32(mul_wrap(hinv[3], x.w) >> bit[3]);
    w[i] = x;
  }
}

// The whole Gram-Schmidt step j of a small system in one launch, one
// workgroup per group of four solves: <V_0, w>, then for kk = 0..j the
// coefficient H(kk,j), w += (int32)((-H(kk,j) V_kk) >> F) and the sums for the
// next coefficient (or <w, w>), then lsqrt, H_inv, the Givens update and
// w = (int32)((H_inv w) >> F) -- what k_dot_first_blk, k_axpy_dot_blk
// (kk = 0..j) and k_scal_givens_blk compute over many workgroups and j + 3
// launches.  Each thread keeps the same rows in every pass, so the updates
// of w need no barrier; the sums are integer and order-free, so every value
// is the multi-workgroup one.
constexpr int kMgsThreads = 1024;

__device__ __forceinline__ void mgs_sum4(uint64_t* acc, uint64_t (*sh)[kMgsThreads / 64],
                                         uint64_t* tot) {
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = wave_sum_u64(acc[c]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c) sh[c][wid] = acc[c];
  __syncthreads();
  if (threadIdx.x < 4) {
    uint64_t t = 0;
    for (int w = 0; w < kMgsThreads / 64; ++w) t += sh[threadIdx.x][w];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kMgsThreads) void k_mgs_step_blk(int n, size_t vs,
                                                              const int4* __restrict__ V,
                                                              int4* __restrict__ w_all,
                                                              const KState* ks, const int* live,
                                                              int j) {
  __shared__ uint64_t sh[4][kMgsThreads / 64];
  __shared__ uint64_t tot[4];
  __shared__ int64_t hsh[4];
  const int g = blockIdx.x;
  const size_t off = static_cast<size_t>(g) * n;
  int4* w = w_all + off;
  int bit[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) bit[c] = ks[4 * g + c].bit;
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int i = threadIdx.x; i < n; i += kMgsThreads) {
    const int4 a = V[off + i], b = w[i];
    acc[0] += prod64(a.x, b.x);
    acc[1] += prod64(a.y, b.y);
    acc[2] += prod64(a.z, b.z);
    acc[3] += prod64(a.w, b.w);
  }
  mgs_sum4(acc, sh, tot);
  for (int kk = 0; kk <= j; ++kk) {
    if (threadIdx.x < 4) {
      const int c = threadIdx.x, r = 4 * g + c;
      const int64_t h = as_i64(tot[c]) >> bit[c];
      hsh[c] = h;
      if (blk_live(ks, live, r)) static_cast<int64_t*>(ks[r].H)[j * (ks[r].m + 1) + kk] = h;
    }
    __syncthreads();
    int64_t a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = as_i64(0 - static_cast<uint64_t>(hsh[c]));
    const bool norm = (kk == j);
    const int4* vk = V + static_cast<size_t>(kk) * vs + off;
    const int4* vk1 = V + static_cast<size_t>(norm ? kk : kk + 1) * vs + off;
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = 0;
    for (int i = threadIdx.x; i < n; i += kMgsThreads) {
      const int4 v = vk[i];
      int4 x = w[i];
      x.x = add32(x.x, low32(mul_wrap(a[0], v.x) >> bit[0]));
      x.y = add32(x.y, low32(mul_wrap(a[1], v.y) >> bit[1]));
      x.z = add32(x.z, low32(mul_wrap(a[2], v.z) >> bit[2]));
      x.w = add32(x.w, low32(mul_wrap(a[3], v.w) >> bit[3]));
      w[i] = x;
      const int4 u = norm ? x : vk1[i];
      acc[0] += prod64(u.x, x.x);
      acc[1] += prod64(u.y, x.y);
      acc[2] += prod64(u.z, x.z);
      acc[3] += prod64(u.w, x.w);
    }
    mgs_sum4(acc, sh, tot);
  }
  if (threadIdx.x < 4) {
    const int c = threadIdx.x, r = 4 * g + c;
    const KState k = ks[r];
    const int64_t hn = lsqrt64(as_i64(tot[c]));
    hsh[c] = hn == 0 ? 0 : ((static_cast<int64_t>(1) << 62) / hn) >> (62 - bit[c] - bit[c]);
    if (blk_live(ks, live, r)) {
      k.iscal[1] = hsh[c];
      givens_int(k, j, hn);
    }
  }
  __syncthreads();
  int64_t hinv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) hinv[c] = hsh[c];
  for (int i = threadIdx.x; i < n; i += kMgsThreads) {
    int4 x = w[i];
    x.x = low32(mul_wrap(hinv[0], x.x) >> bit[0]);
    x.y = low32(mul_wrap(hinv[1], x.y) >> bit[1]);
    x.z = low32(mul_wrap(hinv[2], x.z) >> bit[2]);
    x.w = low32(mul_wrap(hinv[3], x.w) >> bit[3]);
    w[i] = x;
  }
}

// y = H^-1 e for every solve whose x is updated (upd[r]), over the steps its
// last cycle ran (flags[3]).
__global__ void k_trsv_blk(const KState* ks, const int* upd) {
  const int r = blockIdx.x;
  if (threadIdx.x != 0 || !upd[r]) return;
  const KState k = ks[r];
  trsv_int_body(k, k.flags[3]);
}

// x_r += sum_q y_q Z_q(r) 2^-F, in order of q (blockIdx.y = r).
__global__ __launch_bounds__(B) void k_update_x_blk(int n, int groups,
                                                    const int32_t* __restrict__ Z,
                                                    double* __restrict__ x_all,
                                                    const KState* ks, const int* upd) {
  const int r = blockIdx.y;
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n || !upd[r]) return;
  const KState& k = ks[r];
  const int steps = k.flags[3];
  const double fact_inv = 1.0 / static_cast<double>(static_cast<int64_t>(1) << k.bit);
  const size_t vstride = static_cast<size_t>(groups) * n * 4;
  const size_t at = (static_cast<size_t>(r >> 2) * n + i) * 4 + (r & 3);
  double* x = x_all + static_cast<size_t>(r) * n;
  double xi = x[i];
  for (int q = 0; q < steps; ++q)
    xi += k.y[q] * static_cast<double>(Z[q * vstride + at]) * fact_inv;
  x[i] = xi;
}

// Flags and implicit residual of every solve into one array (5 words each:
// converged, overflow, stagnated, steps, then the residual as a double in the
// next two words).
__global__ void k_flags_blk(const KState* ks, int count, int* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= count) return;
  const KState& k = ks[r];
  int* o = out + 6 * r;
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = k.flags[q];
  const double v = k.scal[5];
  __builtin_memcpy(o + 4, &v, sizeof v);
}

// Order-free checksums of every live solve's column of a block vector,
// appended to its trace (the single solve's checksum of the same vector).
__global__ __launch_bounds__(B) void k_checksum_blk_part(int n, const int32_t* __restrict__ v,
                                                         const KState* ks, const int* live) {
  __shared__ uint64_t sh[4][B / 64];
  const int g = blockIdx.y;
  uint64_t acc[4] = {0, 0, 0, 0};
  const int4* v4 = reinterpret_cast<const int4*>(v) + static_cast<size_t>(g) * n;
  for (int i = blockIdx.x * B + threadIdx.x; i < n; i += gridDim.x * B) {
    const int4 x = v4[i];
    const uint64_t base = static_cast<uint64_t>(i) * 0x9E3779B97F4A7C15ULL;
    acc[0] += mix64(base ^ static_cast<uint64_t>(static_cast<uint32_t>(x.x)));
    acc[1] += mix64(base ^ static_cast<uint64_t>(static_cast<uint32_t>(x.y)));
    acc[2] += mix64(base ^ static_cast<uint64_t>(static_cast<uint32_t>(x.z)));
    acc[3] += mix64(base ^ static_cast<uint64_t>(static_cast<uint32_t>(x.w)));
  }
  block_sum4(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      ks[4 * g + c].ipart[2 * kRedBlocks + blockIdx.x] = static_cast<int64_t>(acc[c]);
}

__global__ __launch_bounds__(B) void k_checksum_blk_finalize(const KState* ks, const int* live,
                                                             int G) {
  __shared__ uint64_t sh[B / 64];
  const int r = blockIdx.x;
  if (!blk_live(ks, live, r)) return;
  const KState k = ks[r];
  const uint64_t s = sum_parts(k.ipart + 2 * kRedBlocks, G, sh);
  if (threadIdx.x == 0 && k.trace) {
    const int len = *k.trace_len;
    if (len < k.trace_cap) k.trace[len] = s;
    *k.trace_len = len + 1;
  }
}

inline int grid_for(int n) { return (n + B - 1) / B; }

// Physical blocks for the reductions: about 8 elements per thread, at most
// kRedBlocks.
inline int red_grid(int n) {
  const int g = (n + 8 * B - 1) / (8 * B);
  return g < 1 ? 1 : (g > kRedBlocks ? kRedBlocks : g);
}

}  // namespace

hipError_t launch_outer_residual(const DCsr& a64, const double* b, const double* x,
                                 double* r, const KState& k, Elem e, int seq_red,
                                 hipStream_t s) {
  if (seq_red) {
    if (a64.n > 0)
      hipLaunchKernelGGL(k_residual_only, dim3(grid_for(a64.n)), dim3(B), 0, s, a64.n,
                         a64.rptr, a64.cind, static_cast<const double*>(a64.val), b,
                         x, r);
    hipLaunchKernelGGL(k_seq_norm, dim3(1), dim3(1), 0, s, a64.n, r, k, e, 1,
                       static_cast<double*>(nullptr));
  } else {
    hipLaunchKernelGGL(k_outer_residual, dim3(red_grid(a64.n)), dim3(B), 0, s, a64.n,
                       a64.rptr, a64.cind, static_cast<const double*>(a64.val), b, x,
                       r, k.dpart);
    hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(B), 0, s, k.dpart, k, e, 1,
                       static_cast<double*>(nullptr));
  }
  return hipGetLastError();
}

hipError_t launch_relres(const DCsr& a64, const double* b, const double* x,
                         double* r, double* part, unsigned* counter, double* out,
                         hipStream_t s) {
  (void)counter;
  KState k{};
  hipLaunchKernelGGL(k_outer_residual, dim3(red_grid(a64.n)), dim3(B), 0, s, a64.n,
                     a64.rptr, a64.cind, static_cast<const double*>(a64.val), b, x, r,
                     part);
  hipLaunchKernelGGL(k_norm_finalize, dim3(1), dim3(B), 0, s, part, k, Elem::kF64, 0, out);
  return hipGetLastError();
}

hipError_t launch_v0(Elem e, int n, const double* r, void* v0, const KState& k,
                     hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_v0, dim3(grid_for(n)), dim3(B), 0, s, n, r, v0, k, e);
  return hipGetLastError();
}

// Partial-sum buffers alternate between the two halves of ipart / dpart /
// fpart: pass q writes half (q & 1) and the next pass reads it.
template <typename T>
static T* half(T* base, int q) {
  return base + (q & 1) * kRedBlocks;
}

hipError_t launch_dot_first(Elem e, int n, const void* v0, const void* w,
                            const KState& k, int j, int seq_red, hipStream_t s) {
  const int G = red_grid(n);
  if (e == Elem::kI32) {
    hipLaunchKernelGGL(k_dot_first_int, dim3(G), dim3(B), 0, s, n,
                       static_cast<const int32_t*>(v0), static_cast<const int32_t*>(w),
                       half(k.ipart, 0));
  } else if (e == Elem::kF64) {
    if (seq_red)
      hipLaunchKernelGGL(k_seq_dot_fp<double>, dim3(1), dim3(1), 0, s, n,
                         static_cast<const double*>(v0), static_cast<const double*>(w),
                         k, j, 0);
    else
      hipLaunchKernelGGL(k_dot_first_fp<double>, dim3(G), dim3(B), 0, s, n,
                         static_cast<const double*>(v0), static_cast<const double*>(w),
                         half(k.dpart, 0));
  } else {
    if (seq_red)
      hipLaunchKernelGGL(k_seq_dot_fp<float>, dim3(1), dim3(1), 0, s, n,
                         static_cast<const float*>(v0), static_cast<const float*>(w), k,
                         j, 0);
    else
      hipLaunchKernelGGL(k_dot_first_fp<float>, dim3(G), dim3(B), 0, s, n,
                         static_cast<const float*>(v0), static_cast<const float*>(w),
                         half(k.fpart, 0));
  }
  return hipGetLastError();
}

hipError_t launch_axpy_dot(Elem e, int n, const void* V, void* w, const KState& k,
                           int kk, int j, int seq_red, hipStream_t s) {
  const size_t off = static_cast<size_t>(kk) * n;
  const size_t off1 = static_cast<size_t>(kk + 1) * n;
  const int G = red_grid(n);
  if (e == Elem::kI32) {
    const int32_t* v = static_cast<const int32_t*>(V);
    hipLaunchKernelGGL(k_axpy_dot_int, dim3(G), dim3(B), 0, s, n, v + off, v + off1,
                       static_cast<int32_t*>(w), k, kk, j, half(k.ipart, kk),
                       half(k.ipart, kk + 1));
    return hipGetLastError();
  }
  if (e == Elem::kF64) {
    const double* v = static_cast<const double*>(V);
    double* ww = static_cast<double*>(w);
    if (seq_red) {
      hipLaunchKernelGGL(k_axpy_fp<double>, dim3(grid_for(n)), dim3(B), 0, s, n, v + off,
                         ww, k, kk, j);
      if (kk < j)
        hipLaunchKernelGGL(k_seq_dot_fp<double>, dim3(1), dim3(1), 0, s, n, v + off1, ww,
                           k, j, kk + 1);
      else
        hipLaunchKernelGGL(k_seq_norm_givens_fp<double>, dim3(1), dim3(1), 0, s, n, ww,
                           k, j);
    } else {
      hipLaunchKernelGGL(k_axpy_dot_fp<double>, dim3(G), dim3(B), 0, s, n, v + off,
                         v + off1, ww, k, kk, j, half(k.dpart, kk), half(k.dpart, kk + 1));
    }
    return hipGetLastError();
  }
  const float* v = static_cast<const float*>(V);
  float* ww = static_cast<float*>(w);
  if (seq_red) {
    hipLaunchKernelGGL(k_axpy_fp<float>, dim3(grid_for(n)), dim3(B), 0, s, n, v + off,
                       ww, k, kk, j);
    if (kk < j)
      hipLaunchKernelGGL(k_seq_dot_fp<float>, dim3(1), dim3(1), 0, s, n, v + off1, ww, k,
                         j, kk + 1);
    else
      hipLaunchKernelGGL(k_seq_norm_givens_fp<float>, dim3(1), dim3(1), 0, s, n, ww, k,
                         j);
  } else {
    hipLaunchKernelGGL(k_axpy_dot_fp<float>, dim3(G), dim3(B), 0, s, n, v + off,
                       v + off1, ww, k, kk, j, half(k.fpart, kk), half(k.fpart, kk + 1));
  }
  return hipGetLastError();
}

hipError_t launch_scal(Elem e, int n, void* w, const KState& k, int j, int seq_red,
                       hipStream_t s) {
  const int G = red_grid(n);
  if (e == Elem::kI32) {
    hipLaunchKernelGGL(k_scal_givens_int, dim3(G), dim3(B), 0, s, n,
                       static_cast<int32_t*>(w), k, j, half(k.ipart, j + 1));
  } else if (e == Elem::kF64) {
    if (seq_red)
      hipLaunchKernelGGL(k_scal_fp<double>, dim3(grid_for(n)), dim3(B), 0, s, n,
                         static_cast<double*>(w), k);
    else
      hipLaunchKernelGGL(k_scal_givens_fp<double>, dim3(G), dim3(B), 0, s, n,
                         static_cast<double*>(w), k, j, half(k.dpart, j + 1));
  } else {
    if (seq_red)
      hipLaunchKernelGGL(k_scal_fp<float>, dim3(grid_for(n)), dim3(B), 0, s, n,
                         static_cast<float*>(w), k);
    else
      hipLaunchKernelGGL(k_scal_givens_fp<float>, dim3(G), dim3(B), 0, s, n,
                         static_cast<float*>(w), k, j, half(k.fpart, j + 1));
  }
  return hipGetLastError();
}

hipError_t launch_update_x(Elem e, int n, const void* Z, double* x, const KState& k,
                           int steps, hipStream_t s) {
  if (steps <= 0) return hipSuccess;
  if (e == Elem::kI32) {
    hipLaunchKernelGGL(k_trsv_int, dim3(1), dim3(1), 0, s, k, steps);
    if (n > 0)
      hipLaunchKernelGGL(k_update_x_int, dim3(grid_for(n)), dim3(B), 0, s, n,
                         static_cast<const int32_t*>(Z), x, k, steps);
  } else if (e == Elem::kF64) {
    hipLaunchKernelGGL(k_trsv_fp<double>, dim3(1), dim3(1), 0, s, k, steps, k.y);
    if (n > 0)
      hipLaunchKernelGGL(k_update_x_f64, dim3(grid_for(n)), dim3(B), 0, s, n,
                         static_cast<const double*>(Z), x, k, steps);
  } else {
    hipLaunchKernelGGL(k_trsv_fp<float>, dim3(1), dim3(1), 0, s, k, steps, k.yf);
    if (n > 0)
      hipLaunchKernelGGL(k_update_x_f32, dim3(grid_for(n)), dim3(B), 0, s, n,
                         static_cast<const float*>(Z), x, k, steps);
  }
  return hipGetLastError();
}

hipError_t launch_checksum(int n, const int32_t* v, const KState& k, hipStream_t s) {
  const int G = red_grid(n);
  int64_t* part = k.ipart + 2 * kRedBlocks;  // third half: does not disturb MGS
  hipLaunchKernelGGL(k_checksum_part, dim3(G), dim3(B), 0, s, n, v, part);
  hipLaunchKernelGGL(k_checksum_finalize, dim3(1), dim3(B), 0, s, part, G, k);
  return hipGetLastError();
}

template <typename T, bool kGather>
__global__ __launch_bounds__(B) void k_permute(int n, const int* __restrict__ perm,
                                               const T* __restrict__ src, T* __restrict__ dst) {
  const int k = blockIdx.x * B + threadIdx.x;
  if (k >= n) return;
  if (kGather) dst[k] = src[perm[k]];
  else dst[perm[k]] = src[k];
}

hipError_t launch_permute(Elem e, int n, const int* perm, const void* src, void* dst,
                          bool gather, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (n + B - 1) / B;
  auto go = [&](auto tag) {
    using T = decltype(tag);
    if (gather)
      hipLaunchKernelGGL((k_permute<T, true>), dim3(grid), dim3(B), 0, s, n, perm,
                         static_cast<const T*>(src), static_cast<T*>(dst));
    else
      hipLaunchKernelGGL((k_permute<T, false>), dim3(grid), dim3(B), 0, s, n, perm,
                         static_cast<const T*>(src), static_cast<T*>(dst));
  };
  if (e == Elem::kF64) go(double{});
  else if (e == Elem::kF32) go(float{});
  else go(int32_t{});
  return hipGetLastError();
}

// Jacobi<CSR<T>>::Smooth update (stationary.hpp:56-61): out += d_inv * t,
// with t = in - A out already formed; FP only, no contraction.
template <typename T>
__global__ __launch_bounds__(B) void k_jacobi_update(int n, const T* __restrict__ dinv,
                                                     const T* __restrict__ t, T* out) {
  const int k = blockIdx.x * B + threadIdx.x;
  if (k >= n) return;
  out[k] = out[k] + dinv[k] * t[k];
}

hipError_t launch_jacobi_update(Elem e, int n, const void* dinv, const void* t, void* out,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (n + B - 1) / B;
  if (e == Elem::kF64)
    hipLaunchKernelGGL(k_jacobi_update<double>, dim3(grid), dim3(B), 0, s, n,
                       static_cast<const double*>(dinv), static_cast<const double*>(t),
                       static_cast<double*>(out));
  else if (e == Elem::kF32)
    hipLaunchKernelGGL(k_jacobi_update<float>, dim3(grid), dim3(B), 0, s, n,
                       static_cast<const float*>(dinv), static_cast<const float*>(t),
                       static_cast<float*>(out));
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Streaming copy with 16-byte loads and stores, one word per thread (the
// HBM copy ceiling the SpMV is compared with).  A grid-stride loop over a
// capped grid reached 4.8 TB/s on 4 GiB, one word per thread 6.2 TB/s: the
// loop keeps one load per thread in flight, short-lived waves keep many.
__global__ __launch_bounds__(B) void k_copy_16(size_t n16, const int4* __restrict__ a,
                                               int4* __restrict__ b) {
  const size_t i = blockIdx.x * static_cast<size_t>(B) + threadIdx.x;
  if (i < n16) b[i] = a[i];
}

hipError_t launch_copy(Elem e, int n, const void* src, void* dst, hipStream_t s) {
  const size_t bytes = static_cast<size_t>(n) * (e == Elem::kF64 ? 8 : 4);
  const size_t n4 = bytes / 4;
  if (n4 == 0) return hipSuccess;
  if ((bytes & 15) == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0 &&
      (reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
    const size_t n16 = bytes / 16;
    const size_t blocks = (n16 + B - 1) / B;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy_16, dim3(static_cast<unsigned>(blocks)), dim3(B), 0, s, n16,
                       static_cast<const int4*>(src), static_cast<int4*>(dst));
    return hipGetLastError();
  }
  const int grid = static_cast<int>(std::min<size_t>((n4 + B - 1) / B, 8192));
  hipLaunchKernelGGL(k_copy_bytes, dim3(grid), dim3(B), 0, s, n4,
                     static_cast<const uint32_t*>(src), static_cast<uint32_t*>(dst));
  return hipGetLastError();
}

hipError_t launch_fill_zero(size_t bytes, void* dst, hipStream_t s) {
  const size_t n4 = bytes / 4;
  if (n4 == 0) return hipSuccess;
  const int grid = static_cast<int>(std::min<size_t>((n4 + B - 1) / B, 8192));
  hipLaunchKernelGGL(k_zero_bytes, dim3(grid), dim3(B), 0, s, n4,
                     static_cast<uint32_t*>(dst));
  return hipGetLastError();
}

// ---- block launchers -------------------------------------------------------

hipError_t launch_outer_residual_blk(const DCsr& a64, int count, const double* b,
                                     const double* x, double* r, const KState* ks,
                                     const int* live, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_outer_residual_blk, dim3(red_grid(a64.n), count), dim3(B), 0, s, a64.n,
                     a64.rptr, a64.cind, static_cast<const double*>(a64.val), b, x, r, ks, live);
  hipLaunchKernelGGL(k_norm_start_blk, dim3(count), dim3(B), 0, s, ks, live);
  return hipGetLastError();
}

hipError_t launch_v0_blk(int n, int groups, const double* r, void* v0, const KState* ks,
                         const int* live, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_v0_blk, dim3(grid_for(n), groups), dim3(B), 0, s, n, r,
                     static_cast<int4*>(v0), ks, live);
  return hipGetLastError();
}

hipError_t launch_dot_first_blk(int n, int groups, const void* v0, const void* w,
                                const KState* ks, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_first_blk, dim3(red_grid(n), groups), dim3(B), 0, s, n,
                     static_cast<const int4*>(v0), static_cast<const int4*>(w), ks);
  return hipGetLastError();
}

hipError_t launch_axpy_dot_blk(int n, int groups, const void* V, void* w, const KState* ks,
                               const int* live, int kk, int j, hipStream_t s) {
  const size_t vs = static_cast<size_t>(groups) * n;
  const int4* v = static_cast<const int4*>(V);
  hipLaunchKernelGGL(k_axpy_dot_blk, dim3(red_grid(n), groups), dim3(B), 0, s, n, v + kk * vs,
                     v + (kk + 1) * vs, static_cast<int4*>(w), ks, live, kk, j);
  return hipGetLastError();
}

hipError_t launch_mgs_step_blk(int n, int groups, const void* V, void* w, const KState* ks,
                               const int* live, int j, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const size_t vs = static_cast<size_t>(groups) * n;
  hipLaunchKernelGGL(k_mgs_step_blk, dim3(groups), dim3(kMgsThreads), 0, s, n, vs,
                     static_cast<const int4*>(V), static_cast<int4*>(w), ks, live, j);
  return hipGetLastError();
}

hipError_t launch_scal_blk(int n, int groups, void* w, const KState* ks, const int* live, int j,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_scal_givens_blk, dim3(red_grid(n), groups), dim3(B), 0, s, n,
                     static_cast<int4*>(w), ks, live, j);
  return hipGetLastError();
}

hipError_t launch_update_x_blk(int n, int groups, int count, const void* Z, double* x,
                               const KState* ks, const int* upd, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trsv_blk, dim3(count), dim3(64), 0, s, ks, upd);
  if (n > 0)
    hipLaunchKernelGGL(k_update_x_blk, dim3(grid_for(n), count), dim3(B), 0, s, n, groups,
                       static_cast<const int32_t*>(Z), x, ks, upd);
  return hipGetLastError();
}

hipError_t launch_flags_blk(const KState* ks, int count, int* out, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_flags_blk, dim3((count + 63) / 64), dim3(64), 0, s, ks, count, out);
  return hipGetLastError();
}

hipError_t launch_checksum_blk(int n, int groups, int count, const void* v, const KState* ks,
                               const int* live, hipStream_t s) {
  const int G = red_grid(n);
  hipLaunchKernelGGL(k_checksum_blk_part, dim3(G, groups), dim3(B), 0, s, n,
                     static_cast<const int32_t*>(v), ks, live);
  hipLaunchKernelGGL(k_checksum_blk_finalize, dim3(count), dim3(B), 0, s, ks, live, G);
  return hipGetLastError();
}

}  // namespace sc23ad::dev
