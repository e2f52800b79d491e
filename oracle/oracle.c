/*
 * CPU oracle for the solve phase of int-AMG-FGMRES(m).  TEST INFRASTRUCTURE
 * ONLY: the product never links or calls it (see oracle.h).  Compiled with
 * gcc -O2 -fno-fast-math -ffp-contract=off so every FP operation rounds where
 * the C source says.
 *
 * Integer semantics follow the reference under its C++20 rules: signed
 * arithmetic wraps modulo 2^32 / 2^64, >> is an arithmetic (floor) shift and a
 * narrowing cast keeps the low 32 bits.  Here wrap-around is done in unsigned
 * arithmetic explicitly; gcc converts out-of-range unsigned values to signed
 * modulo 2^N and shifts negative values arithmetically.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* hierarchy file                                                            */

static int rd(FILE* f, void* p, size_t b) { return b == 0 || fread(p, 1, b, f) == b; }

static size_t elem_size(int precision) {
  return precision == 0 ? sizeof(double) : precision == 1 ? sizeof(float) : 4;
}

static int read_mat(FILE* f, or_mat* a, int precision) {
  int hdr[5];
  if (!rd(f, hdr, sizeof hdr)) return 0;
  a->n = hdr[0];
  a->m = hdr[1];
  a->nnz = hdr[2];
  a->bit = hdr[3];
  a->dbit = hdr[4];
  a->rptr = malloc(sizeof(int) * (size_t)(a->n + 1));
  a->cind = malloc(sizeof(int) * (size_t)(a->nnz > 0 ? a->nnz : 1));
  a->val = malloc(elem_size(precision) * (size_t)(a->nnz > 0 ? a->nnz : 1));
  if (!a->rptr || !a->cind || !a->val) return 0;
  return rd(f, a->rptr, sizeof(int) * (size_t)(a->n + 1)) &&
         rd(f, a->cind, sizeof(int) * (size_t)a->nnz) &&
         rd(f, a->val, elem_size(precision) * (size_t)a->nnz);
}

static void free_mat(or_mat* a) {
  free(a->rptr);
  free(a->cind);
  free(a->val);
}

int or_hier_load(const char* path, or_hier** out) {
  FILE* f = fopen(path, "rb");
  if (!f) return OR_ERR_IO;
  char magic[8];
  int hdr[8];
  double dd[2];
  or_hier* h = calloc(1, sizeof *h);
  int ok = rd(f, magic, 8) && memcmp(magic, "SC23AMG1", 8) == 0 &&
           rd(f, hdr, sizeof hdr) && rd(f, dd, sizeof dd);
  if (ok) {
    h->precision = hdr[0];
    h->depth = hdr[1];
    h->smoother = hdr[2];
    h->iter = hdr[3];
    h->b_num = hdr[4];
    h->max_levels = hdr[5];
    h->coarse_n = hdr[6];
    h->n_coarse = hdr[7];
    h->omega = dd[0];
    h->theta = dd[1];
    h->lv = calloc((size_t)h->depth * 5, sizeof(or_mat));
    for (int i = 0; ok && i < h->depth * 5; ++i) ok = read_mat(f, &h->lv[i], h->precision);
    ok = ok && read_mat(f, &h->lc, h->precision) && read_mat(f, &h->uc, h->precision);
  }
  fclose(f);
  if (!ok) {
    or_hier_free(h);
    return OR_ERR_IO;
  }
  *out = h;
  return OR_OK;
}

void or_hier_free(or_hier* h) {
  if (!h) return;
  if (h->lv)
    for (int i = 0; i < h->depth * 5; ++i) free_mat(&h->lv[i]);
  free(h->lv);
  free_mat(&h->lc);
  free_mat(&h->uc);
  free(h);
}

#define LV(h, l, k) (&(h)->lv[(l) * 5 + (k)])
enum { MA = 0, MLD = 1, MU = 2, MR = 3, MP = 4 };

int or_hier_level_n(const or_hier* h, int level) {
  return level < h->depth ? LV(h, level, MA)->n : h->n_coarse;
}

/* ------------------------------------------------------------------------ */
/* integer arithmetic helpers                                                */

static inline int64_t wrap64(uint64_t v) { return (int64_t)v; }
static inline int32_t low32(int64_t v) { return (int32_t)(uint32_t)(uint64_t)v; }
static inline int64_t mul64(int64_t a, int64_t b) {
  return wrap64((uint64_t)a * (uint64_t)b);
}
static inline int64_t shl64(int64_t a, int s) { return wrap64((uint64_t)a << s); }
static inline int32_t add32(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}
static inline int32_t sub32(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a - (uint32_t)b);
}

/* blas1::int32::lsqrt (Library/blas/blas1.hpp:106-119). */
int64_t or_lsqrt(int64_t x) {
  int64_t s, t;
  if (x <= 0) return 0;
  s = 1;
  t = x;
  while (s < t) {
    s <<= 1;
    t >>= 1;
  }
  do {
    t = s;
    s = (x / s + s) >> 1;
  } while (s < t);
  return t;
}

/* CSRflex::SpMV / CSR<Fixed<bit>>::SpMV (Library/matrix/spmat.hpp:77-86,
   128-137): y_i = (int32)((sum_j (int64)a_ij * (int64)x_j) >> bit). */
void or_spmv_int(const or_mat* a, const int32_t* x, int32_t* y) {
  const int32_t* v = a->val;
  for (int i = 0; i < a->n; ++i) {
    uint64_t t = 0;
    for (int p = a->rptr[i]; p < a->rptr[i + 1]; ++p)
      t += (uint64_t)((int64_t)v[p] * (int64_t)x[a->cind[p]]);
    y[i] = low32(wrap64(t) >> a->bit);
  }
}

/* ldtri::CSRflex::SpTRSV (Library/matrix/trimat/ldtri.hpp:172-183), which is
   also BJCSRflex::SpTRSV (:511-525): LD only holds same-block entries, so the
   sequential row order gives the block-parallel result.  The stored inverse
   diagonal is the row's last entry. */
void or_ld_solve_int(const or_mat* l, const int32_t* in, int32_t* out) {
  const int32_t* v = l->val;
  for (int i = 0; i < l->n; ++i) {
    int64_t t = shl64(in[i], l->bit);
    int p;
    for (p = l->rptr[i]; p < l->rptr[i + 1] - 1; ++p)
      t = wrap64((uint64_t)t - (uint64_t)((int64_t)v[p] * (int64_t)out[l->cind[p]]));
    out[i] = low32(mul64(t >> l->bit, v[p]) >> l->dbit);
  }
}

/* ltri::CSRflex::SpTRSV (Library/matrix/trimat/ltri.hpp:125-133). */
void or_ltri_int(const or_mat* l, const int32_t* in, int32_t* out) {
  const int32_t* v = l->val;
  for (int i = 0; i < l->n; ++i) {
    int64_t t = shl64(in[i], l->bit);
    for (int p = l->rptr[i]; p < l->rptr[i + 1]; ++p)
      t = wrap64((uint64_t)t - (uint64_t)((int64_t)v[p] * (int64_t)out[l->cind[p]]));
    out[i] = low32(t >> l->bit);
  }
}

/* dutri::CSRflex::SpTRSV (Library/matrix/trimat/dutri.hpp:146-156): backward,
   inverse diagonal stored first in each row. */
void or_dutri_int(const or_mat* u, const int32_t* in, int32_t* out) {
  const int32_t* v = u->val;
  for (int i = u->n - 1; i >= 0; --i) {
    int64_t t = shl64(in[i], u->bit);
    int p;
    for (p = u->rptr[i + 1] - 1; p >= u->rptr[i] + 1; --p)
      t = wrap64((uint64_t)t - (uint64_t)((int64_t)v[p] * (int64_t)out[u->cind[p]]));
    out[i] = low32(mul64(t >> u->bit, v[p]) >> u->dbit);
  }
}

/* blas1::int32::{Dot,Nrm2,Axpy,Scal,Ggen,Grot} (Library/blas/blas1.hpp:
   230-307). */
int64_t or_dot_int(const int32_t* x, const int32_t* y, int n, int bit) {
  uint64_t s = 0;
  for (int i = 0; i < n; ++i) s += (uint64_t)((int64_t)x[i] * (int64_t)y[i]);
  return wrap64(s) >> bit;
}

int64_t or_nrm2_int(const int32_t* x, int n) {
  uint64_t s = 0;
  for (int i = 0; i < n; ++i) s += (uint64_t)((int64_t)x[i] * (int64_t)x[i]);
  return or_lsqrt(wrap64(s));
}

void or_axpy_int(int64_t a, const int32_t* x, int32_t* y, int n, int bit) {
  for (int i = 0; i < n; ++i) y[i] = add32(y[i], low32(mul64(a, x[i]) >> bit));
}

void or_scal_int(int64_t a, int32_t* x, int n, int bit) {
  for (int i = 0; i < n; ++i) x[i] = low32(mul64(a, x[i]) >> bit);
}

/* Ggen: r = lsqrt(a^2 + b^2); inv = (2^62 / r) >> (62 - 2 bit);
   c = (a inv) >> bit; s = (-(b inv)) >> bit.  r = 0 would divide by zero
   (the reference traps); it is reported through *overflow instead. */
int64_t or_ggen_int(int64_t a, int64_t b, int64_t* c, int64_t* s, int bit,
                    int* overflow) {
  int64_t r = wrap64((uint64_t)mul64(a, a) + (uint64_t)mul64(b, b));
  r = or_lsqrt(r);
  if (r == 0) {
    *overflow = 1;
    *c = 0;
    *s = 0;
    return 0;
  }
  int64_t inv = (((int64_t)1 << 62) / r) >> (62 - bit - bit);
  *c = mul64(a, inv) >> bit;
  *s = wrap64(0 - (uint64_t)mul64(b, inv)) >> bit;
  return r;
}

void or_grot_int(int64_t c, int64_t s, int64_t* a, int64_t* b, int bit) {
  int64_t t = *a;
  *a = wrap64((uint64_t)mul64(c, t) - (uint64_t)mul64(s, *b)) >> bit;
  *b = wrap64((uint64_t)mul64(s, t) + (uint64_t)mul64(c, *b)) >> bit;
}

/* ------------------------------------------------------------------------ */
/* FP reductions in the two documented orders                                */

#define RED_BLOCKS 1024
#define RED_THREADS 256

#define DEFINE_RED(NAME, T, EXPR)                                            \
  static T NAME(const T* x, const T* y, int n, int mode) {                   \
    (void)y;                                                                 \
    if (mode == OR_RED_SEQ) {                                                \
      T s = 0;                                                               \
      for (int i = 0; i < n; ++i) s += EXPR;                                 \
      return s;                                                              \
    }                                                                        \
    static T part[RED_BLOCKS];                                               \
    T sh[RED_THREADS];                                                       \
    const int total = RED_BLOCKS * RED_THREADS;                              \
    for (int blk = 0; blk < RED_BLOCKS; ++blk) {                             \
      for (int t = 0; t < RED_THREADS; ++t) {                                \
        T s = 0;                                                             \
        for (int i = blk * RED_THREADS + t; i < n; i += total) s += EXPR;    \
        sh[t] = s;                                                           \
      }                                                                      \
      for (int st = RED_THREADS / 2; st > 0; st >>= 1)                       \
        for (int t = 0; t < st; ++t) sh[t] = sh[t] + sh[t + st];             \
      part[blk] = sh[0];                                                     \
    }                                                                        \
    for (int st = RED_BLOCKS / 2; st > 0; st >>= 1)                          \
      for (int t = 0; t < st; ++t) part[t] = part[t] + part[t + st];         \
    return part[0];                                                          \
  }

DEFINE_RED(red_sq_f64, double, x[i] * x[i])
DEFINE_RED(red_dot_f64, double, x[i] * y[i])
DEFINE_RED(red_sq_f32, float, x[i] * x[i])
DEFINE_RED(red_dot_f32, float, x[i] * y[i])

double or_nrm2_f64(const double* x, int n, int mode) {
  return sqrt(red_sq_f64(x, NULL, n, mode));
}
double or_dot_f64(const double* x, const double* y, int n, int mode) {
  return red_dot_f64(x, y, n, mode);
}
float or_nrm2sq_f32(const float* x, int n, int mode) {
  return red_sq_f32(x, NULL, n, mode);
}
float or_dot_f32(const float* x, const float* y, int n, int mode) {
  return red_dot_f32(x, y, n, mode);
}

/* ------------------------------------------------------------------------ */
/* FP kernels (CSR<T>::SpMV spmat.hpp:43-52, ldtri::CSR<T>::SpTRSV
   ldtri.hpp:49-58, ltri::CSR<T> ltri.hpp:45-53, dutri::CSR<T> dutri.hpp:45-54),
   generated for double and float.                                           */

#define DEFINE_FP_KERNELS(T, SFX)                                            \
  static void spmv_##SFX(const or_mat* a, const T* x, T* y) {               \
    const T* v = a->val;                                                     \
    for (int i = 0; i < a->n; ++i) {                                         \
      T t = 0;                                                               \
      for (int p = a->rptr[i]; p < a->rptr[i + 1]; ++p) t += v[p] * x[a->cind[p]]; \
      y[i] = t;                                                              \
    }                                                                        \
  }                                                                          \
  static void ld_##SFX(const or_mat* l, const T* in, T* out) {              \
    const T* v = l->val;                                                     \
    for (int i = 0; i < l->n; ++i) {                                         \
      T t = in[i];                                                           \
      int p;                                                                 \
      for (p = l->rptr[i]; p < l->rptr[i + 1] - 1; ++p) t -= v[p] * out[l->cind[p]]; \
      out[i] = t * v[p];                                                     \
    }                                                                        \
  }                                                                          \
  static void ltri_##SFX(const or_mat* l, const T* in, T* out) {            \
    const T* v = l->val;                                                     \
    for (int i = 0; i < l->n; ++i) {                                         \
      T t = in[i];                                                           \
      for (int p = l->rptr[i]; p < l->rptr[i + 1]; ++p) t -= v[p] * out[l->cind[p]]; \
      out[i] = t;                                                            \
    }                                                                        \
  }                                                                          \
  static void dutri_##SFX(const or_mat* u, const T* in, T* out) {           \
    const T* v = u->val;                                                     \
    for (int i = u->n - 1; i >= 0; --i) {                                    \
      T t = in[i];                                                           \
      int p;                                                                 \
      for (p = u->rptr[i + 1] - 1; p >= u->rptr[i] + 1; --p) t -= v[p] * out[u->cind[p]]; \
      out[i] = t * v[p];                                                     \
    }                                                                        \
  }

DEFINE_FP_KERNELS(double, f64)
DEFINE_FP_KERNELS(float, f32)

/* ------------------------------------------------------------------------ */
/* typed dispatch for the preconditioner                                     */

typedef struct {
  int prec;
} ctx_t;

static void op_spmv(int prec, const or_mat* a, const void* x, void* y) {
  if (prec == 2) or_spmv_int(a, x, y);
  else if (prec == 0) spmv_f64(a, x, y);
  else spmv_f32(a, x, y);
}
static void op_ld(int prec, const or_mat* a, const void* x, void* y) {
  if (prec == 2) or_ld_solve_int(a, x, y);
  else if (prec == 0) ld_f64(a, x, y);
  else ld_f32(a, x, y);
}
/* y = x - y  (blas1::Sub, blas1.hpp:40-44; also Axpby(1, x, -1, y) :28-32) */
static void op_sub(int prec, const void* x, void* y, int n) {
  if (prec == 2) {
    const int32_t* a = x; int32_t* b = y;
    for (int i = 0; i < n; ++i) b[i] = sub32(a[i], b[i]);
  } else if (prec == 0) {
    const double* a = x; double* b = y;
    for (int i = 0; i < n; ++i) b[i] = a[i] - b[i];
  } else {
    const float* a = x; float* b = y;
    for (int i = 0; i < n; ++i) b[i] = a[i] - b[i];
  }
}
/* y = x + y  (blas1::Add, blas1.hpp:46-50) */
static void op_add(int prec, const void* x, void* y, int n) {
  if (prec == 2) {
    const int32_t* a = x; int32_t* b = y;
    for (int i = 0; i < n; ++i) b[i] = add32(a[i], b[i]);
  } else if (prec == 0) {
    const double* a = x; double* b = y;
    for (int i = 0; i < n; ++i) b[i] = a[i] + b[i];
  } else {
    const float* a = x; float* b = y;
    for (int i = 0; i < n; ++i) b[i] = a[i] + b[i];
  }
}

static size_t esz(int prec) { return elem_size(prec); }

/* Smooth(in, out, isInited) of SOR / BlockSOR / JacobiSOR
   (stationary.hpp:415-430, 997-1012, 647-660).  Jacobi<CSR<T>>
   (stationary.hpp:50-74; smoother 3, FP only): LD holds d_inv as its only
   entry per row, U is A; first out = d_inv * in, then
   out[i] = out[i] + d_inv[i] * (in[i] - (A out)[i]). */
static void smooth(const or_hier* h, int l, const void* in, void* out,
                   int inited, void* temp) {
  const or_mat* ld = LV(h, l, MLD);
  const or_mat* u = LV(h, l, MU);
  const int prec = h->precision, n = ld->n;
  int k0 = 0;
  if (!inited) {
    op_ld(prec, ld, in, out);
    k0 = 1;
  }
  for (int k = k0; k < h->iter; ++k) {
    op_spmv(prec, u, out, temp);
    op_sub(prec, in, temp, n);
    if (h->smoother == 3 && prec == 0) {
      const double* d = ld->val;
      const double* t = temp;
      double* o = out;
      for (int i = 0; i < n; ++i) o[i] = o[i] + d[ld->rptr[i]] * t[i];
    } else if (h->smoother == 3 && prec == 1) {
      const float* d = ld->val;
      const float* t = temp;
      float* o = out;
      for (int i = 0; i < n; ++i) o[i] = o[i] + d[ld->rptr[i]] * t[i];
    } else {
      op_ld(prec, ld, temp, out);
    }
  }
}

static void coarse_solve(const or_hier* h, const void* in, void* out) {
  const int prec = h->precision;
  void* tmp = malloc(esz(prec) * (size_t)(h->n_coarse > 0 ? h->n_coarse : 1));
  /* direct::LUflex / LU<T>::Solve (lu.hpp:91-94, 33-36) */
  if (prec == 2) {
    or_ltri_int(&h->lc, in, tmp);
    or_dutri_int(&h->uc, tmp, out);
  } else if (prec == 0) {
    ltri_f64(&h->lc, in, tmp);
    dutri_f64(&h->uc, tmp, out);
  } else {
    ltri_f32(&h->lc, in, tmp);
    dutri_f32(&h->uc, tmp, out);
  }
  free(tmp);
}

/* intAMGflex::Precondition / dfAMG::Precondition (INTAMG.hpp:269-293,
   108-133). */
int or_vcycle(const or_hier* h, const void* in, void* out) {
  const int prec = h->precision, d = h->depth;
  const size_t es = esz(prec);
  const int n0 = LV(h, 0, MA)->n;
  char* r = malloc(es * (size_t)n0);
  char* temp = malloc(es * (size_t)n0);
  char** rr = calloc((size_t)d, sizeof(char*));
  char** ee = calloc((size_t)d, sizeof(char*));
  char** xx = calloc((size_t)d, sizeof(char*));
  for (int i = 0; i < d; ++i) {
    const int len = or_hier_level_n(h, i + 1);
    rr[i] = malloc(es * (size_t)(len > 0 ? len : 1));
    ee[i] = malloc(es * (size_t)(len > 0 ? len : 1));
    xx[i] = malloc(es * (size_t)(len > 0 ? len : 1));
  }
  smooth(h, 0, in, out, 0, temp);
  op_spmv(prec, LV(h, 0, MA), out, r);
  op_sub(prec, in, r, n0);
  op_spmv(prec, LV(h, 0, MR), r, ee[0]);
  for (int i = 0; i < d - 1; ++i) {
    const int n = LV(h, i + 1, MA)->n;
    smooth(h, i + 1, ee[i], xx[i], 0, temp);
    op_spmv(prec, LV(h, i + 1, MA), xx[i], rr[i]);
    op_sub(prec, ee[i], rr[i], n);
    op_spmv(prec, LV(h, i + 1, MR), rr[i], ee[i + 1]);
  }
  coarse_solve(h, ee[d - 1], xx[d - 1]);
  for (int i = d - 2; i >= 0; --i) {
    const int n = LV(h, i + 1, MA)->n;
    op_spmv(prec, LV(h, i + 1, MP), xx[i + 1], rr[i]);
    op_add(prec, rr[i], xx[i], n);
    smooth(h, i + 1, ee[i], xx[i], 1, temp);
  }
  op_spmv(prec, LV(h, 0, MP), xx[0], r);
  op_add(prec, r, out, n0);
  smooth(h, 0, in, out, 1, temp);
  for (int i = 0; i < d; ++i) {
    free(rr[i]);
    free(ee[i]);
    free(xx[i]);
  }
  free(rr);
  free(ee);
  free(xx);
  free(r);
  free(temp);
  return OR_OK;
}

int or_level_op(const or_hier* h, int level, int op, const void* in, void* out) {
  const int prec = h->precision;
  if (op == 7) {
    if (level != h->depth) return OR_ERR_PARAM;
    coarse_solve(h, in, out);
    return OR_OK;
  }
  if (level < 0 || level >= h->depth) return OR_ERR_PARAM;
  const int n = LV(h, level, MA)->n;
  void* temp = malloc(esz(prec) * (size_t)n);
  switch (op) {
    case 0: op_spmv(prec, LV(h, level, MA), in, out); break;
    case 1: op_spmv(prec, LV(h, level, MU), in, out); break;
    case 2: op_spmv(prec, LV(h, level, MR), in, out); break;
    case 3: op_spmv(prec, LV(h, level, MP), in, out); break;
    case 4: op_ld(prec, LV(h, level, MLD), in, out); break;
    case 5: smooth(h, level, in, out, 0, temp); break;
    case 6: smooth(h, level, in, out, 1, temp); break;
    default: free(temp); return OR_ERR_PARAM;
  }
  free(temp);
  return OR_OK;
}

int or_default_bits(const or_hier* h, const double* b, double nrm_b) {
  const int n = LV(h, 0, MA)->n;
  double* c = malloc(sizeof(double) * (size_t)n);
  or_vcycle(h, b, c);
  double c_max = 0;
  for (int i = 0; i < n; ++i)
    if (fabs(c[i]) > c_max) c_max = fabs(c[i]);
  free(c);
  /* Work/main.cpp:160 */
  return (c_max / nrm_b < 1) ? 29
                             : 31 - ((int)ceil(log2((double)((int)(c_max / nrm_b) + 1))) + 1);
}

/* ------------------------------------------------------------------------ */
/* solvers                                                                   */

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t or_checksum_i32(const int32_t* v, int n) {
  uint64_t s = 0;
  for (int i = 0; i < n; ++i)
    s += mix64((uint64_t)i * 0x9E3779B97F4A7C15ULL ^ (uint64_t)(uint32_t)v[i]);
  return s;
}

/* C double -> int conversion with the x86 cvttsd2si result for values out of
   range (INT32_MIN), which is what the reference's (int) cast produces there. */
static inline int32_t trunc_i32(double v) {
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
  return (int32_t)v;
}
static inline int64_t trunc_i64(double v) {
  if (!(v > -9223372036854775808.0 && v < 9223372036854775808.0)) return INT64_MIN;
  return (int64_t)v;
}

/* outer residual r = b - A x (CSR<double>::SpMV + Axpby(1, b, -1, r)) */
static void residual(int n, const int* rptr, const int* cind, const double* val,
                     const double* b, const double* x, double* r) {
  for (int i = 0; i < n; ++i) {
    double t = 0;
    for (int p = rptr[i]; p < rptr[i + 1]; ++p) t += val[p] * x[cind[p]];
    r[i] = 1 * b[i] + -1 * t;
  }
}

double or_relres(int n, const int* rptr, const int* cind, const double* val,
                 const double* b, double nrm_b, const double* x, int mode) {
  double* r = malloc(sizeof(double) * (size_t)n);
  residual(n, rptr, cind, val, b, x, r);
  double nr = or_nrm2_f64(r, n, mode);
  free(r);
  return nr / nrm_b;
}

/* FGMRESflex::Solve (Library/solver/gmres.hpp:225-271). Returns the Arnoldi
   steps done; *conv = 1 when |e_{j+1}| <= criteria. */
static int fgmres_int(const or_hier* h, const double* r, double nrm_b, double* x,
                      int m, double eps, int bit, int mode, int32_t* V,
                      int32_t* Z, uint64_t* trace, int cap, int* tlen,
                      int* conv, int* overflow, double e0) {
  const int n = LV(h, 0, MA)->n;
  const double fact = (double)((int64_t)1 << bit);
  int64_t* c = calloc((size_t)m, sizeof(int64_t));
  int64_t* s = calloc((size_t)m, sizeof(int64_t));
  int64_t* e = calloc((size_t)m + 1, sizeof(int64_t));
  int64_t* H = calloc((size_t)(m + 1) * (size_t)m, sizeof(int64_t));
  double* y = calloc((size_t)m, sizeof(double));
  (void)mode;
  const double e_inv = 1.0 / e0;
  const int64_t criteria = trunc_i64(nrm_b * eps * e_inv * fact);
  for (int k = 0; k < n; ++k) V[k] = trunc_i32(r[k] * e_inv * fact);
  e[0] = (int64_t)1 << bit;
  int j;
  *conv = 0;
  for (j = 0; j < m; ++j) {
    int32_t* vj = V + (size_t)j * n;
    int32_t* zj = Z + (size_t)j * n;
    int32_t* w = V + (size_t)(j + 1) * n;
    or_vcycle(h, vj, zj);
    if (trace && *tlen + 2 <= cap) {
      trace[(*tlen)++] = or_checksum_i32(vj, n);
      trace[(*tlen)++] = or_checksum_i32(zj, n);
    }
    or_spmv_int(LV(h, 0, MA), zj, w);
    for (int k = 0; k <= j; ++k) {
      int64_t hk = or_dot_int(V + (size_t)k * n, w, n, bit);
      H[j * (m + 1) + k] = hk;
      or_axpy_int(wrap64(0 - (uint64_t)hk), V + (size_t)k * n, w, n, bit);
    }
    int64_t hn = or_nrm2_int(w, n);
    H[j * (m + 1) + j + 1] = hn;
    if (hn == 0) {
      *overflow = 1;
      break;
    }
    int64_t hinv = (((int64_t)1 << 62) / hn) >> (62 - bit - bit);
    or_scal_int(hinv, w, n, bit);
    for (int k = 0; k < j; ++k)
      or_grot_int(c[k], s[k], &H[j * (m + 1) + k], &H[j * (m + 1) + k + 1], bit);
    H[j * (m + 1) + j] = or_ggen_int(H[j * (m + 1) + j], H[j * (m + 1) + j + 1],
                                     &c[j], &s[j], bit, overflow);
    if (*overflow) break;
    H[j * (m + 1) + j + 1] = 0;
    e[j + 1] = mul64(s[j], e[j]) >> bit;
    e[j] = mul64(c[j], e[j]) >> bit;
    const int64_t ae = e[j + 1] < 0 ? wrap64(0 - (uint64_t)e[j + 1]) : e[j + 1];
    if (ae <= criteria) {
      ++j;
      *conv = 1;
      break;
    }
  }
  if (!*overflow) {
    /* blas2::int32::Trsv (Library/blas/blas2.hpp:51-61) */
    const int nn = m + 1;
    for (int i = j - 1; i >= 0; --i) {
      double temp = e0 * (double)e[i] / fact;
      for (int k = j - 1; k > i; --k) temp -= (double)H[k * nn + i] / fact * y[k];
      y[i] = temp / ((double)H[i * nn + i] / fact);
    }
    /* blas1::int32::Axpy_D (blas1.hpp:252-258) */
    const double fact_inv = 1.0 / (double)((int64_t)1 << bit);
    for (int k = 0; k < j; ++k) {
      const int32_t* zk = Z + (size_t)k * n;
      for (int i = 0; i < n; ++i) x[i] += y[k] * (double)zk[i] * fact_inv;
    }
  }
  free(c);
  free(s);
  free(e);
  free(H);
  free(y);
  return j;
}

/* FGMRES<double> / FGMRES<float>::Solve (gmres.hpp:48-88, 130-170). */
#define DEFINE_FGMRES_FP(T, SFX, PREC_TAG)                                     \
  static int fgmres_##SFX(const or_hier* h, const double* r, double nrm_b,    \
                          double* x, int m, double eps, int mode, T* V, T* Z, \
                          int* conv, double e0d) {                            \
    const int n = LV(h, 0, MA)->n;                                            \
    T* c = calloc((size_t)m, sizeof(T));                                      \
    T* s = calloc((size_t)m, sizeof(T));                                      \
    T* e = calloc((size_t)m + 1, sizeof(T));                                  \
    T* H = calloc((size_t)(m + 1) * (size_t)m, sizeof(T));                    \
    T* y = calloc((size_t)m, sizeof(T));                                      \
    e[0] = (T)e0d;                                                            \
    const double e_inv = 1.0 / e[0];                                          \
    for (int k = 0; k < n; ++k) V[k] = (T)(r[k] * e_inv);                     \
    int j;                                                                    \
    *conv = 0;                                                                \
    for (j = 0; j < m; ++j) {                                                 \
      T* w = V + (size_t)(j + 1) * n;                                         \
      or_vcycle(h, V + (size_t)j * n, Z + (size_t)j * n);                     \
      spmv_##SFX(LV(h, 0, MA), Z + (size_t)j * n, w);                         \
      for (int k = 0; k <= j; ++k) {                                          \
        T hk = red_dot_##SFX(V + (size_t)k * n, w, n, mode);                  \
        H[j * (m + 1) + k] = hk;                                              \
        const T a = -hk;                                                      \
        const T* vk = V + (size_t)k * n;                                      \
        for (int i = 0; i < n; ++i) w[i] += a * vk[i];                        \
      }                                                                       \
      T hn = PREC_TAG##_sqrt(red_sq_##SFX(w, NULL, n, mode));                 \
      H[j * (m + 1) + j + 1] = hn;                                            \
      const T sc = 1 / hn;                                                    \
      for (int i = 0; i < n; ++i) w[i] *= sc;                                 \
      for (int k = 0; k < j; ++k) {                                           \
        T* a = &H[j * (m + 1) + k];                                           \
        T* b = &H[j * (m + 1) + k + 1];                                       \
        T t = a[0];                                                           \
        a[0] = c[k] * t - s[k] * b[0];                                        \
        b[0] = s[k] * t + c[k] * b[0];                                        \
      }                                                                       \
      {                                                                       \
        T a = H[j * (m + 1) + j], b = H[j * (m + 1) + j + 1];                 \
        T rr = PREC_TAG##_sqrt(a * a + b * b);                                \
        c[j] = a / rr;                                                        \
        s[j] = -b / rr;                                                       \
        H[j * (m + 1) + j] = rr;                                              \
      }                                                                       \
      H[j * (m + 1) + j + 1] = 0;                                             \
      e[j + 1] = s[j] * e[j];                                                 \
      e[j] = c[j] * e[j];                                                     \
      if (PREC_TAG##_fabs(e[j + 1]) <= nrm_b * eps) {                         \
        ++j;                                                                  \
        *conv = 1;                                                            \
        break;                                                                \
      }                                                                       \
    }                                                                         \
    /* blas2::Trsv<T> (blas2.hpp:24-34) */                                    \
    for (int i = j - 1; i >= 0; --i) {                                        \
      T temp = e[i];                                                          \
      for (int k = j - 1; k > i; --k) temp -= H[k * (m + 1) + i] * y[k];      \
      y[i] = temp / H[i * (m + 1) + i];                                       \
    }                                                                         \
    for (int k = 0; k < j; ++k) {                                             \
      const T* zk = Z + (size_t)k * n;                                        \
      for (int i = 0; i < n; ++i) x[i] += (double)(y[k] * zk[i]);             \
    }                                                                         \
    free(c);                                                                  \
    free(s);                                                                  \
    free(e);                                                                  \
    free(H);                                                                  \
    free(y);                                                                  \
    return j;                                                                 \
  }

static inline double dbl_sqrt(double v) { return sqrt(v); }
static inline double dbl_fabs(double v) { return fabs(v); }
static inline float flt_sqrt(float v) { return sqrtf(v); }
static inline double flt_fabs(float v) { return fabs((double)v); }
DEFINE_FGMRES_FP(double, f64, dbl)
DEFINE_FGMRES_FP(float, f32, flt)

/* Restarted::Solve (Library/solver/irsolver.hpp:27-43). */
int or_solve(const or_hier* h, int n, const int* rptr, const int* cind,
             const double* val, const double* b, double nrm_b, int m,
             double eps, int bit, int max_outer, int mode, double* x,
             uint64_t* trace, int trace_cap, or_result* res) {
  memset(res, 0, sizeof *res);
  if (h->precision == 2 && (bit < 1 || bit > 30)) return OR_ERR_PARAM;
  const size_t es = esz(h->precision);
  double* r = malloc(sizeof(double) * (size_t)n);
  void* V = malloc(es * (size_t)n * (size_t)(m + 1));
  void* Z = malloc(es * (size_t)n * (size_t)(m + 1));
  double prev = 0;
  int iter = 0, overflow = 0;
  res->status = 6;
  for (int i = 0; i < max_outer; ++i) {
    residual(n, rptr, cind, val, b, x, r);
    const double nrm_r = or_nrm2_f64(r, n, mode);
    if (fabs(nrm_r - prev) / nrm_r < 1.0e-8) {
      res->status = 4;
      break;
    }
    prev = nrm_r;
    int conv = 0, steps;
    ++res->restarts;
    if (h->precision == 2)
      steps = fgmres_int(h, r, nrm_b, x, m, eps, bit, mode, V, Z, trace,
                         trace_cap, &res->trace_len, &conv, &overflow, nrm_r);
    else if (h->precision == 0)
      steps = fgmres_f64(h, r, nrm_b, x, m, eps, mode, V, Z, &conv, nrm_r);
    else
      steps = fgmres_f32(h, r, nrm_b, x, m, eps, mode, V, Z, &conv, nrm_r);
    if (overflow) {
      res->status = OR_ERR_OVERFLOW;
      break;
    }
    iter += conv ? steps : m;
    if (conv) {
      res->status = OR_OK;
      res->converged = 1;
      res->iters = iter;
      break;
    }
  }
  free(r);
  free(V);
  free(Z);
  return OR_OK;
}
