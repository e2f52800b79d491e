/*
 * CPU oracle for the solve phase of int-AMG-FGMRES(m).  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, as the checker.  The product (sc23-ad_amd/) never links or calls it.
 *
 * It is a plain-C restatement of the reference's solve phase, operating on a
 * hierarchy file ("SC23AMG1", written by the product's setup or by the
 * reference harness).  Every function cites the reference lines it follows.
 *
 * FP reductions (norms and dots) come in two orders:
 *   OR_RED_SEQ  left to right, as the reference compiled without OpenMP and
 *               without fast-math sums them; used to pin this restatement
 *               against the reference itself;
 *   OR_RED_GPU  the device library's fixed blocked order (1024 blocks x 256
 *               threads, thread g sums elements g, g+262144, ... in order;
 *               then a halving tree s[t] += s[t+stride] inside each block and
 *               over the 1024 block partials); used to check the GPU bit for bit.
 * Integer reductions are order-free (two's-complement sums mod 2^64).
 */
#ifndef SC23AD_ORACLE_H
#define SC23AD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_RED_SEQ = 0, OR_RED_GPU = 1 };
enum { OR_OK = 0, OR_ERR_IO = 1, OR_ERR_PARAM = 2, OR_ERR_OVERFLOW = 3 };

typedef struct {
  int n, m, nnz, bit, dbit;
  int *rptr, *cind;
  void* val; /* int32_t, float or double by the hierarchy's precision */
} or_mat;

typedef struct {
  int precision; /* 0 FP64, 1 FP32, 2 INT */
  int depth, smoother, iter, b_num, max_levels, coarse_n, n_coarse;
  double omega, theta;
  or_mat* lv; /* depth * 5: A, LD, U, R, P per level */
  or_mat lc, uc;
} or_hier;

int or_hier_load(const char* path, or_hier** out);
void or_hier_free(or_hier* h);
int or_hier_level_n(const or_hier* h, int level); /* rows of level 0..depth */

/* ---- integer kernels ---- */
int64_t or_lsqrt(int64_t x);                             /* blas1.hpp:106-119 */
void or_spmv_int(const or_mat* a, const int32_t* x, int32_t* y);
void or_ld_solve_int(const or_mat* ld, const int32_t* in, int32_t* out);
void or_ltri_int(const or_mat* l, const int32_t* in, int32_t* out);
void or_dutri_int(const or_mat* u, const int32_t* in, int32_t* out);
int64_t or_dot_int(const int32_t* x, const int32_t* y, int n, int bit);
int64_t or_nrm2_int(const int32_t* x, int n);
void or_axpy_int(int64_t a, const int32_t* x, int32_t* y, int n, int bit);
void or_scal_int(int64_t a, int32_t* x, int n, int bit);
int64_t or_ggen_int(int64_t a, int64_t b, int64_t* c, int64_t* s, int bit,
                    int* overflow);
void or_grot_int(int64_t c, int64_t s, int64_t* a, int64_t* b, int bit);

/* ---- FP reductions ---- */
double or_nrm2_f64(const double* x, int n, int mode);
double or_dot_f64(const double* x, const double* y, int n, int mode);
float or_nrm2sq_f32(const float* x, int n, int mode); /* sum of squares */
float or_dot_f32(const float* x, const float* y, int n, int mode);

/* ---- preconditioner ---- */
/* One V-cycle, out = M^-1 in; vectors of the hierarchy's element type. */
int or_vcycle(const or_hier* h, const void* in, void* out);
/* Level operation, op as sc23ad_level_op: 0 A, 1 U, 2 R, 3 P spmv, 4 LD solve,
   5 Smooth(false), 6 Smooth(true), 7 coarse LU (level = depth). */
int or_level_op(const or_hier* h, int level, int op, const void* in, void* out);
/* Work/main.cpp:153-160 with an FP64 hierarchy. */
int or_default_bits(const or_hier* h_fp64, const double* b, double nrm_b);

/* ---- solvers ---- */
typedef struct {
  int iters;      /* Converge.iter (0 when not converged) */
  int converged;
  int restarts;
  int status;     /* OR_OK / OR_ERR_OVERFLOW / 4 stagnation / 6 budget */
  int trace_len;
} or_result;

/* Restarted(FGMRES*) on the row-scaled FP64 fine operator (rptr, cind, val):
   x is the initial guess and the result.  bit is the fractional bit count of
   the integer solver (ignored for FP64 / FP32).  trace (optional) receives
   the order-free checksums of every V-cycle input and output of the integer
   solver, two per Arnoldi step. */
int or_solve(const or_hier* h, int n, const int* rptr, const int* cind,
             const double* val, const double* b, double nrm_b, int m,
             double eps, int bit, int max_outer, int mode, double* x,
             uint64_t* trace, int trace_cap, or_result* res);
double or_relres(int n, const int* rptr, const int* cind, const double* val,
                 const double* b, double nrm_b, const double* x, int mode);
uint64_t or_checksum_i32(const int32_t* v, int n);

#ifdef __cplusplus
}
#endif
#endif
