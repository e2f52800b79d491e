/*
 * sc23ad.h - C ABI of the MI355X-native integer AMG-preconditioned FGMRES(m)
 * solve phase.
 *
 * The reference (suzuki-hpc/SC23-AD) is a header-only C++ library whose call
 * surface is a set of abstract classes.  This ABI replaces them with opaque
 * handles and plain pointers; the precision that the reference selects by
 * template argument (solver 0 / 1 / 2 = FP64 / FP32-mixed / integer) is an
 * argument here.  Every function returns an sc23ad_status; the message of the
 * last failure on the calling thread is available from sc23ad_last_error().
 * Nothing in this ABI calls exit(), where the reference does.
 *
 * Reference interface replaced by each entry point (file:line, relative to the
 * reference repository):
 *   CSRMat, io::ReadMM, CSRMat::LMax1Scaling  -> sc23ad_csrmat_*
 *       (Library/matrix/csrmat.hpp:13-139, Library/io/mm.hpp:24-98)
 *   intAMGflex / dfAMG constructors           -> sc23ad_hier_build, sc23ad_amg_create
 *       (Library/preconditioner/INTAMG.hpp:25-84,149-245)
 *   Preconditioner<T>::Precondition           -> sc23ad_amg_precondition
 *       (Library/preconditioner/preconditioner.hpp:12-17; INTAMG.hpp:108-133,269-293)
 *   SpMat<T>::SpMV, ldtri::SpMat<T>::SpTRSV,  -> sc23ad_amg_level_op
 *   Smoother<T>::Smooth, direct::LU*::Solve
 *       (Library/matrix/spmat.hpp:13-19,43-52,128-137;
 *        Library/matrix/trimat/ldtri.hpp:8-14,172-183,511-525;
 *        Library/smoother/smoother.hpp:9-16, stationary.hpp:415-430,997-1012;
 *        Library/solver/lu.hpp:91-94)
 *   FGMRES<double|float>, FGMRESflex ctors    -> sc23ad_solver_create
 *       (Library/solver/gmres.hpp:28-29,110-111,194-195)
 *   Restarted ctor + Restarted::Solve         -> sc23ad_solver_create / sc23ad_solver_solve
 *       (Library/solver/irsolver.hpp:16,27-43; Solver::Solve solver.hpp:12-30)
 *   fractional-bit rule of the driver         -> sc23ad_default_frac_bits
 *       (Work/main.cpp:153-160)
 *   CSR<Fixed<30>>::SpMV / CSRflex::SpMV       -> sc23ad_spmv_int (raw test hook)
 *       (Library/matrix/spmat.hpp:77-86,128-137)
 */
#ifndef SC23AD_H
#define SC23AD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SC23AD_OK = 0,
  SC23AD_BAD_PARAM = 1,
  SC23AD_SETUP_FAIL = 2,    /* where the reference's setup calls exit()      */
  SC23AD_OVERFLOW = 3,      /* fixed-point overflow / division by zero        */
  SC23AD_STAGNATION = 4,    /* Restarted: |nrm_r - prev| / nrm_r < 1e-8       */
  SC23AD_HIP_ERROR = 5,
  SC23AD_NOT_CONVERGED = 6, /* restart budget exhausted                       */
  SC23AD_IO_ERROR = 7,
  SC23AD_NO_DEVICE = 8      /* device entry point called without a usable GPU */
} sc23ad_status;

typedef enum { SC23AD_FP64 = 0, SC23AD_FP32 = 1, SC23AD_INT = 2 } sc23ad_precision;
/* SOR (seq path), BlockSOR(b_num) (multi path) and the reference's opt-in
   JacobiSOR: 1200-row blocks, Gauss-Seidel across blocks, Jacobi inside one
   (JacobiSORflex / JacobiSOR<T>, Library/smoother/stationary.hpp:447-676),
   and damped Jacobi for the FP64 / FP32 solvers (Jacobi<CSR<T>>,
   stationary.hpp:15-89; omega = JacobiParam::omega). */
typedef enum {
  SC23AD_SOR = 0,
  SC23AD_BLOCK_SOR = 1,
  SC23AD_JACOBI_SOR = 2,
  SC23AD_JACOBI = 3
} sc23ad_smoother;

/* Matrix kinds for sc23ad_csrmat_gen (parameters in brackets). */
typedef enum {
  SC23AD_GEN_POISSON2D = 0,   /* [n]                                        */
  SC23AD_GEN_LAPLACE3D = 1,   /* [n]                                        */
  SC23AD_GEN_DIFFUSION27 = 2, /* [n, block, ax, ay, az, seed]               */
  SC23AD_GEN_CONVDIFF2D = 3,  /* [n, peclet, angle]                         */
  SC23AD_GEN_INCLUSIONS3D = 4 /* [n, nspheres, rmin, rmax, k_inside, seed]  */
} sc23ad_gen_kind;

typedef struct sc23ad_csrmat_s sc23ad_csrmat; /* host CSR, FP64 values          */
typedef struct sc23ad_hier_s sc23ad_hier;     /* host AMG hierarchy              */
typedef struct sc23ad_amg_s sc23ad_amg;       /* hierarchy resident on the GPU   */
typedef struct sc23ad_solver_s sc23ad_solver; /* Restarted(FGMRES(m)) on the GPU */

/* AMGParam + SORParam / BlockSORParam (Library/parameters.hpp:79-120). */
typedef struct {
  int max_levels;  /* AMGParam::level (20 in Work/main.cpp)                  */
  double theta;    /* strength threshold (0.8)                              */
  int cycle;       /* 1 = V-cycle (the only cycle the reference implements)  */
  int smoother;    /* sc23ad_smoother                                       */
  double omega;    /* relaxation weight (1.0)                               */
  int iter;        /* sweeps per smoothing step (1)                         */
  int b_num;       /* BlockSOR block count (40)                             */
  int coarse_n;    /* stop coarsening at <= coarse_n rows (2000)            */
} sc23ad_amg_param;

/* Converge{isConverged, iter} (Library/solver/solver.hpp:12-20), extended. */
typedef struct {
  int is_converged;
  int iter;          /* total Arnoldi steps, 0 when not converged (as reference) */
  int restarts;      /* inner FGMRES cycles run                                   */
  int status;        /* sc23ad_status of the solve                                */
  int overflow_flag; /* a fixed-point step divided by zero or saturated          */
  int fallback_used; /* learned parameters were replaced by the defaults         */
  int frac_bits;     /* F used by the integer solver                              */
  double implicit_relres; /* |e_{j+1}| / nrm_b at exit (FP64 view)                 */
  double solve_ms;   /* device time of the solve (hipEvent)                       */
} sc23ad_converge;

const char* sc23ad_last_error(void);
const char* sc23ad_version(void);
void sc23ad_amg_param_default(sc23ad_amg_param* p, int smoother);

/* ---- host CSR ------------------------------------------------------------ */
int sc23ad_csrmat_create(int n, int m, const int* rptr, const int* cind,
                         const double* val, sc23ad_csrmat** out);
int sc23ad_csrmat_read_mm(const char* path, sc23ad_csrmat** out);
int sc23ad_csrmat_gen(int kind, const double* params, int nparams,
                      sc23ad_csrmat** out);
int sc23ad_csrmat_info(const sc23ad_csrmat* a, int* n, int* m, int* nnz);
int sc23ad_csrmat_export(const sc23ad_csrmat* a, int* rptr, int* cind,
                         double* val);
int sc23ad_csrmat_lmax1_scaling(sc23ad_csrmat* a, double* b);
int sc23ad_csrmat_matvec(const sc23ad_csrmat* a, const double* x, double* y);
void sc23ad_csrmat_destroy(sc23ad_csrmat* a);
/* x* ~ U[-1,1] from mt19937_64(seed) with u = (x >> 11) * 2^-53. */
int sc23ad_xstar(int n, uint64_t seed, double* xstar);
/* Draws of the batched convection-diffusion family: n, peclet, angle per
   system, `count` systems from `seed`. */
int sc23ad_convdiff_draws(int count, uint64_t seed, int* n, double* pe,
                          double* angle);

/* ---- host hierarchy (setup; no GPU needed) ------------------------------- */
int sc23ad_hier_build(const sc23ad_csrmat* a, const sc23ad_amg_param* p,
                      int precision, sc23ad_hier** out);
/* Several precisions from one coarsening chain (the chain is the expensive,
   precision-independent part of the setup): outs[k] receives the hierarchy
   for precisions[k]. */
int sc23ad_hier_build_multi(const sc23ad_csrmat* a, const sc23ad_amg_param* p,
                            const int* precisions, int count, sc23ad_hier** outs);
/* The same setup with the classical interpolation weights and both Galerkin
   products of every coarsening step computed on the GPU
   (get_interpolation_matrix and get_coarse_matrix,
   Library/helper/helper_multigrid.hpp:457-586); the strength graph and the
   C/F passes stay on the host.  Results are identical to
   sc23ad_hier_build_multi.  NO_DEVICE without a GPU. */
int sc23ad_hier_build_multi_dev(const sc23ad_csrmat* a, const sc23ad_amg_param* p,
                                const int* precisions, int count, sc23ad_hier** outs);
/* depth (number of coarsening steps); level_n / level_nnz receive depth+1
   entries (operator rows / nnz per level, the last one the coarsest grid). */
int sc23ad_hier_info(const sc23ad_hier* h, int* depth, int* level_n,
                     int* level_nnz);
int sc23ad_hier_export(const sc23ad_hier* h, const char* path);
void sc23ad_hier_destroy(sc23ad_hier* h);

/* ---- device ---------------------------------------------------------------- */
int sc23ad_device_count(int* count);
int sc23ad_dev_alloc(size_t bytes, void** ptr);
int sc23ad_dev_free(void* ptr);
int sc23ad_dev_upload(void* dst, const void* src, size_t bytes);
int sc23ad_dev_download(void* dst, const void* src, size_t bytes);
/* Device-to-device copy with the library's streaming copy kernel (one read
   and one write of every byte; also the calibration kernel for HBM traffic
   counters).  bytes must be a multiple of 4 and below 8 GiB. */
int sc23ad_dev_copy(void* dst, const void* src, size_t bytes);
int sc23ad_dev_sync(void);

/* Raw fixed-point CSR SpMV on device arrays: y_i = (int32)((sum_j (int64)val_j *
   (int64)x_{cind_j}) >> bit).  All pointers are device pointers. */
int sc23ad_spmv_int(int n, const int* rptr, const int* cind, const int32_t* val,
                    int bit, const int32_t* x, int32_t* y);

/* Upload a hierarchy (kept by the amg; `h` may be destroyed afterwards). */
int sc23ad_amg_create(const sc23ad_hier* h, sc23ad_amg** out);
void sc23ad_amg_destroy(sc23ad_amg* m);
int sc23ad_amg_precision(const sc23ad_amg* m, int* precision);
/* One V-cycle out = M^-1 in.  Vectors are int32 (INT), double (FP64) or
   float (FP32) with the fine-level size; on_device selects device pointers. */
int sc23ad_amg_precondition(sc23ad_amg* m, const void* in, void* out,
                            int on_device);

/* Algorithmic HBM bytes of one V-cycle: the sum over its kernels of the
   SpTRSV and SpMV formulas (values, indices, row pointers, vectors read and
   written once) and the vector passes.  No reference counterpart
   (measurement of intAMGflex::Precondition, INTAMG.hpp:269-293). */
int sc23ad_amg_vcycle_bytes(const sc23ad_amg* m, double* bytes);
/* Mean duration of `reps` back-to-back V-cycles (hipEvents around them on
   the library stream, after one warm-up cycle).  in, out: device vectors of
   the fine level's element type; out receives the V-cycle of in. */
int sc23ad_amg_time_vcycles(sc23ad_amg* m, int reps, const void* in, void* out,
                            double* ms_per_cycle);

/* Level operations (device vectors of the level's element type). */
typedef enum {
  SC23AD_OP_SPMV_A = 0,    /* out = A_l in                                   */
  SC23AD_OP_SPMV_U = 1,    /* out = U_l in                                   */
  SC23AD_OP_SPMV_R = 2,    /* out = R_l in   (level l -> l+1)                */
  SC23AD_OP_SPMV_P = 3,    /* out = P_l in   (level l+1 -> l)                */
  SC23AD_OP_TRSV_LD = 4,   /* out = LD_l^-1 in                               */
  SC23AD_OP_SMOOTH0 = 5,   /* Smooth(in, out, false)                         */
  SC23AD_OP_SMOOTH1 = 6,   /* Smooth(in, out, true): out is read and written  */
  SC23AD_OP_COARSE = 7     /* out = LU^-1 in on the coarsest grid (l = depth) */
} sc23ad_level_op;
int sc23ad_amg_level_op(sc23ad_amg* m, int level, int op, const void* in,
                        void* out);

/* sc23ad_amg_level_op without any allocation (no counterpart in the
   reference): for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid it).  The
   operand and the result are staged, in the level's numbering, in a static
   device array of 16 MiB (two vectors of at most 8 MiB: 2,097,152 int32 or
   FP32 values, 1,048,576 FP64 values); larger levels are rejected with
   SC23AD_BAD_PARAM.  Every call on a GPU shares that GPU's array, so calls
   that overlap on one GPU race; the array keeps the last call's contents,
   all of which a call overwrites before reading.  in, out: device vectors
   as for sc23ad_amg_level_op, whose results this returns. */
int sc23ad_amg_level_op_noalloc(sc23ad_amg* m, int level, int op, const void* in,
                                void* out);

/* FP64 fine operator for the outer residual (the reference's dA). */
int sc23ad_solver_create(int precision, const sc23ad_csrmat* a, sc23ad_amg* m,
                         int restart_m, double eps, int frac_bits,
                         int max_restarts, sc23ad_solver** out);
void sc23ad_solver_destroy(sc23ad_solver* s);
/* Restarted::Solve with x as the initial guess and the result.  on_device
   selects device pointers for b and x (FP64, fine size). */
int sc23ad_solver_solve(sc23ad_solver* s, const double* b, double nrm_b,
                        double* x, int on_device, sc23ad_converge* out);
/* Several independent solves at once (no counterpart in the reference, whose
   drivers solve one system at a time: Work/main.cpp:167-169).  Solver k runs
   on its own stream with its own work buffers, taken by one of `threads` host
   threads (0: min(count, 16)); solvers may share a hierarchy (several
   right-hand sides).  b[k], x[k]: host or device pointers (on_device);
   out[k] (optional) receives solve k's record.  Each solver may appear once. */
int sc23ad_solve_batch(int count, sc23ad_solver* const* solvers, const double* const* b,
                       const double* nrm_b, double* const* x, int on_device, int threads,
                       sc23ad_converge* out);
/* ---- block solves: several right-hand sides of one hierarchy ------------
   No counterpart in the reference, whose drivers solve one right-hand side
   per Restarted::Solve (Work/main.cpp:167-169).  Up to SC23AD_BLOCK_MAX_RHS
   integer solves on one hierarchy run in lockstep: every V-cycle
   (INTAMG.hpp:269-293), fine-level SpMV and Gram-Schmidt pass of the step
   is one launch for all of them, so the matrices and level schedules are
   read once per step and the sequential levels of the LD solves
   (ldtri.hpp:511-525) are walked once for four right-hand sides per
   workgroup.  Each right-hand side's results (x, iterations, statuses) are
   those of its own sc23ad_solver_solve, bit for bit. */
#define SC23AD_BLOCK_MAX_RHS 64
typedef struct sc23ad_block_solver_s sc23ad_block_solver;
/* One V-cycle on nrhs vectors in the block layout: element (row i, vector r)
   at ((r / 4) * n + i) * 4 + r % 4 (int32; ceil(nrhs / 4) * 4 * n words).
   Integer hierarchies only. */
int sc23ad_amg_precondition_block(sc23ad_amg* m, int nrhs, const int32_t* in, int32_t* out,
                                  int on_device);
/* Mean duration of `reps` back-to-back block V-cycles on nrhs device
   vectors in the block layout (hipEvents on the library stream, after one
   warm-up cycle); out receives the V-cycles of in.  No reference
   counterpart (measurement of intAMGflex::Precondition, INTAMG.hpp:269-293,
   on several right-hand sides at once). */
int sc23ad_amg_time_vcycles_block(sc23ad_amg* m, int nrhs, int reps, const int32_t* in,
                                  int32_t* out, double* ms_per_cycle);
/* 1 when every LD solve of the hierarchy has a block kernel (otherwise those
   levels take the right-hand sides one at a time; results are the same). */
int sc23ad_amg_block_native(const sc23ad_amg* m, int* native);
/* frac_bits: F of each right-hand side (nrhs entries). */
int sc23ad_block_solver_create(const sc23ad_csrmat* a, sc23ad_amg* m, int nrhs, int restart_m,
                               double eps, const int* frac_bits, int max_restarts,
                               sc23ad_block_solver** out);
void sc23ad_block_solver_destroy(sc23ad_block_solver* s);
/* b, x: nrhs FP64 columns of n (column r at r * n); x is the initial guess and
   the result; nrm_b: nrhs norms; out (optional): nrhs records. */
int sc23ad_block_solver_solve(sc23ad_block_solver* s, const double* b, const double* nrm_b,
                              double* x, int on_device, sc23ad_converge* out);
/* Several block solvers (e.g. one per system of a batch of uneven systems),
   each on its own stream, taken by `threads` host threads (0: up to 16):
   b[k], nrm_b[k], x[k], out[k] as for sc23ad_block_solver_solve. */
int sc23ad_block_solve_batch(int count, sc23ad_block_solver* const* solvers,
                             const double* const* b, const double* const* nrm_b,
                             double* const* x, int on_device, int threads,
                             sc23ad_converge* const* out);
/* buf == NULL: record V-cycle checksums (as sc23ad_solver_trace) of every
   right-hand side during the next solves, cap per right-hand side; else
   read right-hand side rhs's trace. */
int sc23ad_block_solver_trace(sc23ad_block_solver* s, int rhs, uint64_t* buf, int cap,
                              int* len);
int sc23ad_block_solver_history(sc23ad_block_solver* s, int rhs, double* buf, int cap,
                                int* len);

/* FP reduction order of the solver's norms and inner products:
   0 = the fixed blocked order (default, 1024 x 256 partial sums, halving
   trees); 1 = strictly left to right in one thread, the order of the
   reference built without OpenMP (slow; for bit-exact comparison with it). */
int sc23ad_solver_set_reduction(sc23ad_solver* s, int sequential);
/* Record an order-free checksum of every V-cycle input and output (two
   uint64 per Arnoldi step) during the next solves; cap = capacity. */
int sc23ad_solver_trace(sc23ad_solver* s, uint64_t* buf, int cap, int* len);
/* Time every fine-level SpMV of the Arnoldi steps (A0 Z_j, gmres.hpp:240)
   and every fine-level LD solve of the smoother (ldtri.hpp:172-183,
   511-525) with hipEvents during the next solves; report their count and
   mean duration.  Single-launch operations bind the events to the kernel's
   dispatch (hipExtLaunchKernel: execution time only); multi-launch ones are
   timed on the stream around their launches. */
int sc23ad_solver_time_spmv(sc23ad_solver* s, int enable);
int sc23ad_solver_spmv_stats(sc23ad_solver* s, int* launches, double* mean_ms);
int sc23ad_solver_trsv_stats(sc23ad_solver* s, int* launches, double* mean_ms);
/* Stall stop (no counterpart in the reference, whose only early exit is
   Restarted's stagnation test, irsolver.hpp:35): with steps > 0 a solve whose
   implicit residual has not fallen below its best value by a relative 1e-3
   for `steps` consecutive Arnoldi steps stops with SC23AD_STAGNATION and
   leaves x at its last restart.  0 (the default) keeps the reference's
   behaviour.  The learned setup uses it to fall back to the defaults early. */
int sc23ad_solver_set_stall(sc23ad_solver* s, int steps);
int sc23ad_block_solver_set_stall(sc23ad_block_solver* s, int steps);
/* |e_{j+1}| / nrm_b after every Arnoldi step of the last solve (the
   reference's "# [j]" convergence history, gmres.hpp:253-255). */
int sc23ad_solver_history(sc23ad_solver* s, double* buf, int cap, int* len);
/* Explicit relative residual ||b - A x|| / nrm_b on the device (FP64). */
int sc23ad_solver_relres(sc23ad_solver* s, const double* b, double nrm_b,
                         const double* x, int on_device, double* relres);

/* Fractional bits of the integer solver by the driver's rule (Work/main.cpp:
   153-160): c = M_fp64^-1 b; F = 29 if max|c|/nrm_b < 1 else
   31 - (ceil(log2(floor(max|c|/nrm_b) + 1)) + 1).  m_fp64 must be an FP64 amg. */
int sc23ad_default_frac_bits(sc23ad_amg* m_fp64, const double* b, double nrm_b,
                             int* bits);

/* ---- learned per-level parameters ------------------------------------------
   No counterpart in the reference, which uses one omega for every level
   (SORParam / BlockSORParam, Work/main.cpp:82-83) and one fractional bit
   length F for every vector (the FGMRESflex ctor, gmres.hpp:194-195; rule
   main.cpp:153-160).  A small network (8 -> 32 -> 32 -> 2, tanh, sigmoid
   outputs) maps 8 features of each level's operator to the level's smoother
   weight omega_l in [0.6, 1.9] and the fractional bits of the next coarser
   level's vectors relative to F, in [-21, 1]; the integer kernels are
   unchanged (sc23ad_hier_retune_levels folds the bits into the R / P
   shifts). */
typedef struct sc23ad_mlp_s sc23ad_mlp;
/* Hierarchy with other smoother parameters from base's coarsening (no
   re-coarsening): level_omega / level_bits (nlevels entries, either may be
   NULL; bits <= 0 = no cap). */
int sc23ad_hier_retune(const sc23ad_hier* base, const double* level_omega,
                       const int* level_bits, int nlevels, int precision,
                       sc23ad_hier** out);
/* Hierarchy with per-level smoother weights and per-level fractional bits of
   the level vectors, from base's coarsening: level l's integer vectors carry
   F + level_vbits[l] fractional bits (level_vbits: nlevels + 1 entries for
   levels 0..depth, level_vbits[0] = 0).  The smoothers, residuals and the
   coarse solve are homogeneous in the vector scale, so the integer kernels
   are unchanged: only R_l's shift drops and P_l's grows by
   d = level_vbits[l+1] - level_vbits[l] (INTAMG.hpp:278-290; a step outside
   the shifts' range is cut to it and counted by sc23ad_hier_saturated).
   level_omega as for sc23ad_hier_retune; either array may be NULL. */
int sc23ad_hier_retune_levels(const sc23ad_hier* base, const double* level_omega,
                              const int* level_vbits, int nlevels, int precision,
                              sc23ad_hier** out);
/* Levels whose requested bits exceeded the flex rule's maximum (the rule's
   value was kept: the default is the fallback for a saturated prediction). */
int sc23ad_hier_saturated(const sc23ad_hier* h, int* levels);
/* 8 features per level (feat: cap x 8 floats, NULL to query nlevels). */
int sc23ad_hier_features(const sc23ad_hier* h, float* feat, int cap, int* nlevels);
int sc23ad_mlp_create(uint64_t seed, sc23ad_mlp** out);
void sc23ad_mlp_destroy(sc23ad_mlp* m);
int sc23ad_mlp_nparams(int* n);
/* Copy the parameters out (get) and/or in (set; resets the optimiser). */
int sc23ad_mlp_params(sc23ad_mlp* m, float* get, const float* set);
/* y (batch x 2, in [0,1]) = network(x (batch x 8)); host arrays. */
int sc23ad_mlp_forward(sc23ad_mlp* m, const float* x, int batch, float* y);
/* Mean over the batch of sum_k (y_k - t_k)^2 and its gradient (nparams). */
int sc23ad_mlp_grad(sc23ad_mlp* m, const float* x, const float* t, int batch, float* loss,
                    float* grad);
/* `steps` full-batch Adam steps on the GPU; *loss = loss afterwards. */
int sc23ad_mlp_train(sc23ad_mlp* m, const float* x, const float* t, int batch, int steps,
                     float lr, float* loss);
/* The network's omega_l (levels 0 .. nlevels-1) and vector bits
   (level_vbits[l] for levels 0 .. nlevels, level_vbits[0] = 0) for h;
   cap: room in level_omega (level_vbits holds cap + 1).  Either may be NULL. */
int sc23ad_learned_levels(sc23ad_mlp* m, const sc23ad_hier* h, double* level_omega,
                          int* level_vbits, int cap, int* nlevels);
/* Order-free checksum used by the traces: sum_i mix64(i * 0x9E3779B97F4A7C15 ^
   (uint32)v_i) mod 2^64, mix64 = splitmix64's finaliser.  Host arrays. */
uint64_t sc23ad_checksum_i32(const int32_t* v, int n);

#ifdef __cplusplus
}
#endif

#endif /* SC23AD_H */
