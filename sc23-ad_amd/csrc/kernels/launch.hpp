// Interface between the host-side orchestration (compiled by g++) and the
// HIP kernels for gfx950 (compiled by hipcc).  Every launcher enqueues on the
// given stream and returns hipSuccess or the launch error.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace sc23ad::dev {

// Fixed blocked order of every FP reduction on the device: thread g of a
// kRedBlocks x kRedThreads grid sums elements g, g + kRedTotal, ... in
// increasing order, then a halving tree s[t] += s[t + stride] combines the
// threads of a block and, in the last block to finish, the block partials.
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
constexpr int kRedTotal = kRedBlocks * kRedThreads;

// Device view of a CSR matrix.  val points at int32 / float / double values.
constexpr int kSellWindow = 1024;
struct DCsr {
  int n = 0, m = 0, nnz = 0, bit = 0, dbit = -1;
  const int* rptr = nullptr;
  const int* cind = nullptr;
  const void* val = nullptr;
  int group = 8;  // lanes per row for the integer SpMV (power of two <= 64)
  // SELL-64 copy (rows with near-uniform lengths): slices of 64 rows, entry k
  // of row r of slice s at soff[s] + 64 k + r, padded with column -1 / value 0.
  // srow (optional): the slices hold rows srow[0], srow[1], ... (rows sorted
  // by length inside windows of kSellWindow rows), else rows in order.
  const long long* soff = nullptr;
  const int* scol = nullptr;
  const void* sval = nullptr;
  const int* srow = nullptr;
  // integer SELL-64 with 16-bit column offsets col - row (used instead of
  // scol when set; square matrices, no srow)
  const int16_t* scol16 = nullptr;
  // SELL workgroups numbered per XCD (each XCD a contiguous row range):
  // -1 = the SC23AD_XCD_REMAP default, 0 / 1 = chosen per matrix by the tuner
  int remap = -1;
};

// Dispatch-recorded timing.  When armed, the next launch by a timing-aware
// launcher on this thread (the SpMV launchers and the entry-parallel LD solve)
// goes through hipExtLaunchKernel with these events, so their interval is the
// kernel's own execution time, without the queue gap before it.  The launch
// disarms them; take_dispatch_events() returns and disarms what is left.
struct DispatchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
void arm_dispatch_events(hipEvent_t start, hipEvent_t stop);
DispatchEvents take_dispatch_events();

enum class Elem : int { kF64 = 0, kF32 = 1, kI32 = 2 };
enum class SpmvMode : int { kPlain = 0, kSub = 1, kAdd = 2 };
// Triangular factor layouts:
//   kLD  L + D/omega with the inverse diagonal stored last (ldtri)
//   kL   strictly lower, unit diagonal (ltri)
//   kDU  inverse diagonal first, then the upper entries; solved backward (dutri)
enum class TriKind : int { kLD = 0, kL = 1, kDU = 2 };

// y = spmv(x)  |  y = z - spmv(x)  |  y = z + spmv(x)   (y may alias z)
hipError_t launch_spmv(Elem e, const DCsr& a, SpmvMode mode, const void* x,
                       const void* z, void* y, hipStream_t s);

// Level-scheduled triangular solve.  `perm` lists the rows chain by chain and
// level by level; `lev_ptr` delimits the levels inside perm.
struct TriSegment {
  int chain_kernel;  // 1: one workgroup per chain walks levels [lev_begin_c, lev_end_c)
                     // 0: one level, rows perm[row_begin, row_end), many workgroups
  int nchain;
  const int* chain_lev;  // device, 2 * nchain ints: level range of each chain
  int row_begin, row_end;
  int group = 1;  // chain_kernel 0, integer factors: lanes per row (power of two <= 64)
};
hipError_t launch_trsv_segment(Elem e, TriKind kind, const DCsr& a,
                               const int* perm, const int* lev_ptr,
                               const TriSegment& seg, const void* in, void* out,
                               hipStream_t s);

// Level-ordered layout of a triangular factor (see k_tri_chain_pre): row
// lrow[idx] at position idx, its dependency entries lcol/lval[lbeg[idx] ..
// lbeg[idx+1]) in subtraction order, its inverse diagonal ldiag[idx].
struct TriLayout {
  const int* lrow = nullptr;
  const int* lbeg = nullptr;
  const int* lcol = nullptr;
  const void* lval = nullptr;
  const void* ldiag = nullptr;
  const int* ecol = nullptr;   // first 8 entries of each position, fixed stride 8
  const void* eval = nullptr;
  const int* lev_ptr = nullptr;
  // ring mode: entries as level-ordered positions instead of rows
  const int* lpos = nullptr;
  const int* epos = nullptr;
  int ring = 0;  // > 0: LDS ring of this many values (power of two)
  // entry-parallel mode (integer factors, factor numbered in level order):
  // packed entries {dep, val, slot, 0} (see k_tri_chain_ep), entry offset of
  // each level, waves per chain
  const void* lent = nullptr;
  const int* lev_eptr = nullptr;
  int ep_waves = 0;  // > 0: use the entry-parallel kernel
  int ep_far = 0;    // entries with dep < 0 (read from global memory) exist
  int ep_narrow = 0; // 4 entries and 1 row per lane in registers (else 8 and 2)
  int ep_maxlev = 0; // most levels in one chain
  int ep_copies = 1; // accumulator copies per row (1 or kEpCopies)
  int ep_acc = 0;    // accumulator slots in LDS
  // row-parallel mode (LD factors with short rows): rp_lanes lanes per row,
  // each holding rp_e entries (int2 {dep, val}, padded) of its row, rp_rows
  // rows per lane (one lane per row only), rp_waves waves per chain; same
  // chains, ring and far rule as the entry-parallel mode
  const void* rent = nullptr;
  int rp_waves = 0;  // > 0: use the row-parallel kernel
  int rp_e = 0;
  int rp_rows = 0;
  int rp_lanes = 1;
  int rp_farq = 0;   // most far entries in one row (stored first; one-lane layouts)
  int rp_stage = 0;  // FP rows over several lanes: products staged in LDS (else shuffled)
  // the block kernel's own row-parallel layout (four right-hand sides per
  // workgroup, four values per ring slot): entries as rent, ring slots, far
  // flag and most far entries per row, entries per lane, lanes per row,
  // rows per lane, waves
  const void* rent_blk = nullptr;
  int ring_blk = 0;
  int far_blk = 0;
  int farq_blk = 0;
  int e_blk = 0;
  int lanes_blk = 1;
  int rows_blk = 1;
  int waves_blk = 0;
  // FP row-parallel mode (one lane per row, rp_e entries in stored order):
  // dependency codes and values by position, and each row's entry count
  const int* rpf_dep = nullptr;
  const void* rpf_val = nullptr;
  const int* rpf_cnt = nullptr;
  int n = 0;
};
// Entry-parallel integer solve (one workgroup per chain).
constexpr int kEpEntriesPerLane = 8;  // entry-parallel kernel: entries held in registers
constexpr int kEpFarLevels = 40;  // least level distance of a dependency read from global memory
constexpr int kEpMaxRows = 1024;  // most rows in one level of one chain
// Rows with many entries: each row gets kEpCopies accumulators and lane t
// adds into copy t % kEpCopies, so the lanes of one instruction, which hold
// consecutive entries of few rows, collide kEpCopies times less in LDS.
constexpr int kEpCopies = 8;
hipError_t launch_trsv_ep(TriKind kind, const TriLayout& L, const TriSegment& seg, int bit,
                          int dbit, const void* in, void* out, hipStream_t s);
// Row-parallel integer LD solve (one workgroup per chain, 1-16 lanes per row).
constexpr int kRpMaxEntries = 12;  // one lane per row
constexpr int kRpMaxLanes = 16;    // several lanes per row: 8 or 12 entries each
hipError_t launch_trsv_rp(const TriLayout& L, const TriSegment& seg, int bit, int dbit,
                          const void* in, void* out, hipStream_t s);
hipError_t launch_trsv_rpf(Elem e, const TriLayout& L, const TriSegment& seg, const void* in,
                           void* out, hipStream_t s);
inline size_t rp_lds_bytes(int ring, int maxlev) {
  return sizeof(int32_t) * static_cast<size_t>(ring) +
         sizeof(int) * (static_cast<size_t>(maxlev) + 1);
}
// ---- block kernels (block.hip): several right-hand sides of one hierarchy --
// A block vector of n rows and `groups` groups of four right-hand sides holds
// element (row i, right-hand side r) at ((r / 4) * n + i) * 4 + r % 4.
// Least level distance of a dependency the block kernel reads from global
// memory (it fences every 8 levels; its far values are loaded at most 4
// levels ahead).
constexpr int kRpbFarLevels = 12;
// LDS of the FP several-lanes-per-row LD solve: the ring, each thread's e
// products (staged for the row's ordered subtraction), the level offsets.
inline size_t rpg_lds_bytes(size_t elem, int ring, int waves, int e, int maxlev, bool stage) {
  return elem * (static_cast<size_t>(ring) + (stage ? static_cast<size_t>(waves) * 64 * e : 0)) +
         sizeof(int) * (static_cast<size_t>(maxlev) + 1);
}
inline size_t rpb_lds_bytes(int ring, int maxlev) {
  return 16 * static_cast<size_t>(ring) + sizeof(int) * (static_cast<size_t>(maxlev) + 1);
}
hipError_t launch_permute_blk(int n, int groups, const int* perm, const void* src, void* dst,
                              bool gather, hipStream_t s);
hipError_t launch_copy_blk(int n, int groups, const void* src, void* dst, hipStream_t s);
// right-hand side r of a block <-> a plain int32 vector (get: block -> col)
hipError_t launch_blk_column(int n, int r, void* blk, void* col, bool get, hipStream_t s);
// y = spmv(x) | z - spmv(x) | z + spmv(x) for every right-hand side (y may alias z)
hipError_t launch_spmv_blk(const DCsr& a, int groups, SpmvMode mode, const void* x,
                           const void* z, void* y, hipStream_t s);
// Row-parallel LD solve of a block (the layouts of launch_trsv_rp that the
// block kernel has; see trsv_rp_blk_supported).
bool trsv_rp_blk_supported(const TriLayout& L);
hipError_t launch_trsv_rp_blk(const TriLayout& L, const TriSegment& seg, int groups, int bit,
                              int dbit, const void* in, void* out, hipStream_t s);
// Tiled coarse solve of a block: xp holds groups * n * 4 words, tickets two
// zero-initialised counters that the kernel leaves at zero.
hipError_t launch_trsv_tiles_blk(TriKind kind, int n, int nb, int groups, const int32_t* tiles,
                                 const long long* toff, const int32_t* dinv, int bit, int dbit,
                                 const void* in, void* out, unsigned long long* xp, int epoch,
                                 int* err, unsigned* tickets, hipStream_t s);

// LDS bytes of the entry-parallel kernel: acc_slots uint64 accumulators and
// kEpDummy spare ones, the ring, the chain's level offsets.
constexpr int kEpDummy = 64;
inline size_t ep_lds_bytes(int ring, int maxlev, int acc_slots) {
  return sizeof(uint64_t) * (static_cast<size_t>(acc_slots) + kEpDummy) +
         sizeof(int32_t) * static_cast<size_t>(ring) +
         2 * sizeof(int) * (static_cast<size_t>(maxlev) + 1);
}
hipError_t launch_trsv_layout(Elem e, TriKind kind, const TriLayout& L,
                              const TriSegment& seg, int bit, int dbit, const void* in,
                              void* out, hipStream_t s);

// Blocked single-workgroup solve of a whole triangular factor with its
// solution and accumulators in LDS (the coarsest grid).  bptr[i*(nb+1)+b] is
// the first dependency entry of row i with column >= 64 b; nb = ceil(n/64).
constexpr int kTriBlockWidth = 64;
constexpr size_t kTriBlockedMaxLds = 160 * 1024;
size_t tri_blocked_lds(int n, size_t elem, size_t acc);
// Integer factor stored as dense 64-wide slices per column block (see
// k_tri_dense_int); dinv[i] is row i's stored inverse diagonal (LD / DU).
// Tiled integer solve of a coarse factor, one workgroup per 64-row block
// (see k_tri_tiles_int); xp: n words {epoch, value} publishing the solved
// rows, epoch: a value no word holds yet, err: incremented when a wait gives
// up.
constexpr int kTileMaxBlocks = 64;
hipError_t launch_trsv_tiles(TriKind kind, int n, int nb, const int32_t* tiles,
                             const long long* toff, const int32_t* dinv, int bit, int dbit,
                             const void* in, void* out, unsigned long long* xp, int epoch,
                             int* err, unsigned* tickets, hipStream_t s);
// Tiled FP coarse solve (one wave per 64-row block, each row's terms in the
// sequential order): tiles as for launch_trsv_tiles with FP values, tmask
// one 64-bit mask of present entries per tile row, flags nb words raised to
// epoch when a block is done (see k_tri_tiles_fp).
hipError_t launch_trsv_tiles_fp(Elem e, TriKind kind, int n, int nb, const void* tiles,
                                const long long* toff, const uint64_t* tmask, const void* dinv,
                                const void* in, void* out, unsigned long long* flags, int epoch,
                                int* err, unsigned* tickets, hipStream_t s);
hipError_t launch_trsv_blocked(Elem e, TriKind kind, const DCsr& a, const int* bptr,
                               int nb, const void* in, void* out, hipStream_t s);

// ---- AMG setup (setup.hip) -----------------------------------------------
// Classical interpolation weights of the F rows `fine` into pval (P's CSR
// values, row i starting at pptr[i]); status[0] |= 1 where an F-F connection
// shares no C point, status[1] counts the amk = 0 substitutions.
hipError_t launch_interp_weights(int nfine, const int* fine, const int* arp, const int* aci,
                                 const double* av, const int* csp, const int* csc,
                                 const double* csv, const int* dsp, const int* dsc,
                                 const double* dsv, const int* wkp, const double* wkv,
                                 const int* pptr, double* pval, int* status, hipStream_t s);
// Row-merge product rows [row0, row1): row i hashes its terms in
// tkey/tval[toff[i] - base_t ...) (toff[i+1] - toff[i] slots) and writes its
// surviving entries, sorted, to ocol/oval[ooff[i] - base_o ...) and their
// number to cnt[i - row0].
hipError_t launch_rowmerge(int row0, int row1, const int* lp, const int* lc, const double* lv,
                           const int* rp, const int* rc, const double* rv,
                           const long long* toff, const long long* ooff, long long base_t,
                           long long base_o, int* tkey, double* tval, int* ocol, double* oval,
                           int* cnt, hipStream_t s);
// Copies rows [row0, row1) of the compacted output to
// cind/val[optr[i - row0] ...).
hipError_t launch_rowmerge_gather(int row0, int row1, const long long* ooff, long long base_o,
                                  const int* ocol, const double* oval, const int* optr,
                                  int* cind, double* val, hipStream_t s);

// ---- Krylov kernels ------------------------------------------------------
// Scalars and small arrays of one FGMRES(m) solver, resident on the device.
struct KState {
  int m;
  int bit;          // fractional bits F (integer solver)
  double nrm_b;
  double eps;
  // integer Arnoldi (int64) or FP64 / FP32 Arnoldi (stored as double / float)
  void* H;          // (m+1) * m, column j at offset j*(m+1)
  void* c;          // m
  void* s;          // m
  void* e;          // m+1
  double* y;        // m (int / f64) ; float solver keeps y as float in yf
  float* yf;
  // reductions
  int64_t* ipart;   // 3 x kRedBlocks partial int64 sums (two alternate, one spare)
  double* dpart;    // 2 x kRedBlocks partial double sums
  float* fpart;     // 2 x kRedBlocks partial float sums
  unsigned* counter;  // unused (kept for layout)
  // scalars
  double* scal;     // [0] nrm_r, [1] prev, [2] e0, [3] e_inv, [4] hn (fp), [5] implicit res
  int64_t* iscal;   // [0] criteria, [1] hinv, [2] hn
  int* flags;       // [0] converged, [1] overflow, [2] stagnated, [3] steps done
  uint64_t* trace;  // checksum trace buffer (optional)
  int trace_cap;
  int* trace_len;   // device counter
};

// Outer residual r = b - A x (FP64, row sums in column order), its norm in the
// fixed blocked order (seq_red: strictly left to right, one thread), the
// Restarted stagnation test and the FGMRES start scalars for precision e.
hipError_t launch_outer_residual(const DCsr& a64, const double* b,
                                 const double* x, double* r, const KState& k,
                                 Elem e, int seq_red, hipStream_t s);
// V_0 from r: int: trunc(r * e_inv * 2^F); f64: r * e_inv; f32: (float)(r*e_inv)
hipError_t launch_v0(Elem e, int n, const double* r, void* v0, const KState& k,
                     hipStream_t s);
// H(0,j) = <V_0, w>  (int: >> F)
hipError_t launch_dot_first(Elem e, int n, const void* v0, const void* w,
                            const KState& k, int j, int seq_red, hipStream_t s);
// w += -H(k,j) V_k, then H(k+1,j) = <V_{k+1}, w> (k < j) or, for k == j, the
// norm of w, H_inv / 1/h, the Givens rotations, e and the convergence flag.
hipError_t launch_axpy_dot(Elem e, int n, const void* V, void* w, const KState& k,
                           int kk, int j, int seq_red, hipStream_t s);
// h = norm of w from the last pass's partials; w = H_inv w (int) or
// w *= 1/h (fp); the Givens rotations, e and the convergence flag of step j.
hipError_t launch_scal(Elem e, int n, void* w, const KState& k, int j, int seq_red,
                       hipStream_t s);
// y = H^-1 e (FP64 / float as the reference) over the first `steps`
// columns, then x += sum_k y_k Z_k in order of k.
hipError_t launch_update_x(Elem e, int n, const void* Z, double* x,
                           const KState& k, int steps, hipStream_t s);
// order-free checksum of an int32 vector appended to k.trace
hipError_t launch_checksum(int n, const int32_t* v, const KState& k,
                           hipStream_t s);
// explicit residual ||b - A x|| in the blocked order into *out (device)
hipError_t launch_relres(const DCsr& a64, const double* b, const double* x,
                         double* r, double* part, unsigned* counter,
                         double* out, hipStream_t s);
// ---- block Krylov kernels (several solves on one hierarchy, in lockstep) ---
// ks: one KState per right-hand side (4 * groups of them, in device memory);
// live[r] != 0: solve r takes part in the current cycle; upd[r] != 0: its x
// is updated at the end of the cycle over the steps it ran (flags[3]).
// Block vectors as in block.hip; b, x, r: `count` FP64 columns of n.
hipError_t launch_outer_residual_blk(const DCsr& a64, int count, const double* b,
                                     const double* x, double* r, const KState* ks,
                                     const int* live, hipStream_t s);
hipError_t launch_v0_blk(int n, int groups, const double* r, void* v0, const KState* ks,
                         const int* live, hipStream_t s);
hipError_t launch_dot_first_blk(int n, int groups, const void* v0, const void* w,
                                const KState* ks, hipStream_t s);
hipError_t launch_axpy_dot_blk(int n, int groups, const void* V, void* w, const KState* ks,
                               const int* live, int kk, int j, hipStream_t s);
// The whole Gram-Schmidt step j (dot_first, axpy_dot kk = 0..j, scal) in
// one launch of one workgroup per group: for systems of at most
// kMgsFusedRows rows.
constexpr int kMgsFusedRows = 8192;
hipError_t launch_mgs_step_blk(int n, int groups, const void* V, void* w, const KState* ks,
                               const int* live, int j, hipStream_t s);
hipError_t launch_scal_blk(int n, int groups, void* w, const KState* ks, const int* live, int j,
                           hipStream_t s);
hipError_t launch_update_x_blk(int n, int groups, int count, const void* Z, double* x,
                               const KState* ks, const int* upd, hipStream_t s);
// flags[0..3] and the implicit residual (a double in words 4-5) of each solve
hipError_t launch_flags_blk(const KState* ks, int count, int* out, hipStream_t s);
hipError_t launch_checksum_blk(int n, int groups, int count, const void* v, const KState* ks,
                               const int* live, hipStream_t s);

// Device scratch of the allocation-free level operation (noalloc.hip): the
// current device's address of a file-scope array of kNoAllocScratchBytes
// bytes (two staging vectors of half that size each), looked up once per
// device.  Every call on a device shares it.
constexpr size_t kNoAllocScratchBytes = size_t{16} << 20;
hipError_t noalloc_scratch(void** ptr);

// V-cycle vector helpers
hipError_t launch_copy(Elem e, int n, const void* src, void* dst, hipStream_t s);
// gather: dst[k] = src[perm[k]]; otherwise scatter: dst[perm[k]] = src[k].
// out[k] += dinv[k] * t[k] (FP; the Jacobi smoother's update).
hipError_t launch_jacobi_update(Elem e, int n, const void* dinv, const void* t, void* out,
                                hipStream_t s);
hipError_t launch_permute(Elem e, int n, const int* perm, const void* src, void* dst,
                          bool gather, hipStream_t s);
hipError_t launch_fill_zero(size_t bytes, void* dst, hipStream_t s);

// Learned-parameter network (mlp.hip): parameters P (mlp_param_count()
// floats), inputs X (B x 8), targets T and outputs Y (B x 2).
int mlp_param_count();
hipError_t launch_mlp_forward(int B, const float* X, const float* P, float* Y, hipStream_t s);
hipError_t launch_mlp_grad(int B, const float* X, const float* T, const float* P, float* G,
                           float* loss, hipStream_t s);
hipError_t launch_adam(int n, float* P, const float* G, float* M, float* V, float lr, int step,
                       hipStream_t s);

}  // namespace sc23ad::dev
