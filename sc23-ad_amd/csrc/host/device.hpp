// Device-resident hierarchy and solver: the host-side sequencing of the HIP
// kernels for one V-cycle (intAMGflex / dfAMG::Precondition,
// Library/preconditioner/INTAMG.hpp:108-133,269-293) and for
// Restarted(FGMRES(m)) (Library/solver/irsolver.hpp:27-43,
// Library/solver/gmres.hpp:48-88,130-170,225-271).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "../kernels/launch.hpp"
#include "hierarchy.hpp"

namespace sc23ad {

struct DevError {
  int code;
  std::string what;
};

void hip_check(hipError_t e, const char* what);
// The stream library work goes to: the calling thread's current stream (see
// StreamScope), else the library's default stream on device 0.
hipStream_t lib_stream();
// Makes `s` the calling thread's current stream for the scope's lifetime.
class StreamScope {
 public:
  explicit StreamScope(hipStream_t s);
  ~StreamScope();
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;

 private:
  hipStream_t prev_;
};

class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  void alloc(size_t bytes);
  void release();
  void upload(const void* src, size_t bytes, size_t off = 0);
  void download(void* dst, size_t bytes, size_t off = 0) const;
  template <typename T>
  T* as() const { return static_cast<T*>(p_); }
  void* get() const { return p_; }
  size_t bytes() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

struct DevMat {
  dev::DCsr d;
  DevBuf rptr, cind, val;
  DevBuf soff, scol, sval, srow;  // SELL-64 copy (optional; srow: row order)
  DevBuf scol16;                   // 16-bit column offsets of the SELL-64 copy
  // second SELL-64 copy of a rectangular matrix, rows ordered by their first
  // column (gathers of neighbouring rows share cache lines); d2 describes it
  DevBuf soff2, scol2, sval2, srow2;
  dev::DCsr d2;
  bool has2 = false;
  size_t algo_bytes_spmv(size_t elem) const;
};

// Level-set schedule of one triangular factor.
struct TriPlan {
  dev::TriKind kind = dev::TriKind::kLD;
  bool identity = false;  // level order == row order (the matrix was renumbered)
  DevBuf perm, lev_ptr, chain_lev;
  std::vector<dev::TriSegment> segs;
  int nlevels = 0;    // longest dependency chain (levels of the deepest chain)
  int nlaunch = 0;
  // blocked single-workgroup solve (coarsest grid)
  bool blocked = false;
  DevBuf bptr;
  int nb = 0;
  // 64 x 64 tiles, one workgroup per block (integer coarse factors); the
  // hand-off flags live with the caller's work context
  bool dense = false;
  DevBuf slice, boff, ddinv;
  bool dense_fp = false;  // FP tiles (slice / boff / ddinv as FP) with present-entry masks
  DevBuf tmask;
  // level-ordered layout used by the level-scheduled kernels
  DevBuf lrow, lbeg, lcol, lval, ldiag, ecol, eval, lpos, epos, lent, lev_eptr, rent;
  DevBuf rpf_dep, rpf_val, rpf_cnt;  // FP row-parallel layout
  DevBuf rent_blk;                    // row-parallel entries for the block kernel's ring
  dev::TriLayout layout;
};

// hipEvent pairs timing launches on the library stream: bound to the kernel's
// own dispatch when the launcher supports it (the kernel's execution time),
// else recorded around the launch on the stream (stream_timed counts those).
// Elapsed times are read by collect() once the stream has been synchronised.
struct EventTimer {
  bool on = false;
  int launches = 0;
  int stream_timed = 0;
  double ms_total = 0;
  // aware: the next launch binds the armed events to its dispatch (no
  // marker on the stream before it); else the stream interval is timed.
  void begin(bool aware);
  void end();
  void collect();
  void reset();
  ~EventTimer();

 private:
  std::vector<hipEvent_t> ev;
  bool armed_aware = false;
  int used = 0;
};

// One level on the device.  Its unknowns are numbered in the smoother's
// level order (position in the LD schedule); perm[new] = original row.
struct DevLevel {
  DevMat a, ld, u, r, p;
  TriPlan ld_plan;
  DevBuf perm;
  int n = 0;
};

// Launch-shape knobs (chosen by the autotuner, or defaults).
struct LaunchCfg {
  int single_wg_max_rows = 1024;  // levels up to this size stay in one workgroup
  int spmv_group = 0;             // 0: choose from nnz/row
};

class DevAmg {
 public:
  dev::Elem elem = dev::Elem::kI32;
  int depth = 0;
  int iter = 1;
  std::vector<DevLevel> lv;
  DevMat lc, uc;
  TriPlan lc_plan, uc_plan;
  int n_coarse = 0;
  bool jacobi = false;   // Jacobi smoother: LD holds d_inv, U is A
  DevMat a0;             // fine-level operator in the original numbering (Arnoldi)
  LaunchCfg cfg;

  static std::unique_ptr<DevAmg> upload(const IntHierarchy& h, const LaunchCfg& c);
  template <typename T>
  static std::unique_ptr<DevAmg> upload_fp(const FpHierarchy<T>& h, const LaunchCfg& c);

  size_t esize() const { return elem == dev::Elem::kF64 ? 8 : 4; }
  int n0() const { return lv.empty() ? 0 : lv[0].n; }
  int level_n(int l) const { return l < depth ? lv[l].n : n_coarse; }

  // Work vectors and coarse-solve hand-off flags of one V-cycle in flight.
  // Every concurrent user of the hierarchy (a solver, a batch member) owns
  // one; the hierarchy's own serves the preconditioner and level-op calls.
  struct Work {
    std::vector<DevBuf> ee, xx, rr, tmp;  // index = level (1..depth); tmp[0] fine
    DevBuf r0, tmpc, vin0, vout0;         // fine residual, coarse temp, level-0 order in/out
    DevBuf lflags, uflags, err;           // tiled coarse solves: {epoch, value} words, errors
    DevBuf tickets;                       // tiled coarse solves: L counters [0,1], U [2,3]
    int lepoch = 0, uepoch = 0;           // last flag values raised
    EventTimer* trsv_timer = nullptr;     // times the level-0 LD solves (owner's hook)
  };
  std::unique_ptr<Work> make_work() const;
  Work& own_work() const { return *work_; }

  void spmv(const DevMat& m, dev::SpmvMode mode, const void* x, const void* z,
            void* y) const;
  bool trsv_timing_aware(const TriPlan& plan) const;
  void trsv(const TriPlan& plan, const DevMat& m, const void* in, void* out,
            unsigned long long* flags, int* epoch, int* err, unsigned* tickets) const;
  void smooth(int l, const void* in, void* out, bool inited, Work& w) const;
  void coarse(const void* in, void* out, Work& w) const;
  // One V-cycle on vectors in the original numbering.
  void vcycle(const void* in, void* out, Work& w) const;
  // Level operation (sc23ad_level_op) on vectors in the original numbering;
  // ti / to: staging vectors of the operand and the result in the level's
  // numbering (at least es * n bytes each; nullptr: allocated here).
  void level_op(int level, int op, const void* in, void* out, Work& w,
                void* ti = nullptr, void* to = nullptr) const;
  // Algorithmic HBM bytes of one V-cycle (sum over its kernels).
  double vcycle_bytes() const;
  // Throws if a kernel with a bounded wait gave up (after a stream sync);
  // the error word is cleared first, so only the call that timed out fails.
  void check_errors(Work& w) const;

  // ---- block V-cycle: 4 * groups right-hand sides at once (integer only) ---
  // Vectors hold element (row i, right-hand side r) at ((r / 4) * n + i) * 4
  // + r % 4 (see block.hip).  One BlockWork per concurrent block user.
  struct BlockWork {
    int groups = 0;
    std::vector<DevBuf> ee, xx, rr, tmp;   // per level (1..depth); tmp[0] fine
    DevBuf r0, vin0, vout0, tmpc;
    DevBuf lflags, uflags, tickets, err;   // tiled coarse solves
    int lepoch = 0, uepoch = 0;
    DevBuf col_in, col_out;                // one right-hand side (per-column fallback)
  };
  std::unique_ptr<BlockWork> make_block_work(int groups) const;
  void vcycle_blk(const void* in, void* out, BlockWork& w) const;
  void spmv_blk(const DevMat& m, dev::SpmvMode mode, const void* x, const void* z, void* y,
                int groups) const;
  // Whether every LD solve of the hierarchy has a block kernel (else those
  // levels solve one right-hand side at a time).
  bool block_native() const;
  void check_errors(BlockWork& w) const;

 private:
  std::unique_ptr<Work> work_;
  void tune();  // SpMV launch shapes per matrix
  void vcycle_lo(const void* in, void* out, Work& w) const;  // vectors in level order
  void vcycle_lo_blk(const void* in, void* out, BlockWork& w) const;
  void smooth_blk(int l, const void* in, void* out, bool inited, BlockWork& w) const;
  void trsv_blk(const TriPlan& plan, const DevMat& m, const void* in, void* out, BlockWork& w,
                unsigned long long* flags, int* epoch, unsigned* tickets) const;
  void coarse_blk(const void* in, void* out, BlockWork& w) const;
  const int* perm_of(int level) const {  // nullptr: original order (coarsest grid)
    return level < depth ? lv[level].perm.as<int>() : nullptr;
  }
};

// FP64 copy of the fine operator for the outer residual (the reference's dA).
void upload_fp64_operator(DevMat& dm, const HostCsr& a);

struct SolverOpts {
  int seq_reduction = 0;  // 1: FP reductions strictly left to right
  // > 0: stop with STAGNATION once the implicit residual has not fallen
  // below its best value by a relative 1e-3 for this many consecutive
  // Arnoldi steps (the learned setup's fallback trigger); 0: never
  int stall_steps = 0;
};

// Stall test of one solve: steps since the implicit residual last fell below
// its best value by a relative kStallGain.
struct StallWatch {
  static constexpr double kStallGain = 1e-3;
  double best = 0;
  int since = 0;
  bool started = false;
  void reset() { started = false; since = 0; }
  // true when the solve has stalled for `steps` steps (steps <= 0: never)
  bool step(double rel, int steps) {
    if (!started || rel < best * (1.0 - kStallGain)) {
      best = started ? std::min(best, rel) : rel;
      started = true;
      since = 0;
      return false;
    }
    return steps > 0 && ++since >= steps;
  }
};

class DevSolver {
 public:
  dev::Elem elem;
  int m = 30;
  double eps = 1e-10;
  int bit = 29;
  int max_outer = 33;
  DevAmg* amg = nullptr;  // not owned
  DevMat a64;             // FP64 fine operator (outer residual)
  SolverOpts opts;
  DevBuf V, Z, r, b, x, state, pinned_dummy;
  hipStream_t stream = nullptr;             // this solver's stream (default priority)
  std::unique_ptr<DevAmg::Work> work;       // its own V-cycle work context
  dev::KState k{};
  int* host_flags = nullptr;  // pinned: flags[0..3], then the implicit residual
  // tracing / timing
  bool tracing = false;
  int trace_cap = 0;
  EventTimer spmv_timer;  // fine-level A SpMV of the Arnoldi steps (benchmark hook)
  EventTimer trsv0_timer; // level-0 LD solves of this solver's V-cycles (benchmark hook)
  std::vector<double> history;          // |e_{j+1}| / nrm_b per Arnoldi step

  DevSolver(dev::Elem e, const HostCsr& a, DevAmg* amg, int m, double eps,
            int bit, int max_outer);
  ~DevSolver();
  struct Result {
    int converged = 0, iter = 0, restarts = 0, status = 0, overflow = 0;
    double implicit_relres = 0, solve_ms = 0;
  };
  Result solve(const double* b_dev, double nrm_b, double* x_dev);
  double relres(const double* b_dev, double nrm_b, const double* x_dev);
  void set_trace(int cap);
  void read_trace(uint64_t* out, int cap, int* len);

 private:
  void read_flags();
};

// k integer solves on one hierarchy in lockstep: Restarted(FGMRESflex(m)) for
// every right-hand side (irsolver.hpp:27-43, gmres.hpp:225-271), with the
// V-cycles, the fine-level SpMV and the Gram-Schmidt passes of all of them
// in shared launches.  Each right-hand side's arithmetic is its single
// solve's, so its results are identical to DevSolver's.
class DevBlockSolver {
 public:
  int count = 0;   // right-hand sides
  int groups = 0;  // ceil(count / 4)
  int m = 30;
  double eps = 1e-10;
  int max_outer = 33;
  std::vector<int> bits;  // F per right-hand side
  DevAmg* amg = nullptr;  // not owned
  DevMat a64;
  DevBuf V, Z, r, states, kst, live, upd, flagbuf, trace_buf;
  hipStream_t stream = nullptr;
  std::unique_ptr<DevAmg::BlockWork> work;
  std::vector<dev::KState> ks;
  int* host_flags = nullptr;  // pinned: 6 words per right-hand side
  int* host_mask = nullptr;   // pinned: live then upd masks
  int stall_steps = 0;        // see SolverOpts::stall_steps
  bool tracing = false;
  int trace_cap = 0;
  std::vector<std::vector<double>> history;

  DevBlockSolver(const HostCsr& a, DevAmg* amg, int count, int m, double eps,
                 const int* bits, int max_outer);
  ~DevBlockSolver();
  // b, x: `count` FP64 columns of n on the device; res: count records
  void solve(const double* b_dev, const double* nrm_b, double* x_dev, DevSolver::Result* res);
  void set_trace(int cap);
  void read_trace(int rhs, uint64_t* out, int cap, int* len);

 private:
  void read_flags();
  void put_masks(const std::vector<int>& live_v, const std::vector<int>& upd_v);
};

}  // namespace sc23ad
