// Device half of the C ABI: device buffers, the raw integer SpMV, the
// device-resident preconditioner and the solver.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <memory>
#include <new>

#include "../../../include/sc23ad.h"
#include "capi_internal.hpp"
#include "device.hpp"
#include "learn.hpp"

using namespace sc23ad;

struct sc23ad_amg_s {
  int precision = 2;
  std::unique_ptr<DevAmg> d;
};

struct sc23ad_mlp_s {
  DevBuf P, G, M, V, loss;
  int step = 0;
};

struct sc23ad_solver_s {
  int precision = 2;
  std::unique_ptr<DevSolver> s;
  DevBuf hb, hx;  // staging for host-pointer solves
};

struct sc23ad_block_solver_s {
  std::unique_ptr<DevBlockSolver> s;
  DevBuf hb, hx;  // staging for host-pointer solves
};

namespace {

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const DevError& e) {
    return fail(e.code, e.what);
  } catch (const SetupError& e) {
    return fail(e.code, e.what);
  } catch (const std::bad_alloc&) {
    return fail(SC23AD_HIP_ERROR, "out of host memory");
  }
}

dev::Elem elem_of(int precision) {
  return precision == SC23AD_INT ? dev::Elem::kI32
         : precision == SC23AD_FP64 ? dev::Elem::kF64
                                    : dev::Elem::kF32;
}

}  // namespace

extern "C" {

int sc23ad_device_count(int* count) {
  if (!count) return fail(SC23AD_BAD_PARAM, "device_count: null");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return SC23AD_OK;
}

int sc23ad_hier_build_multi_dev(const sc23ad_csrmat* a, const sc23ad_amg_param* p,
                                const int* precisions, int count, sc23ad_hier** outs) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return fail(SC23AD_NO_DEVICE, "hier_build_multi_dev: no GPU");
  return build_multi_impl(a, p, precisions, count, outs, true);
}

int sc23ad_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(SC23AD_BAD_PARAM, "dev_alloc: null");
  return guarded([&]() -> int {
    lib_stream();
    hip_check(hipMalloc(ptr, bytes ? bytes : 4), "hipMalloc");
    return SC23AD_OK;
  });
}

int sc23ad_dev_free(void* ptr) {
  if (ptr) (void)hipFree(ptr);
  return SC23AD_OK;
}

int sc23ad_dev_upload(void* dst, const void* src, size_t bytes) {
  return guarded([&]() -> int {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, lib_stream()), "upload");
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    return SC23AD_OK;
  });
}

int sc23ad_dev_download(void* dst, const void* src, size_t bytes) {
  return guarded([&]() -> int {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, lib_stream()), "download");
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    return SC23AD_OK;
  });
}

int sc23ad_dev_copy(void* dst, const void* src, size_t bytes) {
  if ((bytes & 3) != 0 || (!dst && bytes) || (!src && bytes))
    return fail(SC23AD_BAD_PARAM, "dev_copy: bytes must be a multiple of 4");
  return guarded([&]() -> int {
    hip_check(dev::launch_copy(dev::Elem::kI32, static_cast<int>(bytes / 4), src, dst,
                               lib_stream()),
              "copy");
    return SC23AD_OK;
  });
}

int sc23ad_dev_sync(void) {
  return guarded([&]() -> int {
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    return SC23AD_OK;
  });
}

int sc23ad_spmv_int(int n, const int* rptr, const int* cind, const int32_t* val,
                    int bit, const int32_t* x, int32_t* y) {
  if (n < 0 || bit < 0 || bit > 62 || (n > 0 && (!rptr || !cind || !val || !x || !y)))
    return fail(SC23AD_BAD_PARAM, "spmv_int: bad arguments");
  return guarded([&]() -> int {
    dev::DCsr a;
    a.n = n;
    a.rptr = rptr;
    a.cind = cind;
    a.val = val;
    a.bit = bit;
    a.group = 8;
    hip_check(dev::launch_spmv(dev::Elem::kI32, a, dev::SpmvMode::kPlain, x, nullptr, y,
                               lib_stream()),
              "spmv_int");
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    return SC23AD_OK;
  });
}

int sc23ad_amg_create(const sc23ad_hier* h, sc23ad_amg** out) {
  if (!h || !out) return fail(SC23AD_BAD_PARAM, "amg_create: bad arguments");
  return guarded([&]() -> int {
    auto a = std::make_unique<sc23ad_amg_s>();
    a->precision = h->precision;
    LaunchCfg cfg;
    if (h->precision == SC23AD_INT)
      a->d = DevAmg::upload(h->ih, cfg);
    else if (h->precision == SC23AD_FP64)
      a->d = DevAmg::upload_fp(h->dh, cfg);
    else
      a->d = DevAmg::upload_fp(h->fh, cfg);
    *out = a.release();
    return SC23AD_OK;
  });
}

void sc23ad_amg_destroy(sc23ad_amg* m) { delete m; }

int sc23ad_amg_precision(const sc23ad_amg* m, int* precision) {
  if (!m || !precision) return fail(SC23AD_BAD_PARAM, "amg_precision: null");
  *precision = m->precision;
  return SC23AD_OK;
}

int sc23ad_amg_precondition(sc23ad_amg* m, const void* in, void* out, int on_device) {
  if (!m || !in || !out) return fail(SC23AD_BAD_PARAM, "precondition: null");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    const size_t bytes = a.esize() * static_cast<size_t>(a.n0());
    if (on_device) {
      a.vcycle(in, out, a.own_work());
    } else {
      DevBuf di(bytes), dout(bytes);
      di.upload(in, bytes);
      a.vcycle(di.get(), dout.get(), a.own_work());
      dout.download(out, bytes);
    }
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    a.check_errors(a.own_work());
    return SC23AD_OK;
  });
}

int sc23ad_amg_vcycle_bytes(const sc23ad_amg* m, double* bytes) {
  if (!m || !bytes) return fail(SC23AD_BAD_PARAM, "vcycle_bytes: null");
  *bytes = m->d->vcycle_bytes();
  return SC23AD_OK;
}

int sc23ad_amg_time_vcycles(sc23ad_amg* m, int reps, const void* in, void* out,
                            double* ms_per_cycle) {
  if (!m || !in || !out || !ms_per_cycle || reps < 1)
    return fail(SC23AD_BAD_PARAM, "time_vcycles: bad arguments");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    a.vcycle(in, out, a.own_work());  // warm-up
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, lib_stream()), "record");
    for (int r = 0; r < reps; ++r) a.vcycle(in, out, a.own_work());
    hip_check(hipEventRecord(e1, lib_stream()), "record");
    hip_check(hipEventSynchronize(e1), "sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_cycle = ms / reps;
    a.check_errors(a.own_work());
    return SC23AD_OK;
  });
}

int sc23ad_amg_level_op(sc23ad_amg* m, int level, int op, const void* in, void* out) {
  if (!m || !in || !out) return fail(SC23AD_BAD_PARAM, "level_op: null");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    if (op == SC23AD_OP_COARSE) {
      if (level != a.depth) return fail(SC23AD_BAD_PARAM, "level_op: coarse solve is level depth");
    } else {
      if (level < 0 || level >= a.depth) return fail(SC23AD_BAD_PARAM, "level_op: bad level");
      if (op < SC23AD_OP_SPMV_A || op > SC23AD_OP_SMOOTH1)
        return fail(SC23AD_BAD_PARAM, "level_op: bad op");
    }
    a.level_op(level, op, in, out, a.own_work());
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    a.check_errors(a.own_work());
    return SC23AD_OK;
  });
}

int sc23ad_amg_level_op_noalloc(sc23ad_amg* m, int level, int op, const void* in, void* out) {
  if (!m || !in || !out) return fail(SC23AD_BAD_PARAM, "level_op_noalloc: null");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    if (op == SC23AD_OP_COARSE) {
      if (level != a.depth)
        return fail(SC23AD_BAD_PARAM, "level_op_noalloc: coarse solve is level depth");
    } else {
      if (level < 0 || level >= a.depth)
        return fail(SC23AD_BAD_PARAM, "level_op_noalloc: bad level");
      if (op < SC23AD_OP_SPMV_A || op > SC23AD_OP_SMOOTH1)
        return fail(SC23AD_BAD_PARAM, "level_op_noalloc: bad op");
    }
    const int nin = op == SC23AD_OP_SPMV_P ? a.level_n(level + 1) : a.level_n(level);
    const int nout = op == SC23AD_OP_SPMV_R ? a.level_n(level + 1) : a.level_n(level);
    const size_t half = dev::kNoAllocScratchBytes / 2;
    const size_t need = a.esize() * static_cast<size_t>(std::max(nin, nout));
    if (need > half)
      return fail(SC23AD_BAD_PARAM,
                  "level_op_noalloc: a level vector of " + std::to_string(need) +
                      " bytes exceeds the static scratch's " + std::to_string(half) +
                      " bytes per vector; use sc23ad_amg_level_op");
    lib_stream();  // the device is current
    void* scratch = nullptr;
    hip_check(dev::noalloc_scratch(&scratch), "static scratch address");
    char* base = static_cast<char*>(scratch);
    a.level_op(level, op, in, out, a.own_work(), base, base + half);
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    a.check_errors(a.own_work());
    return SC23AD_OK;
  });
}

int sc23ad_solver_create(int precision, const sc23ad_csrmat* a, sc23ad_amg* m,
                         int restart_m, double eps, int frac_bits, int max_restarts,
                         sc23ad_solver** out) {
  if (!a || !m || !out || restart_m < 1 || !(eps > 0))
    return fail(SC23AD_BAD_PARAM, "solver_create: bad arguments");
  if (precision != m->precision)
    return fail(SC23AD_BAD_PARAM, "solver_create: precision differs from the amg's");
  if (precision == SC23AD_INT && (frac_bits < 1 || frac_bits > 30))
    return fail(SC23AD_BAD_PARAM, "solver_create: frac_bits must be in [1, 30]");
  if (a->a.n != m->d->n0()) return fail(SC23AD_BAD_PARAM, "solver_create: size mismatch");
  return guarded([&]() -> int {
    auto s = std::make_unique<sc23ad_solver_s>();
    s->precision = precision;
    const int outer = max_restarts > 0 ? max_restarts : 1000 / restart_m;
    s->s = std::make_unique<DevSolver>(elem_of(precision), a->a, m->d.get(), restart_m,
                                       eps, frac_bits, outer);
    *out = s.release();
    return SC23AD_OK;
  });
}

void sc23ad_solver_destroy(sc23ad_solver* s) { delete s; }

namespace {

// One solve; host pointers are staged through the solver's buffers on its
// own stream, so that solves of different solvers can run concurrently.
int solve_one(sc23ad_solver* s, const double* b, double nrm_b, double* x, int on_device,
              sc23ad_converge* out) {
  if (!s || !b || !x) return fail(SC23AD_BAD_PARAM, "solve: null");
  return guarded([&]() -> int {
    DevSolver& d = *s->s;
    StreamScope scope(d.stream);
    const int n = d.amg->n0();
    const size_t bytes = sizeof(double) * n;
    DevSolver::Result r;
    if (on_device) {
      r = d.solve(b, nrm_b, x);
    } else {
      if (s->hb.bytes() < bytes) s->hb.alloc(bytes);
      if (s->hx.bytes() < bytes) s->hx.alloc(bytes);
      s->hb.upload(b, bytes);
      s->hx.upload(x, bytes);
      r = d.solve(s->hb.as<double>(), nrm_b, s->hx.as<double>());
      s->hx.download(x, bytes);
    }
    if (out) {
      memset(out, 0, sizeof *out);
      out->is_converged = r.converged;
      out->iter = r.iter;
      out->restarts = r.restarts;
      out->status = r.status;
      out->overflow_flag = r.overflow;
      out->frac_bits = d.bit;
      out->implicit_relres = r.implicit_relres;
      out->solve_ms = r.solve_ms;
    }
    return SC23AD_OK;
  });
}

}  // namespace

int sc23ad_solver_solve(sc23ad_solver* s, const double* b, double nrm_b, double* x,
                        int on_device, sc23ad_converge* out) {
  return solve_one(s, b, nrm_b, x, on_device, out);
}

int sc23ad_solve_batch(int count, sc23ad_solver* const* solvers, const double* const* b,
                       const double* nrm_b, double* const* x, int on_device, int threads,
                       sc23ad_converge* out) {
  if (count < 0 || (count > 0 && (!solvers || !b || !nrm_b || !x)))
    return fail(SC23AD_BAD_PARAM, "solve_batch: bad arguments");
  if (count == 0) return SC23AD_OK;
  for (int k = 0; k < count; ++k)
    for (int q = 0; q < k; ++q)
      if (solvers[k] == solvers[q])
        return fail(SC23AD_BAD_PARAM, "solve_batch: a solver appears twice");
  int nt = threads > 0 ? threads : 16;
  nt = std::max(1, std::min(nt, count));
  std::atomic<int> next{0};
  std::vector<int> rc(count, SC23AD_OK);
  std::vector<std::string> msg(count);
  auto worker = [&]() {
    for (int k = next.fetch_add(1); k < count; k = next.fetch_add(1)) {
      rc[k] = solve_one(solvers[k], b[k], nrm_b[k], x[k], on_device, out ? out + k : nullptr);
      if (rc[k] != SC23AD_OK) msg[k] = g_last_error;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  for (int k = 0; k < count; ++k)
    if (rc[k] != SC23AD_OK) return fail(rc[k], "solve_batch: system " + std::to_string(k) + ": " + msg[k]);
  return SC23AD_OK;
}

// ---- block solves: several right-hand sides of one hierarchy --------------

int sc23ad_amg_precondition_block(sc23ad_amg* m, int nrhs, const int32_t* in, int32_t* out,
                                  int on_device) {
  if (!m || !in || !out || nrhs < 1) return fail(SC23AD_BAD_PARAM, "precondition_block: bad arguments");
  if (m->precision != SC23AD_INT)
    return fail(SC23AD_BAD_PARAM, "precondition_block: integer hierarchies only");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    const int groups = (nrhs + 3) / 4;
    const size_t bytes = 16 * static_cast<size_t>(groups) * a.n0();
    auto w = a.make_block_work(groups);
    if (on_device) {
      a.vcycle_blk(in, out, *w);
    } else {
      DevBuf di(bytes), dout(bytes);
      di.upload(in, bytes);
      a.vcycle_blk(di.get(), dout.get(), *w);
      dout.download(out, bytes);
    }
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    a.check_errors(*w);
    return SC23AD_OK;
  });
}

int sc23ad_amg_time_vcycles_block(sc23ad_amg* m, int nrhs, int reps, const int32_t* in,
                                  int32_t* out, double* ms_per_cycle) {
  if (!m || !in || !out || !ms_per_cycle || nrhs < 1 || reps < 1)
    return fail(SC23AD_BAD_PARAM, "time_vcycles_block: bad arguments");
  if (m->precision != SC23AD_INT)
    return fail(SC23AD_BAD_PARAM, "time_vcycles_block: integer hierarchies only");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    auto w = a.make_block_work((nrhs + 3) / 4);
    a.vcycle_blk(in, out, *w);  // warm-up
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, lib_stream()), "record");
    for (int r = 0; r < reps; ++r) a.vcycle_blk(in, out, *w);
    hip_check(hipEventRecord(e1, lib_stream()), "record");
    hip_check(hipEventSynchronize(e1), "sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_cycle = ms / reps;
    a.check_errors(*w);
    return SC23AD_OK;
  });
}

int sc23ad_amg_block_native(const sc23ad_amg* m, int* native) {
  if (!m || !native) return fail(SC23AD_BAD_PARAM, "amg_block_native: null");
  *native = (m->precision == SC23AD_INT && m->d->block_native()) ? 1 : 0;
  return SC23AD_OK;
}

int sc23ad_block_solver_create(const sc23ad_csrmat* a, sc23ad_amg* m, int nrhs, int restart_m,
                               double eps, const int* frac_bits, int max_restarts,
                               sc23ad_block_solver** out) {
  if (!a || !m || !out || !frac_bits || nrhs < 1 || nrhs > SC23AD_BLOCK_MAX_RHS ||
      restart_m < 1 || !(eps > 0))
    return fail(SC23AD_BAD_PARAM, "block_solver_create: bad arguments");
  if (m->precision != SC23AD_INT)
    return fail(SC23AD_BAD_PARAM, "block_solver_create: integer hierarchies only");
  for (int q = 0; q < nrhs; ++q)
    if (frac_bits[q] < 1 || frac_bits[q] > 30)
      return fail(SC23AD_BAD_PARAM, "block_solver_create: frac_bits must be in [1, 30]");
  if (a->a.n != m->d->n0()) return fail(SC23AD_BAD_PARAM, "block_solver_create: size mismatch");
  return guarded([&]() -> int {
    auto s = std::make_unique<sc23ad_block_solver_s>();
    const int outer = max_restarts > 0 ? max_restarts : 1000 / restart_m;
    s->s = std::make_unique<DevBlockSolver>(a->a, m->d.get(), nrhs, restart_m, eps, frac_bits,
                                            outer);
    *out = s.release();
    return SC23AD_OK;
  });
}

void sc23ad_block_solver_destroy(sc23ad_block_solver* s) { delete s; }

namespace {

int block_solve_one(sc23ad_block_solver* s, const double* b, const double* nrm_b, double* x,
                    int on_device, sc23ad_converge* out) {
  if (!s || !b || !nrm_b || !x) return fail(SC23AD_BAD_PARAM, "block solve: null");
  return guarded([&]() -> int {
    DevBlockSolver& d = *s->s;
    StreamScope scope(d.stream);
    const size_t bytes = sizeof(double) * d.amg->n0() * static_cast<size_t>(d.count);
    std::vector<DevSolver::Result> r(d.count);
    if (on_device) {
      d.solve(b, nrm_b, x, r.data());
    } else {
      if (s->hb.bytes() < bytes) s->hb.alloc(bytes);
      if (s->hx.bytes() < bytes) s->hx.alloc(bytes);
      s->hb.upload(b, bytes);
      s->hx.upload(x, bytes);
      d.solve(s->hb.as<double>(), nrm_b, s->hx.as<double>(), r.data());
      s->hx.download(x, bytes);
    }
    if (out)
      for (int q = 0; q < d.count; ++q) {
        memset(out + q, 0, sizeof *out);
        out[q].is_converged = r[q].converged;
        out[q].iter = r[q].iter;
        out[q].restarts = r[q].restarts;
        out[q].status = r[q].status;
        out[q].overflow_flag = r[q].overflow;
        out[q].frac_bits = d.bits[q];
        out[q].implicit_relres = r[q].implicit_relres;
        out[q].solve_ms = r[q].solve_ms;
      }
    return SC23AD_OK;
  });
}

}  // namespace

int sc23ad_block_solver_solve(sc23ad_block_solver* s, const double* b, const double* nrm_b,
                              double* x, int on_device, sc23ad_converge* out) {
  return block_solve_one(s, b, nrm_b, x, on_device, out);
}

int sc23ad_block_solve_batch(int count, sc23ad_block_solver* const* solvers,
                             const double* const* b, const double* const* nrm_b,
                             double* const* x, int on_device, int threads,
                             sc23ad_converge* const* out) {
  if (count < 0 || (count > 0 && (!solvers || !b || !nrm_b || !x)))
    return fail(SC23AD_BAD_PARAM, "block_solve_batch: bad arguments");
  if (count == 0) return SC23AD_OK;
  for (int k = 0; k < count; ++k)
    for (int q = 0; q < k; ++q)
      if (solvers[k] == solvers[q])
        return fail(SC23AD_BAD_PARAM, "block_solve_batch: a solver appears twice");
  int nt = threads > 0 ? threads : 16;
  nt = std::max(1, std::min(nt, count));
  std::atomic<int> next{0};
  std::vector<int> rc(count, SC23AD_OK);
  std::vector<std::string> msg(count);
  auto worker = [&]() {
    for (int k = next.fetch_add(1); k < count; k = next.fetch_add(1)) {
      rc[k] = block_solve_one(solvers[k], b[k], nrm_b[k], x[k], on_device,
                              out ? out[k] : nullptr);
      if (rc[k] != SC23AD_OK) msg[k] = g_last_error;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  for (int k = 0; k < count; ++k)
    if (rc[k] != SC23AD_OK)
      return fail(rc[k], "block_solve_batch: system " + std::to_string(k) + ": " + msg[k]);
  return SC23AD_OK;
}

int sc23ad_block_solver_trace(sc23ad_block_solver* s, int rhs, uint64_t* buf, int cap,
                              int* len) {
  if (!s) return fail(SC23AD_BAD_PARAM, "block trace: null");
  return guarded([&]() -> int {
    if (buf == nullptr) {
      s->s->set_trace(cap);
      return SC23AD_OK;
    }
    if (rhs < 0 || rhs >= s->s->count) return fail(SC23AD_BAD_PARAM, "block trace: bad rhs");
    int l = 0;
    s->s->read_trace(rhs, buf, cap, &l);
    if (len) *len = l;
    return SC23AD_OK;
  });
}

int sc23ad_block_solver_history(sc23ad_block_solver* s, int rhs, double* buf, int cap,
                                int* len) {
  if (!s || !len || rhs < 0 || rhs >= s->s->count)
    return fail(SC23AD_BAD_PARAM, "block history: bad arguments");
  const auto& h = s->s->history[rhs];
  const int n = static_cast<int>(h.size());
  *len = n;
  if (buf)
    for (int i = 0; i < n && i < cap; ++i) buf[i] = h[i];
  return SC23AD_OK;
}

int sc23ad_solver_trace(sc23ad_solver* s, uint64_t* buf, int cap, int* len) {
  if (!s) return fail(SC23AD_BAD_PARAM, "trace: null");
  return guarded([&]() -> int {
    if (buf == nullptr) {
      s->s->set_trace(cap);
    } else {
      int l = 0;
      s->s->read_trace(buf, cap, &l);
      if (len) *len = l;
    }
    return SC23AD_OK;
  });
}

int sc23ad_solver_set_reduction(sc23ad_solver* s, int sequential) {
  if (!s) return fail(SC23AD_BAD_PARAM, "set_reduction: null");
  s->s->opts.seq_reduction = sequential ? 1 : 0;
  return SC23AD_OK;
}

int sc23ad_solver_set_stall(sc23ad_solver* s, int steps) {
  if (!s || steps < 0) return fail(SC23AD_BAD_PARAM, "set_stall: bad arguments");
  s->s->opts.stall_steps = steps;
  return SC23AD_OK;
}

int sc23ad_block_solver_set_stall(sc23ad_block_solver* s, int steps) {
  if (!s || steps < 0) return fail(SC23AD_BAD_PARAM, "block set_stall: bad arguments");
  s->s->stall_steps = steps;
  return SC23AD_OK;
}

int sc23ad_solver_history(sc23ad_solver* s, double* buf, int cap, int* len) {
  if (!s || !len) return fail(SC23AD_BAD_PARAM, "history: null");
  const auto& h = s->s->history;
  const int n = static_cast<int>(h.size());
  *len = n;
  if (buf)
    for (int i = 0; i < n && i < cap; ++i) buf[i] = h[i];
  return SC23AD_OK;
}

int sc23ad_solver_time_spmv(sc23ad_solver* s, int enable) {
  if (!s) return fail(SC23AD_BAD_PARAM, "time_spmv: null");
  return guarded([&]() -> int {
    for (EventTimer* t : {&s->s->spmv_timer, &s->s->trsv0_timer}) {
      t->collect();
      t->reset();
      t->on = enable != 0;
    }
    return SC23AD_OK;
  });
}

static int timer_stats(EventTimer& t, int* launches, double* mean_ms) {
  return guarded([&]() -> int {
    t.collect();
    if (launches) *launches = t.launches;
    if (mean_ms) *mean_ms = t.launches ? t.ms_total / t.launches : 0.0;
    return SC23AD_OK;
  });
}

int sc23ad_solver_spmv_stats(sc23ad_solver* s, int* launches, double* mean_ms) {
  if (!s) return fail(SC23AD_BAD_PARAM, "spmv_stats: null");
  return timer_stats(s->s->spmv_timer, launches, mean_ms);
}

int sc23ad_solver_trsv_stats(sc23ad_solver* s, int* launches, double* mean_ms) {
  if (!s) return fail(SC23AD_BAD_PARAM, "trsv_stats: null");
  return timer_stats(s->s->trsv0_timer, launches, mean_ms);
}

int sc23ad_solver_relres(sc23ad_solver* s, const double* b, double nrm_b,
                         const double* x, int on_device, double* relres) {
  if (!s || !b || !x || !relres) return fail(SC23AD_BAD_PARAM, "relres: null");
  return guarded([&]() -> int {
    DevSolver& d = *s->s;
    const size_t bytes = sizeof(double) * d.amg->n0();
    if (on_device) {
      *relres = d.relres(b, nrm_b, x);
    } else {
      DevBuf db(bytes), dx(bytes);
      db.upload(b, bytes);
      dx.upload(x, bytes);
      *relres = d.relres(db.as<double>(), nrm_b, dx.as<double>());
    }
    return SC23AD_OK;
  });
}

int sc23ad_default_frac_bits(sc23ad_amg* m, const double* b, double nrm_b, int* bits) {
  if (!m || !b || !bits) return fail(SC23AD_BAD_PARAM, "default_frac_bits: null");
  if (m->precision != SC23AD_FP64)
    return fail(SC23AD_BAD_PARAM, "default_frac_bits needs an FP64 amg");
  return guarded([&]() -> int {
    DevAmg& a = *m->d;
    const int n = a.n0();
    DevBuf db(sizeof(double) * n), dc(sizeof(double) * n);
    db.upload(b, sizeof(double) * n);
    a.vcycle(db.get(), dc.get(), a.own_work());
    a.check_errors(a.own_work());
    std::vector<double> c(n);
    dc.download(c.data(), sizeof(double) * n);
    double c_max = 0;
    for (int i = 0; i < n; ++i)
      if (std::fabs(c[i]) > c_max) c_max = std::fabs(c[i]);
    // Work/main.cpp:160
    *bits = (c_max / nrm_b < 1)
                ? 29
                : 31 - (static_cast<int>(std::ceil(std::log2(static_cast<int>(c_max / nrm_b) + 1))) + 1);
    return SC23AD_OK;
  });
}


// ---- learned per-level smoother parameters ----------------------------------

int sc23ad_mlp_create(uint64_t seed, sc23ad_mlp** out) {
  if (!out) return fail(SC23AD_BAD_PARAM, "mlp_create: null");
  return guarded([&]() -> int {
    if (dev::mlp_param_count() != kMlpParams)
      return fail(SC23AD_BAD_PARAM, "mlp_create: parameter layout mismatch");
    auto m = std::make_unique<sc23ad_mlp>();
    std::vector<float> p(kMlpParams), z(kMlpParams, 0.0f);
    mlp_init(seed, p.data());
    for (DevBuf* b : {&m->P, &m->G, &m->M, &m->V}) b->alloc(sizeof(float) * kMlpParams);
    m->P.upload(p.data(), sizeof(float) * kMlpParams);
    m->M.upload(z.data(), sizeof(float) * kMlpParams);
    m->V.upload(z.data(), sizeof(float) * kMlpParams);
    m->loss.alloc(sizeof(float));
    *out = m.release();
    return SC23AD_OK;
  });
}

void sc23ad_mlp_destroy(sc23ad_mlp* m) { delete m; }

int sc23ad_mlp_nparams(int* n) {
  if (!n) return fail(SC23AD_BAD_PARAM, "mlp_nparams: null");
  *n = kMlpParams;
  return SC23AD_OK;
}

int sc23ad_mlp_params(sc23ad_mlp* m, float* get, const float* set) {
  if (!m) return fail(SC23AD_BAD_PARAM, "mlp_params: null");
  return guarded([&]() -> int {
    if (set) {
      m->P.upload(set, sizeof(float) * kMlpParams);
      m->step = 0;
      std::vector<float> z(kMlpParams, 0.0f);
      m->M.upload(z.data(), sizeof(float) * kMlpParams);
      m->V.upload(z.data(), sizeof(float) * kMlpParams);
    }
    if (get) m->P.download(get, sizeof(float) * kMlpParams);
    return SC23AD_OK;
  });
}

int sc23ad_mlp_forward(sc23ad_mlp* m, const float* x, int batch, float* y) {
  if (!m || batch < 0 || (batch > 0 && (!x || !y))) return fail(SC23AD_BAD_PARAM, "mlp_forward: bad arguments");
  return guarded([&]() -> int {
    if (batch == 0) return SC23AD_OK;
    DevBuf dx(sizeof(float) * batch * kMlpIn), dy(sizeof(float) * batch * kMlpOut);
    dx.upload(x, sizeof(float) * batch * kMlpIn);
    hip_check(dev::launch_mlp_forward(batch, dx.as<float>(), m->P.as<float>(), dy.as<float>(),
                                      lib_stream()),
              "mlp forward");
    dy.download(y, sizeof(float) * batch * kMlpOut);
    return SC23AD_OK;
  });
}

int sc23ad_mlp_grad(sc23ad_mlp* m, const float* x, const float* t, int batch, float* loss,
                    float* grad) {
  if (!m || batch < 1 || !x || !t) return fail(SC23AD_BAD_PARAM, "mlp_grad: bad arguments");
  return guarded([&]() -> int {
    DevBuf dx(sizeof(float) * batch * kMlpIn), dt(sizeof(float) * batch * kMlpOut);
    dx.upload(x, sizeof(float) * batch * kMlpIn);
    dt.upload(t, sizeof(float) * batch * kMlpOut);
    hip_check(dev::launch_mlp_grad(batch, dx.as<float>(), dt.as<float>(), m->P.as<float>(),
                                   m->G.as<float>(), m->loss.as<float>(), lib_stream()),
              "mlp grad");
    if (loss) m->loss.download(loss, sizeof(float));
    if (grad) m->G.download(grad, sizeof(float) * kMlpParams);
    return SC23AD_OK;
  });
}

int sc23ad_mlp_train(sc23ad_mlp* m, const float* x, const float* t, int batch, int steps,
                     float lr, float* loss) {
  if (!m || batch < 1 || !x || !t || steps < 0 || !(lr > 0))
    return fail(SC23AD_BAD_PARAM, "mlp_train: bad arguments");
  return guarded([&]() -> int {
    DevBuf dx(sizeof(float) * batch * kMlpIn), dt(sizeof(float) * batch * kMlpOut);
    dx.upload(x, sizeof(float) * batch * kMlpIn);
    dt.upload(t, sizeof(float) * batch * kMlpOut);
    for (int k = 0; k < steps; ++k) {
      hip_check(dev::launch_mlp_grad(batch, dx.as<float>(), dt.as<float>(), m->P.as<float>(),
                                     m->G.as<float>(), m->loss.as<float>(), lib_stream()),
                "mlp grad");
      hip_check(dev::launch_adam(kMlpParams, m->P.as<float>(), m->G.as<float>(),
                                 m->M.as<float>(), m->V.as<float>(), lr, ++m->step,
                                 lib_stream()),
                "adam");
    }
    if (loss) {  // loss at the final parameters
      hip_check(dev::launch_mlp_grad(batch, dx.as<float>(), dt.as<float>(), m->P.as<float>(),
                                     m->G.as<float>(), m->loss.as<float>(), lib_stream()),
                "mlp grad");
      m->loss.download(loss, sizeof(float));
    }
    return SC23AD_OK;
  });
}

int sc23ad_learned_levels(sc23ad_mlp* m, const sc23ad_hier* h, double* level_omega,
                          int* level_vbits, int cap, int* nlevels) {
  if (!m || !h || !nlevels) return fail(SC23AD_BAD_PARAM, "learned_levels: null");
  int nl = 0;
  int rc = sc23ad_hier_features(h, nullptr, 0, &nl);
  if (rc != SC23AD_OK) return rc;
  *nlevels = nl;
  if (nl == 0 || (!level_omega && !level_vbits)) return SC23AD_OK;
  std::vector<float> f(static_cast<size_t>(nl) * kLevelFeatures), y(static_cast<size_t>(nl) * kMlpOut);
  rc = sc23ad_hier_features(h, f.data(), nl, &nl);
  if (rc != SC23AD_OK) return rc;
  rc = sc23ad_mlp_forward(m, f.data(), nl, y.data());
  if (rc != SC23AD_OK) return rc;
  if (level_vbits && cap > 0) level_vbits[0] = 0;  // the fine level keeps the solver's F
  for (int l = 0; l < nl && l < cap; ++l) {
    if (level_omega) level_omega[l] = mlp_omega(y[2 * l]);
    if (level_vbits && l + 1 <= cap) level_vbits[l + 1] = mlp_vbits(y[2 * l + 1]);
  }
  return SC23AD_OK;
}

}  // extern "C"
