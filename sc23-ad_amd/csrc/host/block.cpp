// Several right-hand sides of one hierarchy at once: the block V-cycle
// (intAMGflex::Precondition, Library/preconditioner/INTAMG.hpp:269-293,
// applied to 4 * groups vectors) and k Restarted(FGMRESflex(m)) solves in
// lockstep (Library/solver/irsolver.hpp:27-43, gmres.hpp:225-271).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "device.hpp"

namespace sc23ad {

using dev::Elem;
using dev::SpmvMode;

std::unique_ptr<DevAmg::BlockWork> DevAmg::make_block_work(int groups) const {
  if (elem != Elem::kI32) throw DevError{1, "block V-cycle: integer hierarchies only"};
  auto w = std::make_unique<BlockWork>();
  w->groups = groups;
  const size_t row = 16 * static_cast<size_t>(groups);  // bytes of one block row
  w->ee.resize(depth + 1);
  w->xx.resize(depth + 1);
  w->rr.resize(depth + 1);
  w->tmp.resize(depth + 1);
  int nmax = std::max(n0(), 1);
  for (int l = 1; l <= depth; ++l) {
    const size_t n = static_cast<size_t>(std::max(level_n(l), 1));
    nmax = std::max(nmax, static_cast<int>(n));
    for (auto* v : {&w->ee, &w->xx, &w->rr, &w->tmp}) (*v)[l].alloc(row * n);
  }
  const size_t n0s = static_cast<size_t>(std::max(n0(), 1));
  w->r0.alloc(row * n0s);
  w->vin0.alloc(row * n0s);
  w->vout0.alloc(row * n0s);
  w->tmp[0].alloc(row * n0s);
  const size_t nc = static_cast<size_t>(std::max(n_coarse, 1));
  w->tmpc.alloc(row * nc);
  for (DevBuf* f : {&w->lflags, &w->uflags}) {
    f->alloc(sizeof(unsigned long long) * 4 * groups * nc);
    hip_check(dev::launch_fill_zero(f->bytes(), f->get(), lib_stream()), "zero flags");
  }
  w->tickets.alloc(4 * sizeof(unsigned));
  hip_check(dev::launch_fill_zero(w->tickets.bytes(), w->tickets.get(), lib_stream()), "zero");
  w->err.alloc(sizeof(int));
  hip_check(dev::launch_fill_zero(w->err.bytes(), w->err.get(), lib_stream()), "zero");
  if (!block_native()) {
    w->col_in.alloc(sizeof(int32_t) * nmax);
    w->col_out.alloc(sizeof(int32_t) * nmax);
  }
  hip_check(hipStreamSynchronize(lib_stream()), "block work");
  return w;
}

bool DevAmg::block_native() const {
  for (int l = 0; l < depth; ++l) {
    const TriPlan& p = lv[l].ld_plan;
    if (!(p.segs.size() == 1 && dev::trsv_rp_blk_supported(p.layout))) return false;
  }
  return lc_plan.dense && uc_plan.dense;
}

void DevAmg::spmv_blk(const DevMat& m, SpmvMode mode, const void* x, const void* z, void* y,
                      int groups) const {
  hip_check(dev::launch_spmv_blk(m.d, groups, mode, x, z, y, lib_stream()), "spmv block");
}

void DevAmg::trsv_blk(const TriPlan& plan, const DevMat& m, const void* in, void* out,
                      BlockWork& w, unsigned long long* flags, int* epoch,
                      unsigned* tickets) const {
  const int groups = w.groups;
  if (plan.dense && flags) {
    *epoch = *epoch == INT32_MAX ? 1 : *epoch + 1;
    hip_check(dev::launch_trsv_tiles_blk(plan.kind, m.d.n, plan.nb, groups,
                                         plan.slice.as<int32_t>(), plan.boff.as<long long>(),
                                         plan.ddinv.as<int32_t>(), m.d.bit, m.d.dbit, in, out,
                                         flags, *epoch, w.err.as<int>(), tickets, lib_stream()),
              "trsv tiles block");
    return;
  }
  if (plan.segs.size() == 1 && dev::trsv_rp_blk_supported(plan.layout)) {
    hip_check(dev::launch_trsv_rp_blk(plan.layout, plan.segs[0], groups, m.d.bit, m.d.dbit, in,
                                      out, lib_stream()),
              "trsv rp block");
    return;
  }
  // one right-hand side at a time through the single-vector kernels
  const int n = m.d.n;
  for (int r = 0; r < 4 * groups; ++r) {
    hip_check(dev::launch_blk_column(n, r, const_cast<void*>(in), w.col_in.get(), true,
                                     lib_stream()),
              "column");
    trsv(plan, m, w.col_in.get(), w.col_out.get(), nullptr, nullptr, nullptr, nullptr);
    hip_check(dev::launch_blk_column(n, r, out, w.col_out.get(), false, lib_stream()), "column");
  }
}

// SOR / BlockSOR Smooth (stationary.hpp:415-430, 997-1012) on a block.
void DevAmg::smooth_blk(int l, const void* in, void* out, bool inited, BlockWork& w) const {
  const DevLevel& L = lv[l];
  int k0 = 0;
  if (!inited) {
    trsv_blk(L.ld_plan, L.ld, in, out, w, nullptr, nullptr, nullptr);
    k0 = 1;
  }
  void* t = w.tmp[l].get();
  for (int k = k0; k < iter; ++k) {
    spmv_blk(L.u, SpmvMode::kSub, out, in, t, w.groups);
    trsv_blk(L.ld_plan, L.ld, t, out, w, nullptr, nullptr, nullptr);
  }
}

void DevAmg::coarse_blk(const void* in, void* out, BlockWork& w) const {
  unsigned* tk = w.tickets.as<unsigned>();
  trsv_blk(lc_plan, lc, in, w.tmpc.get(), w, w.lflags.as<unsigned long long>(), &w.lepoch, tk);
  trsv_blk(uc_plan, uc, w.tmpc.get(), out, w, w.uflags.as<unsigned long long>(), &w.uepoch,
           tk + 2);
}

void DevAmg::vcycle_blk(const void* in, void* out, BlockWork& w) const {
  const int n = n0();
  hip_check(dev::launch_permute_blk(n, w.groups, perm_of(0), in, w.vin0.get(), true,
                                    lib_stream()),
            "permute in");
  vcycle_lo_blk(w.vin0.get(), w.vout0.get(), w);
  hip_check(dev::launch_permute_blk(n, w.groups, perm_of(0), w.vout0.get(), out, false,
                                    lib_stream()),
            "permute out");
}

void DevAmg::vcycle_lo_blk(const void* in, void* out, BlockWork& w) const {
  const int g = w.groups;
  smooth_blk(0, in, out, false, w);
  spmv_blk(lv[0].a, SpmvMode::kSub, out, in, w.r0.get(), g);
  if (depth == 0) return;
  spmv_blk(lv[0].r, SpmvMode::kPlain, w.r0.get(), nullptr, w.ee[1].get(), g);
  for (int l = 1; l < depth; ++l) {
    smooth_blk(l, w.ee[l].get(), w.xx[l].get(), false, w);
    spmv_blk(lv[l].a, SpmvMode::kSub, w.xx[l].get(), w.ee[l].get(), w.rr[l].get(), g);
    spmv_blk(lv[l].r, SpmvMode::kPlain, w.rr[l].get(), nullptr, w.ee[l + 1].get(), g);
  }
  coarse_blk(w.ee[depth].get(), w.xx[depth].get(), w);
  for (int l = depth - 1; l >= 1; --l) {
    spmv_blk(lv[l].p, SpmvMode::kAdd, w.xx[l + 1].get(), w.xx[l].get(), w.xx[l].get(), g);
    smooth_blk(l, w.ee[l].get(), w.xx[l].get(), true, w);
  }
  spmv_blk(lv[0].p, SpmvMode::kAdd, w.xx[1].get(), out, out, g);
  smooth_blk(0, in, out, true, w);
}

void DevAmg::check_errors(BlockWork& w) const {
  int e = 0;
  hip_check(hipMemcpyAsync(&e, w.err.get(), sizeof(int), hipMemcpyDeviceToHost, lib_stream()),
            "error flag");
  hip_check(hipStreamSynchronize(lib_stream()), "sync");
  if (e != 0) {
    hip_check(dev::launch_fill_zero(sizeof(int), w.err.get(), lib_stream()), "error reset");
    hip_check(hipStreamSynchronize(lib_stream()), "sync");
    throw DevError{5, "coarse solve: a block waited too long for its dependencies"};
  }
}

// ---------------------------------------------------------------------------

DevBlockSolver::DevBlockSolver(const HostCsr& a, DevAmg* amg_, int count_, int m_, double eps_,
                               const int* bits_, int max_outer_)
    : count(count_), groups((count_ + 3) / 4), m(m_), eps(eps_), max_outer(max_outer_),
      bits(bits_, bits_ + count_), amg(amg_) {
  lib_stream();
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  StreamScope scope(stream);
  work = amg->make_block_work(groups);
  upload_fp64_operator(a64, a);
  const int n = amg->n0();
  const int K = 4 * groups;
  const size_t vs = 16 * static_cast<size_t>(groups) * std::max(n, 1);
  V.alloc(vs * (m + 1));
  Z.alloc(vs * (m + 1));
  r.alloc(sizeof(double) * static_cast<size_t>(std::max(n, 1)) * count);
  // per right-hand side: the state block of DevSolver (H, c, s, e, y, partial
  // sums, scalars, flags, trace length)
  const size_t H = 8 * static_cast<size_t>(m + 1) * m;
  auto up = [](size_t b) { return (b + 7) & ~static_cast<size_t>(7); };
  const size_t per = up(H) + 2 * up(8 * m) + up(8 * (m + 1)) + up(8 * m) + up(4 * m) +
                     up(8 * 3 * dev::kRedBlocks) + up(8 * 2 * dev::kRedBlocks) +
                     up(4 * 2 * dev::kRedBlocks) + up(64) + up(8 * 8) + up(8 * 4) + up(4 * 8) +
                     up(16);
  states.alloc(per * K);
  hip_check(dev::launch_fill_zero(states.bytes(), states.get(), lib_stream()), "zero states");
  ks.assign(K, dev::KState{});
  for (int q = 0; q < K; ++q) {
    char* p = states.as<char>() + per * q;
    auto take = [&](size_t bytes) {
      char* c = p;
      p += up(bytes);
      return c;
    };
    dev::KState& k = ks[q];
    k.m = m;
    k.bit = q < count ? bits[q] : 29;
    k.eps = eps;
    k.H = take(H);
    k.c = take(8 * m);
    k.s = take(8 * m);
    k.e = take(8 * (m + 1));
    k.y = reinterpret_cast<double*>(take(8 * m));
    k.yf = reinterpret_cast<float*>(take(4 * m));
    k.ipart = reinterpret_cast<int64_t*>(take(8 * 3 * dev::kRedBlocks));
    k.dpart = reinterpret_cast<double*>(take(8 * 2 * dev::kRedBlocks));
    k.fpart = reinterpret_cast<float*>(take(4 * 2 * dev::kRedBlocks));
    k.counter = reinterpret_cast<unsigned*>(take(64));
    k.scal = reinterpret_cast<double*>(take(8 * 8));
    k.iscal = reinterpret_cast<int64_t*>(take(8 * 4));
    k.flags = reinterpret_cast<int*>(take(4 * 8));
    k.trace_len = reinterpret_cast<int*>(take(16));
    k.trace = nullptr;
    k.trace_cap = 0;
  }
  kst.alloc(sizeof(dev::KState) * K);
  live.alloc(sizeof(int) * K);
  upd.alloc(sizeof(int) * K);
  flagbuf.alloc(sizeof(int) * 6 * K);
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_flags), sizeof(int) * 6 * K,
                          hipHostMallocDefault),
            "hipHostMalloc");
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&host_mask), sizeof(int) * 2 * K,
                          hipHostMallocDefault),
            "hipHostMalloc");
  history.assign(count, {});
  hip_check(hipStreamSynchronize(lib_stream()), "block solver init");
}

DevBlockSolver::~DevBlockSolver() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (host_flags) (void)hipHostFree(host_flags);
  if (host_mask) (void)hipHostFree(host_mask);
  work.reset();
  if (stream) (void)hipStreamDestroy(stream);
}

void DevBlockSolver::set_trace(int cap) {
  StreamScope scope(stream);
  const int K = 4 * groups;
  tracing = cap > 0;
  trace_cap = std::max(cap, 0);
  if (tracing) trace_buf.alloc(sizeof(uint64_t) * static_cast<size_t>(cap) * K);
  for (int q = 0; q < K; ++q) {
    ks[q].trace = tracing ? trace_buf.as<uint64_t>() + static_cast<size_t>(cap) * q : nullptr;
    ks[q].trace_cap = tracing ? cap : 0;
    hip_check(dev::launch_fill_zero(16, ks[q].trace_len, lib_stream()), "trace reset");
  }
  hip_check(hipStreamSynchronize(lib_stream()), "trace");
}

void DevBlockSolver::read_trace(int rhs, uint64_t* out, int cap, int* len) {
  StreamScope scope(stream);
  int l = 0;
  hip_check(hipMemcpyAsync(&l, ks[rhs].trace_len, sizeof(int), hipMemcpyDeviceToHost,
                           lib_stream()),
            "trace len");
  hip_check(hipStreamSynchronize(lib_stream()), "sync");
  const int cnt = tracing ? std::min(std::min(l, cap), trace_cap) : 0;
  if (cnt > 0)
    trace_buf.download(out, sizeof(uint64_t) * cnt,
                       sizeof(uint64_t) * static_cast<size_t>(trace_cap) * rhs);
  *len = cnt;
}

void DevBlockSolver::read_flags() {
  const int K = 4 * groups;
  hip_check(dev::launch_flags_blk(kst.as<dev::KState>(), K, flagbuf.as<int>(), lib_stream()),
            "flags");
  hip_check(hipMemcpyAsync(host_flags, flagbuf.get(), sizeof(int) * 6 * K,
                           hipMemcpyDeviceToHost, lib_stream()),
            "flags copy");
  hip_check(hipStreamSynchronize(lib_stream()), "flags sync");
}

void DevBlockSolver::put_masks(const std::vector<int>& live_v, const std::vector<int>& upd_v) {
  const int K = 4 * groups;
  for (int q = 0; q < K; ++q) {
    host_mask[q] = live_v[q];
    host_mask[K + q] = upd_v[q];
  }
  hip_check(hipMemcpyAsync(live.get(), host_mask, sizeof(int) * K, hipMemcpyHostToDevice,
                           lib_stream()),
            "live mask");
  hip_check(hipMemcpyAsync(upd.get(), host_mask + K, sizeof(int) * K, hipMemcpyHostToDevice,
                           lib_stream()),
            "update mask");
  // the pinned words are read by the copies above before they are rewritten
  hip_check(hipStreamSynchronize(lib_stream()), "mask sync");
}

void DevBlockSolver::solve(const double* b_dev, const double* nrm_b, double* x_dev,
                           DevSolver::Result* res) {
  StreamScope scope(stream);
  hipStream_t s = lib_stream();
  const int n = amg->n0();
  const int K = 4 * groups;
  const size_t vs = 16 * static_cast<size_t>(groups) * n;  // bytes of one block vector
  for (int q = 0; q < K; ++q) ks[q].nrm_b = q < count ? nrm_b[q] : 1.0;
  hip_check(hipMemcpyAsync(kst.get(), ks.data(), sizeof(dev::KState) * K,
                           hipMemcpyHostToDevice, s),
            "states");
  for (int q = 0; q < K; ++q) {
    // Restarted: prev = 0 before the first pass
    hip_check(dev::launch_fill_zero(8 * sizeof(double), ks[q].scal, s), "reset scal");
    hip_check(dev::launch_fill_zero(8 * sizeof(int), ks[q].flags, s), "reset flags");
  }
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  hip_check(hipEventRecord(e0, s), "record");
  for (int q = 0; q < count; ++q) {
    res[q] = DevSolver::Result{};
    res[q].status = 6;
    history[q].clear();
  }
  std::vector<int> active(K, 0), iter(count, 0), none(K, 0);
  std::vector<StallWatch> stall(count);
  for (int q = 0; q < count; ++q) active[q] = 1;
  char* Vb = V.as<char>();
  char* Zb = Z.as<char>();
  // systems of at most this many rows run each Gram-Schmidt step in one
  // launch; SC23AD_MGS_FUSED_MAX sets the bound (0: never)
  int mgs_fused_rows = dev::kMgsFusedRows;
  if (const char* e = std::getenv("SC23AD_MGS_FUSED_MAX")) mgs_fused_rows = std::atoi(e);
  for (int outer = 0; outer < max_outer; ++outer) {
    put_masks(active, none);
    hip_check(dev::launch_outer_residual_blk(a64.d, K, b_dev, x_dev, r.as<double>(),
                                             kst.as<dev::KState>(), live.as<int>(), s),
              "outer residual");
    read_flags();
    int nact = 0;
    for (int q = 0; q < count; ++q) {
      if (!active[q]) continue;
      if (host_flags[6 * q + 2]) {  // stagnation (irsolver.hpp:35)
        res[q].status = 4;
        active[q] = 0;
        continue;
      }
      ++res[q].restarts;
      ++nact;
    }
    if (nact == 0) break;
    std::vector<int> cyc = active;  // solves of this cycle
    put_masks(cyc, none);
    hip_check(dev::launch_v0_blk(n, groups, r.as<double>(), Vb, kst.as<dev::KState>(),
                                 live.as<int>(), s),
              "v0");
    std::vector<int> running = cyc, conv(count, 0), ovf(count, 0), stalled(count, 0);
    for (int j = 0; j < m; ++j) {
      void* vj = Vb + vs * j;
      void* zj = Zb + vs * j;
      void* w = Vb + vs * (j + 1);
      amg->vcycle_blk(vj, zj, *work);
      if (tracing) {
        hip_check(dev::launch_checksum_blk(n, groups, K, vj, kst.as<dev::KState>(),
                                           live.as<int>(), s),
                  "checksum");
        hip_check(dev::launch_checksum_blk(n, groups, K, zj, kst.as<dev::KState>(),
                                           live.as<int>(), s),
                  "checksum");
      }
      amg->spmv_blk(amg->a0, SpmvMode::kPlain, zj, nullptr, w, groups);
      if (n <= mgs_fused_rows) {
        hip_check(dev::launch_mgs_step_blk(n, groups, Vb, w, kst.as<dev::KState>(),
                                           live.as<int>(), j, s),
                  "mgs step");
      } else {
        hip_check(dev::launch_dot_first_blk(n, groups, Vb, w, kst.as<dev::KState>(), s), "dot");
        for (int kk = 0; kk <= j; ++kk)
          hip_check(dev::launch_axpy_dot_blk(n, groups, Vb, w, kst.as<dev::KState>(),
                                             live.as<int>(), kk, j, s),
                    "axpy_dot");
        hip_check(
            dev::launch_scal_blk(n, groups, w, kst.as<dev::KState>(), live.as<int>(), j, s),
            "scal");
      }
      read_flags();
      int still = 0;
      bool dropped = false;
      for (int q = 0; q < count; ++q) {
        if (!running[q]) continue;
        double rel = 0;
        std::memcpy(&rel, host_flags + 6 * q + 4, sizeof rel);
        history[q].push_back(rel);
        if (host_flags[6 * q + 1]) {
          ovf[q] = 1;
          running[q] = 0;
        } else if (host_flags[6 * q]) {
          conv[q] = j + 1;
          running[q] = 0;
        } else if (stall[q].step(rel, stall_steps)) {
          stalled[q] = 1;  // abandoned: leaves the cycle, x not updated
          running[q] = 0;
          cyc[q] = 0;
          dropped = true;
        } else {
          ++still;
        }
      }
      if (still == 0) break;
      if (dropped) put_masks(cyc, none);
    }
    std::vector<int> up(K, 0);
    for (int q = 0; q < count; ++q) up[q] = cyc[q] && !ovf[q];
    put_masks(cyc, up);
    hip_check(dev::launch_update_x_blk(n, groups, K, Zb, x_dev, kst.as<dev::KState>(),
                                       upd.as<int>(), s),
              "update x");
    int left = 0;
    for (int q = 0; q < count; ++q) {
      if (stalled[q]) {
        res[q].status = 4;
        active[q] = 0;
        continue;
      }
      if (!cyc[q]) continue;
      if (ovf[q]) {
        res[q].status = 3;
        res[q].overflow = 1;
        active[q] = 0;
      } else if (conv[q]) {
        iter[q] += conv[q];
        res[q].status = 0;
        res[q].converged = 1;
        res[q].iter = iter[q];
        active[q] = 0;
      } else {
        iter[q] += m;
        ++left;
      }
    }
    if (left == 0) break;
  }
  hip_check(hipEventRecord(e1, s), "record");
  hip_check(hipEventSynchronize(e1), "sync");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  read_flags();
  for (int q = 0; q < count; ++q) {
    res[q].solve_ms = ms;
    std::memcpy(&res[q].implicit_relres, host_flags + 6 * q + 4, sizeof(double));
  }
  amg->check_errors(*work);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

}  // namespace sc23ad
