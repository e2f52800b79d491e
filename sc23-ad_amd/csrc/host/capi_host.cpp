// Host half of the C ABI: CSR containers, generators, MatrixMarket input and
// the hierarchy build.  Device entry points live in capi_device.cpp.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <variant>

#include "../../../include/sc23ad.h"
#include "capi_internal.hpp"
#include "device.hpp"
#include "learn.hpp"
#include "gen.hpp"
#include "hier_io.hpp"
#include "hierarchy.hpp"
#include "mm.hpp"

using namespace sc23ad;

namespace sc23ad {
thread_local std::string g_last_error;
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace sc23ad

extern "C" {

const char* sc23ad_last_error(void) { return g_last_error.c_str(); }
const char* sc23ad_version(void) { return "sc23ad 0.1 (gfx950)"; }

void sc23ad_amg_param_default(sc23ad_amg_param* p, int smoother) {
  if (!p) return;
  p->max_levels = 20;
  p->theta = 0.8;
  p->cycle = 1;
  p->smoother = smoother;
  p->omega = 1.0;
  p->iter = 1;
  p->b_num = 40;
  p->coarse_n = 2000;
}

int sc23ad_csrmat_create(int n, int m, const int* rptr, const int* cind,
                         const double* val, sc23ad_csrmat** out) {
  if (!out || n < 0 || m < 0 || !rptr || (n > 0 && (!cind || !val)))
    return fail(SC23AD_BAD_PARAM, "csrmat_create: bad arguments");
  const int nnz = rptr[n];
  if (rptr[0] != 0 || nnz < 0)
    return fail(SC23AD_BAD_PARAM, "csrmat_create: rptr must start at 0");
  for (int i = 0; i < n; ++i) {
    if (rptr[i + 1] < rptr[i])
      return fail(SC23AD_BAD_PARAM, "csrmat_create: rptr not monotone");
    for (int p = rptr[i]; p < rptr[i + 1]; ++p) {
      if (cind[p] < 0 || cind[p] >= m)
        return fail(SC23AD_BAD_PARAM, "csrmat_create: column out of range");
      if (p > rptr[i] && cind[p] <= cind[p - 1])
        return fail(SC23AD_BAD_PARAM,
                    "csrmat_create: columns must be strictly ascending");
    }
  }
  auto* a = new (std::nothrow) sc23ad_csrmat;
  if (!a) return fail(SC23AD_BAD_PARAM, "out of host memory");
  a->a.n = n;
  a->a.m = m;
  a->a.rptr.assign(rptr, rptr + n + 1);
  a->a.cind.assign(cind, cind + nnz);
  a->a.val.assign(val, val + nnz);
  *out = a;
  return SC23AD_OK;
}

int sc23ad_csrmat_read_mm(const char* path, sc23ad_csrmat** out) {
  if (!path || !out) return fail(SC23AD_BAD_PARAM, "read_mm: bad arguments");
  HostCsr a;
  std::string err;
  if (!read_matrix_market(path, a, err)) return fail(SC23AD_IO_ERROR, err);
  auto* m = new sc23ad_csrmat;
  m->a = std::move(a);
  *out = m;
  return SC23AD_OK;
}

int sc23ad_csrmat_gen(int kind, const double* p, int np, sc23ad_csrmat** out) {
  if (!out || (np > 0 && !p)) return fail(SC23AD_BAD_PARAM, "gen: bad arguments");
  auto need = [&](int k) { return np >= k; };
  HostCsr a;
  switch (kind) {
    case SC23AD_GEN_POISSON2D:
      if (!need(1) || p[0] < 1) return fail(SC23AD_BAD_PARAM, "gen: poisson2d [n]");
      a = gen_poisson2d(static_cast<int>(p[0]));
      break;
    case SC23AD_GEN_LAPLACE3D:
      if (!need(1) || p[0] < 1) return fail(SC23AD_BAD_PARAM, "gen: laplace3d [n]");
      a = gen_laplace3d(static_cast<int>(p[0]));
      break;
    case SC23AD_GEN_DIFFUSION27:
      if (!need(6) || p[0] < 1 || p[1] < 1)
        return fail(SC23AD_BAD_PARAM, "gen: diffusion27 [n,block,ax,ay,az,seed]");
      a = gen_diffusion27(static_cast<int>(p[0]), static_cast<int>(p[1]), p[2],
                          p[3], p[4], static_cast<uint64_t>(p[5]));
      break;
    case SC23AD_GEN_CONVDIFF2D:
      if (!need(3) || p[0] < 1)
        return fail(SC23AD_BAD_PARAM, "gen: convdiff2d [n,peclet,angle]");
      a = gen_convdiff2d(static_cast<int>(p[0]), p[1], p[2]);
      break;
    case SC23AD_GEN_INCLUSIONS3D:
      if (!need(6) || p[0] < 1)
        return fail(SC23AD_BAD_PARAM,
                    "gen: inclusions3d [n,nsph,rmin,rmax,kin,seed]");
      a = gen_inclusions3d(static_cast<int>(p[0]), static_cast<int>(p[1]), p[2],
                           p[3], p[4], static_cast<uint64_t>(p[5]));
      break;
    default:
      return fail(SC23AD_BAD_PARAM, "gen: unknown kind");
  }
  auto* m = new sc23ad_csrmat;
  m->a = std::move(a);
  *out = m;
  return SC23AD_OK;
}

int sc23ad_csrmat_info(const sc23ad_csrmat* a, int* n, int* m, int* nnz) {
  if (!a) return fail(SC23AD_BAD_PARAM, "info: null matrix");
  if (n) *n = a->a.n;
  if (m) *m = a->a.m;
  if (nnz) *nnz = a->a.nnz();
  return SC23AD_OK;
}

int sc23ad_csrmat_export(const sc23ad_csrmat* a, int* rptr, int* cind,
                         double* val) {
  if (!a) return fail(SC23AD_BAD_PARAM, "export: null matrix");
  if (rptr) memcpy(rptr, a->a.rptr.data(), sizeof(int) * a->a.rptr.size());
  if (cind) memcpy(cind, a->a.cind.data(), sizeof(int) * a->a.cind.size());
  if (val) memcpy(val, a->a.val.data(), sizeof(double) * a->a.val.size());
  return SC23AD_OK;
}

int sc23ad_csrmat_lmax1_scaling(sc23ad_csrmat* a, double* b) {
  if (!a) return fail(SC23AD_BAD_PARAM, "lmax1: null matrix");
  lmax1_scaling(a->a, b);
  return SC23AD_OK;
}

int sc23ad_csrmat_matvec(const sc23ad_csrmat* a, const double* x, double* y) {
  if (!a || !x || !y) return fail(SC23AD_BAD_PARAM, "matvec: bad arguments");
  csr_matvec(a->a, x, y);
  return SC23AD_OK;
}

void sc23ad_csrmat_destroy(sc23ad_csrmat* a) { delete a; }

int sc23ad_xstar(int n, uint64_t seed, double* xstar) {
  if (n < 0 || (n > 0 && !xstar)) return fail(SC23AD_BAD_PARAM, "xstar: bad arguments");
  gen_xstar(n, seed, xstar);
  return SC23AD_OK;
}

int sc23ad_convdiff_draws(int count, uint64_t seed, int* n, double* pe,
                          double* angle) {
  if (count < 0 || !n || !pe || !angle)
    return fail(SC23AD_BAD_PARAM, "convdiff_draws: bad arguments");
  auto d = draw_convdiff_family(count, seed);
  for (int s = 0; s < count; ++s) {
    n[s] = d[s].n;
    pe[s] = d[s].pe;
    angle[s] = d[s].angle;
  }
  return SC23AD_OK;
}

static AmgParams to_params(const sc23ad_amg_param* p) {
  AmgParams prm;
  prm.max_levels = p->max_levels;
  prm.theta = p->theta;
  prm.cycle = p->cycle;
  prm.smoother = static_cast<SmootherKind>(p->smoother);
  prm.omega = p->omega;
  prm.iter = p->iter;
  prm.b_num = p->b_num;
  prm.coarse_n = p->coarse_n;
  return prm;
}

// A hierarchy of `precision` from a kept setup state with parameters prm
// (the state's own or per-level overrides of them).
static std::unique_ptr<sc23ad_hier> derive_hierarchy(
    const std::shared_ptr<const Sc23adSetupState>& st, const AmgParams& prm, int precision) {
  auto h = std::make_unique<sc23ad_hier>();
  h->precision = precision;
  h->warnings = st->chain.warnings;
  h->fine_nnz = st->fine.nnz();
  h->setup = st;
  if (precision == SC23AD_INT)
    h->ih = build_int_hierarchy(st->fine, st->chain, prm);
  else if (precision == SC23AD_FP64)
    h->dh = build_fp_hierarchy<double>(st->fine, st->chain, prm);
  else
    h->fh = build_fp_hierarchy<float>(st->fine, st->chain, prm);
  return h;
}

int sc23ad_hier_build_multi(const sc23ad_csrmat* a, const sc23ad_amg_param* p,
                            const int* precisions, int count, sc23ad_hier** outs) {
  return sc23ad::build_multi_impl(a, p, precisions, count, outs, false);
}

}  // extern "C"

namespace sc23ad {

int build_multi_impl(const sc23ad_csrmat* a, const sc23ad_amg_param* p, const int* precisions,
                     int count, sc23ad_hier** outs, bool device_setup) {
  if (!a || !p || !precisions || !outs || count < 1)
    return fail(SC23AD_BAD_PARAM, "hier_build: bad arguments");
  for (int k = 0; k < count; ++k) {
    outs[k] = nullptr;
    if (precisions[k] < 0 || precisions[k] > 2)
      return fail(SC23AD_BAD_PARAM, "hier_build: precision must be 0, 1 or 2");
  }
  if (p->smoother < 0 || p->smoother > 3)
    return fail(SC23AD_BAD_PARAM, "hier_build: smoother must be 0, 1, 2 or 3");
  for (int k = 0; k < count; ++k)
    if (p->smoother == 3 && precisions[k] == SC23AD_INT)
      return fail(SC23AD_BAD_PARAM, "hier_build: the Jacobi smoother is FP64 / FP32 only");
  if (a->a.n != a->a.m) return fail(SC23AD_BAD_PARAM, "hier_build: matrix must be square");
  AmgParams prm = to_params(p);
  prm.device_setup = device_setup;
  try {
    auto st = std::make_shared<Sc23adSetupState>();
    st->fine = a->a;
    st->prm = prm;
    st->chain = build_chain(a->a, prm);
    for (int k = 0; k < count; ++k) {
      outs[k] = derive_hierarchy(st, prm, precisions[k]).release();
    }
  } catch (const SetupError& e) {
    for (int k = 0; k < count; ++k) {
      delete outs[k];
      outs[k] = nullptr;
    }
    return fail(e.code, "hier_build: " + e.what);
  } catch (const std::bad_alloc&) {
    for (int k = 0; k < count; ++k) {
      delete outs[k];
      outs[k] = nullptr;
    }
    return fail(SC23AD_SETUP_FAIL, "hier_build: out of host memory");
  } catch (const DevError& e) {
    for (int k = 0; k < count; ++k) {
      delete outs[k];
      outs[k] = nullptr;
    }
    return fail(e.code, "hier_build: " + e.what);
  }
  return SC23AD_OK;
}

}  // namespace sc23ad

extern "C" {

int sc23ad_hier_build(const sc23ad_csrmat* a, const sc23ad_amg_param* p,
                      int precision, sc23ad_hier** out) {
  if (!out) return fail(SC23AD_BAD_PARAM, "hier_build: bad arguments");
  return sc23ad_hier_build_multi(a, p, &precision, 1, out);
}

int sc23ad_hier_retune(const sc23ad_hier* base, const double* level_omega,
                       const int* level_bits, int nlevels, int precision,
                       sc23ad_hier** out) {
  if (!base || !out || nlevels < 0 || precision < 0 || precision > 2)
    return fail(SC23AD_BAD_PARAM, "hier_retune: bad arguments");
  *out = nullptr;
  if (!base->setup) return fail(SC23AD_BAD_PARAM, "hier_retune: hierarchy has no setup state");
  AmgParams prm = base->setup->prm;
  if (level_omega) {
    prm.level_omega.assign(level_omega, level_omega + nlevels);
    for (double w : prm.level_omega)
      if (!(w > 0 && w < 2)) return fail(SC23AD_BAD_PARAM, "hier_retune: omega outside (0, 2)");
  }
  if (level_bits) prm.level_bits.assign(level_bits, level_bits + nlevels);
  try {
    *out = derive_hierarchy(base->setup, prm, precision).release();
  } catch (const SetupError& e) {
    return fail(e.code, "hier_retune: " + e.what);
  } catch (const std::bad_alloc&) {
    return fail(SC23AD_SETUP_FAIL, "hier_retune: out of host memory");
  }
  return SC23AD_OK;
}

int sc23ad_hier_retune_levels(const sc23ad_hier* base, const double* level_omega,
                              const int* level_vbits, int nlevels, int precision,
                              sc23ad_hier** out) {
  if (!base || !out || nlevels < 0 || precision < 0 || precision > 2)
    return fail(SC23AD_BAD_PARAM, "hier_retune_levels: bad arguments");
  *out = nullptr;
  if (!base->setup)
    return fail(SC23AD_BAD_PARAM, "hier_retune_levels: hierarchy has no setup state");
  AmgParams prm = base->setup->prm;
  if (level_omega) {
    prm.level_omega.assign(level_omega, level_omega + nlevels);
    for (double w : prm.level_omega)
      if (!(w > 0 && w < 2))
        return fail(SC23AD_BAD_PARAM, "hier_retune_levels: omega outside (0, 2)");
  }
  if (level_vbits) {
    if (level_vbits[0] != 0)
      return fail(SC23AD_BAD_PARAM, "hier_retune_levels: level 0 keeps the solver's F");
    prm.level_vbits.assign(level_vbits, level_vbits + nlevels + 1);
    for (int v : prm.level_vbits)
      if (v < -30 || v > 30)
        return fail(SC23AD_BAD_PARAM, "hier_retune_levels: vector bits outside [-30, 30]");
  }
  try {
    *out = derive_hierarchy(base->setup, prm, precision).release();
  } catch (const SetupError& e) {
    return fail(e.code, "hier_retune_levels: " + e.what);
  } catch (const std::bad_alloc&) {
    return fail(SC23AD_SETUP_FAIL, "hier_retune_levels: out of host memory");
  }
  return SC23AD_OK;
}

int sc23ad_hier_saturated(const sc23ad_hier* h, int* levels) {
  if (!h || !levels) return fail(SC23AD_BAD_PARAM, "hier_saturated: null");
  *levels = h->precision == SC23AD_INT ? h->ih.saturated : 0;
  return SC23AD_OK;
}

int sc23ad_hier_features(const sc23ad_hier* h, float* feat, int cap, int* nlevels) {
  if (!h || !nlevels) return fail(SC23AD_BAD_PARAM, "hier_features: null");
  if (!h->setup) return fail(SC23AD_BAD_PARAM, "hier_features: hierarchy has no setup state");
  const AmgChain& ch = h->setup->chain;
  *nlevels = ch.depth;
  if (!feat) return SC23AD_OK;
  for (int l = 0; l < ch.depth && l < cap; ++l) {
    const HostCsr& a = l == 0 ? h->setup->fine : ch.coarse[l - 1];
    level_features(a, l, ch.depth, feat + static_cast<size_t>(l) * kLevelFeatures);
  }
  return SC23AD_OK;
}

int sc23ad_hier_info(const sc23ad_hier* h, int* depth, int* level_n,
                     int* level_nnz) {
  if (!h) return fail(SC23AD_BAD_PARAM, "hier_info: null hierarchy");
  auto fill = [&](const auto& hh) {
    if (depth) *depth = hh.depth;
    for (int l = 0; l < hh.depth; ++l) {
      if (level_n) level_n[l] = hh.lv[l].a.n;
      if (level_nnz) level_nnz[l] = hh.lv[l].a.nnz();
    }
    if (level_n) level_n[hh.depth] = hh.n_coarse;
    if (level_nnz) level_nnz[hh.depth] = hh.lc.nnz() + hh.uc.nnz();
  };
  if (h->precision == SC23AD_INT)
    fill(h->ih);
  else if (h->precision == SC23AD_FP64)
    fill(h->dh);
  else
    fill(h->fh);
  return SC23AD_OK;
}

int sc23ad_hier_export(const sc23ad_hier* h, const char* path) {
  if (!h || !path) return fail(SC23AD_BAD_PARAM, "hier_export: bad arguments");
  bool ok = h->precision == SC23AD_INT    ? write_hierarchy(path, h->ih)
            : h->precision == SC23AD_FP64 ? write_hierarchy(path, h->dh)
                                          : write_hierarchy(path, h->fh);
  return ok ? SC23AD_OK : fail(SC23AD_IO_ERROR, "hier_export: write failed");
}

void sc23ad_hier_destroy(sc23ad_hier* h) { delete h; }

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t sc23ad_checksum_i32(const int32_t* v, int n) {
  uint64_t s = 0;
  for (int i = 0; i < n; ++i)
    s += mix64(static_cast<uint64_t>(i) * 0x9E3779B97F4A7C15ULL ^
               static_cast<uint64_t>(static_cast<uint32_t>(v[i])));
  return s;
}

}  // extern "C"
