// Test-infrastructure harness around the reference library (suzuki-hpc/SC23-AD,
// header-only C++20).  It is compiled only by oracle/Makefile, directly
// against the reference headers where they lie (-I <reference>/Library), into
// oracle/_ref/.  Nothing from the reference is copied into this repository.
//
// What it provides:
//  * ref_build_hier(): builds the reference's intAMGflex / dfAMG<double|float>
//    with SOR, BlockSOR or JacobiSOR smoothers on a given (already row-scaled)
//    matrix and
//    writes the resulting hierarchy in the "SC23AMG1" format, so the product's
//    host setup can be compared with it byte for byte.
//  * ref_amg_*(): the reference V-cycle on given inputs.
//  * ref_solve(): the reference driver flow of Work/main.cpp:124-171 (and
//    :85-122 for solvers 0/1), with a checksum trace of every V-cycle input and
//    output of the integer solver.
//  * main(): a seq.exe / multi.exe stand-in for the CPU baseline that reads a
//    system file written by the benchmark and times Restarted::Solve exactly
//    as Work/main.cpp:167-169 does.
//
// Reference members are private; they are read through the standard
// explicit-instantiation access idiom (an explicit instantiation may name
// private members), which changes nothing in the reference code.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "timer.hpp"
#include "utils/utils.hpp"
#include "matrix/spmat.hpp"
#include "matrix/trimat/spmat.hpp"
#include "blas/blas1.hpp"
#include "solver/gmres.hpp"
#include "solver/irsolver.hpp"
#include "preconditioner/INTAMG.hpp"
#include "smoother/stationary.hpp"

using namespace senk;

namespace {

template <typename Tag, typename Tag::type M>
struct Access {
  friend typename Tag::type get(Tag) { return M; }
};
#define SC23_ACCESS(NAME, CLASS, TYPE, MEMBER)         \
  struct NAME {                                        \
    using type = TYPE CLASS::*;                        \
    friend type get(NAME);                             \
  };                                                   \
  template struct Access<NAME, &CLASS::MEMBER>;

// ---- matrix classes ----
#define SC23_CSR_FIELDS(PFX, CLASS, VT)                \
  SC23_ACCESS(PFX##_n, CLASS, int, N)                  \
  SC23_ACCESS(PFX##_m, CLASS, int, M)                  \
  SC23_ACCESS(PFX##_val, CLASS, VT*, val)              \
  SC23_ACCESS(PFX##_cind, CLASS, int*, cind)           \
  SC23_ACCESS(PFX##_rptr, CLASS, int*, rptr)

SC23_CSR_FIELDS(csrflex, CSRflex, int)
SC23_ACCESS(csrflex_bit, CSRflex, int, bit)
using CSRFix30 = CSR<Fixed<30>>;
SC23_CSR_FIELDS(csrfix, CSRFix30, int)
using CSRd = CSR<double>;
using CSRf = CSR<float>;
SC23_CSR_FIELDS(csrd, CSRd, double)
SC23_CSR_FIELDS(csrf, CSRf, float)

SC23_CSR_FIELDS(ldflex, ldtri::CSRflex, int)
SC23_ACCESS(ldflex_bit, ldtri::CSRflex, int, bit)
SC23_ACCESS(ldflex_dbit, ldtri::CSRflex, int, dbit)
SC23_CSR_FIELDS(bjflex, ldtri::BJCSRflex, int)
SC23_ACCESS(bjflex_bit, ldtri::BJCSRflex, int, bit)
SC23_ACCESS(bjflex_dbit, ldtri::BJCSRflex, int, dbit)
using LDd = ldtri::CSR<double>;
using LDf = ldtri::CSR<float>;
using BJd = ldtri::BJCSR<double>;
using BJf = ldtri::BJCSR<float>;
SC23_CSR_FIELDS(mcflex, ldtri::MCCSRflex, int)
SC23_ACCESS(mcflex_bit, ldtri::MCCSRflex, int, bit)
SC23_ACCESS(mcflex_dbit, ldtri::MCCSRflex, int, dbit)
using MCd = ldtri::MCCSR<double>;
using MCf = ldtri::MCCSR<float>;
SC23_CSR_FIELDS(mcd, MCd, double)
SC23_CSR_FIELDS(mcf, MCf, float)
SC23_CSR_FIELDS(ldd, LDd, double)
SC23_CSR_FIELDS(ldf, LDf, float)
SC23_CSR_FIELDS(bjd, BJd, double)
SC23_CSR_FIELDS(bjf, BJf, float)

SC23_CSR_FIELDS(lflex, ltri::CSRflex, int)
SC23_ACCESS(lflex_bit, ltri::CSRflex, int, bit)
SC23_CSR_FIELDS(uflex, dutri::CSRflex, int)
SC23_ACCESS(uflex_bit, dutri::CSRflex, int, bit)
SC23_ACCESS(uflex_dbit, dutri::CSRflex, int, dbit)
using Ld = ltri::CSR<double>;
using Lf = ltri::CSR<float>;
using Ud = dutri::CSR<double>;
using Uf = dutri::CSR<float>;
SC23_CSR_FIELDS(ld_l, Ld, double)
SC23_CSR_FIELDS(lf_l, Lf, float)
SC23_CSR_FIELDS(ud_u, Ud, double)
SC23_CSR_FIELDS(uf_u, Uf, float)

// ---- smoothers ----
using SORi = SOR<ldtri::CSRflex, CSRflex>;
using BSORi = BlockSORflex<ldtri::BJCSRflex, CSRflex>;
using SORd = SOR<ldtri::CSR<double>, CSR<double>>;
using SORf = SOR<ldtri::CSR<float>, CSR<float>>;
using BSORd = BlockSOR<ldtri::BJCSR<double>, CSR<double>>;
using BSORf = BlockSOR<ldtri::BJCSR<float>, CSR<float>>;
using JSORi = JacobiSORflex<ldtri::MCCSRflex, CSRflex>;
using JSORd = JacobiSOR<ldtri::MCCSR<double>, CSR<double>>;
using JSORf = JacobiSOR<ldtri::MCCSR<float>, CSR<float>>;
SC23_ACCESS(jsori_ld, JSORi, ldtri::SpMat<int>*, LD)
SC23_ACCESS(jsori_u, JSORi, SpMat<int>*, U)
SC23_ACCESS(jsord_ld, JSORd, ldtri::SpMat<double>*, LD)
SC23_ACCESS(jsord_u, JSORd, SpMat<double>*, U)
SC23_ACCESS(jsorf_ld, JSORf, ldtri::SpMat<float>*, LD)
SC23_ACCESS(jsorf_u, JSORf, SpMat<float>*, U)
using JACd = Jacobi<CSR<double>>;
using JACf = Jacobi<CSR<float>>;
SC23_ACCESS(jacd_lu, JACd, SpMat<double>*, LU)
SC23_ACCESS(jacd_dinv, JACd, double*, d_inv)
SC23_ACCESS(jacf_lu, JACf, SpMat<float>*, LU)
SC23_ACCESS(jacf_dinv, JACf, float*, d_inv)
SC23_ACCESS(sori_ld, SORi, ldtri::SpMat<int>*, LD)
SC23_ACCESS(sori_u, SORi, SpMat<int>*, U)
SC23_ACCESS(bsori_ld, BSORi, ldtri::SpMat<int>*, LD)
SC23_ACCESS(bsori_u, BSORi, SpMat<int>*, U)
SC23_ACCESS(sord_ld, SORd, ldtri::SpMat<double>*, LD)
SC23_ACCESS(sord_u, SORd, SpMat<double>*, U)
SC23_ACCESS(sorf_ld, SORf, ldtri::SpMat<float>*, LD)
SC23_ACCESS(sorf_u, SORf, SpMat<float>*, U)
SC23_ACCESS(bsord_ld, BSORd, ldtri::SpMat<double>*, LD)
SC23_ACCESS(bsord_u, BSORd, SpMat<double>*, U)
SC23_ACCESS(bsorf_ld, BSORf, ldtri::SpMat<float>*, LD)
SC23_ACCESS(bsorf_u, BSORf, SpMat<float>*, U)

// ---- LU ----
SC23_ACCESS(luflex_l, direct::LUflex, ltri::CSRflex*, L)
SC23_ACCESS(luflex_u, direct::LUflex, dutri::CSRflex*, U)
using LUd = direct::LU<double>;
using LUf = direct::LU<float>;
SC23_ACCESS(lud_l, LUd, ltri::CSR<double>*, L)
SC23_ACCESS(lud_u, LUd, dutri::CSR<double>*, U)
SC23_ACCESS(luf_l, LUf, ltri::CSR<float>*, L)
SC23_ACCESS(luf_u, LUf, dutri::CSR<float>*, U)

// ---- preconditioners ----
using IAmgSOR = intAMGflex<CSRflex, CSRflex, SORi>;
using IAmgBJ = intAMGflex<CSRflex, CSRflex, BSORi>;
SC23_ACCESS(ias_a, IAmgSOR, SpMat<int>*, A)
SC23_ACCESS(ias_ca, IAmgSOR, CSRflex**, C_A)
SC23_ACCESS(ias_p, IAmgSOR, CSRflex**, Intrpl)
SC23_ACCESS(ias_r, IAmgSOR, CSRflex**, Intrpl_t)
SC23_ACCESS(ias_sm, IAmgSOR, SORi**, smoother)
SC23_ACCESS(ias_lu, IAmgSOR, direct::LUflex*, lu_solver)
SC23_ACCESS(ias_depth, IAmgSOR, int, depth)
SC23_ACCESS(ias_lastn, IAmgSOR, int, last_N)
SC23_ACCESS(iab_a, IAmgBJ, SpMat<int>*, A)
SC23_ACCESS(iab_ca, IAmgBJ, CSRflex**, C_A)
SC23_ACCESS(iab_p, IAmgBJ, CSRflex**, Intrpl)
SC23_ACCESS(iab_r, IAmgBJ, CSRflex**, Intrpl_t)
SC23_ACCESS(iab_sm, IAmgBJ, BSORi**, smoother)
SC23_ACCESS(iab_lu, IAmgBJ, direct::LUflex*, lu_solver)
SC23_ACCESS(iab_depth, IAmgBJ, int, depth)
SC23_ACCESS(iab_lastn, IAmgBJ, int, last_N)
using IAmgJS = intAMGflex<CSRflex, CSRflex, JSORi>;
SC23_ACCESS(iaj_a, IAmgJS, SpMat<int>*, A)
SC23_ACCESS(iaj_ca, IAmgJS, CSRflex**, C_A)
SC23_ACCESS(iaj_p, IAmgJS, CSRflex**, Intrpl)
SC23_ACCESS(iaj_r, IAmgJS, CSRflex**, Intrpl_t)
SC23_ACCESS(iaj_sm, IAmgJS, JSORi**, smoother)
SC23_ACCESS(iaj_lu, IAmgJS, direct::LUflex*, lu_solver)
SC23_ACCESS(iaj_depth, IAmgJS, int, depth)
SC23_ACCESS(iaj_lastn, IAmgJS, int, last_N)

template <typename T, typename SM>
using DAmg = dfAMG<CSR<T>, SM, T>;
using DAmgSORd = DAmg<double, SORd>;
using DAmgBJd = DAmg<double, BSORd>;
using DAmgSORf = DAmg<float, SORf>;
using DAmgBJf = DAmg<float, BSORf>;
using DAmgJSd = DAmg<double, JSORd>;
using DAmgJSf = DAmg<float, JSORf>;
using DAmgJacd = DAmg<double, JACd>;
using DAmgJacf = DAmg<float, JACf>;
#define SC23_DAMG_FIELDS(PFX, CLASS, T, SM)                  \
  SC23_ACCESS(PFX##_a, CLASS, SpMat<T>*, A)                  \
  SC23_ACCESS(PFX##_ca, CLASS, CSR<T>**, C_A)                \
  SC23_ACCESS(PFX##_p, CLASS, CSR<T>**, Intrpl)              \
  SC23_ACCESS(PFX##_r, CLASS, CSR<T>**, Intrpl_t)            \
  SC23_ACCESS(PFX##_sm, CLASS, SM**, smoother)               \
  SC23_ACCESS(PFX##_lu, CLASS, direct::LU<T>*, lu_solver)    \
  SC23_ACCESS(PFX##_depth, CLASS, int, depth)                \
  SC23_ACCESS(PFX##_lastn, CLASS, int, last_N)
SC23_DAMG_FIELDS(das, DAmgSORd, double, SORd)
SC23_DAMG_FIELDS(dab, DAmgBJd, double, BSORd)
SC23_DAMG_FIELDS(fas, DAmgSORf, float, SORf)
SC23_DAMG_FIELDS(fab, DAmgBJf, float, BSORf)
SC23_DAMG_FIELDS(daj, DAmgJSd, double, JSORd)
SC23_DAMG_FIELDS(faj, DAmgJSf, float, JSORf)
SC23_DAMG_FIELDS(dac, DAmgJacd, double, JACd)
SC23_DAMG_FIELDS(fac, DAmgJacf, float, JACf)

// ---------------------------------------------------------------------------
struct Out {
  FILE* f;
  void raw(const void* p, size_t b) {
    if (b) fwrite(p, 1, b, f);
  }
  void i32(int v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
};

template <typename VT>
void put_mat(Out& o, int n, int m, const int* rptr, const int* cind,
             const VT* val, int bit, int dbit) {
  const int nnz = rptr[n];
  o.i32(n);
  o.i32(m);
  o.i32(nnz);
  o.i32(bit);
  o.i32(dbit);
  o.raw(rptr, 4 * (size_t)(n + 1));
  o.raw(cind, 4 * (size_t)nnz);
  o.raw(val, sizeof(VT) * (size_t)nnz);
}

#define PUT_CSR(o, obj, PFX, bit, dbit)                                        \
  put_mat(o, (obj).*get(PFX##_n()), (obj).*get(PFX##_m()),                     \
          (obj).*get(PFX##_rptr()), (obj).*get(PFX##_cind()),                  \
          (obj).*get(PFX##_val()), bit, dbit)

void put_header(Out& o, int precision, int depth, int smoother, int b_num,
                int n_coarse, double omega, double theta) {
  o.raw("SC23AMG1", 8);
  o.i32(precision);
  o.i32(depth);
  o.i32(smoother);
  o.i32(1);       // iter
  o.i32(b_num);
  o.i32(20);      // max_levels
  o.i32(2000);    // coarse_n
  o.i32(n_coarse);
  o.f64(omega);
  o.f64(theta);
}

CSRMat* make_mat(int n, const int* rptr, const int* cind, const double* val) {
  int nnz = rptr[n];
  double* v = utils::SafeMalloc<double>(nnz);
  int* c = utils::SafeMalloc<int>(nnz);
  int* r = utils::SafeMalloc<int>(n + 1);
  memcpy(v, val, sizeof(double) * nnz);
  memcpy(c, cind, sizeof(int) * nnz);
  memcpy(r, rptr, sizeof(int) * (n + 1));
  return new CSRMat(v, c, r, n, n, MMShape::General, MMType::Real);
}

template <typename IA, typename SM, typename LDT>
void dump_int(Out& o, IA* amg, int smoother, int b_num, double omega,
              double theta, CSRFix30* ia, SM** (IA::*sm_ptr),
              CSRflex** (IA::*ca_ptr), CSRflex** (IA::*p_ptr),
              CSRflex** (IA::*r_ptr), direct::LUflex* (IA::*lu_ptr),
              int(IA::*depth_ptr), int(IA::*lastn_ptr),
              ldtri::SpMat<int>*(SM::*sm_ld), SpMat<int>*(SM::*sm_u),
              int ldk) {
  const int depth = amg->*depth_ptr;
  put_header(o, 2, depth, smoother, b_num, amg->*lastn_ptr, omega, theta);
  for (int l = 0; l < depth; ++l) {
    if (l == 0) {
      PUT_CSR(o, *ia, csrfix, 30, -1);
    } else {
      CSRflex& c = *(amg->*ca_ptr)[l - 1];
      PUT_CSR(o, c, csrflex, c.*get(csrflex_bit()), -1);
    }
    SM* s = (amg->*sm_ptr)[l];
    ldtri::SpMat<int>* ld = s->*sm_ld;
    if (ldk == 1) {
      auto& L = *static_cast<ldtri::BJCSRflex*>(ld);
      PUT_CSR(o, L, bjflex, L.*get(bjflex_bit()), L.*get(bjflex_dbit()));
    } else if (ldk == 2) {
      auto& L = *static_cast<ldtri::MCCSRflex*>(ld);
      PUT_CSR(o, L, mcflex, L.*get(mcflex_bit()), L.*get(mcflex_dbit()));
    } else {
      auto& L = *static_cast<ldtri::CSRflex*>(ld);
      PUT_CSR(o, L, ldflex, L.*get(ldflex_bit()), L.*get(ldflex_dbit()));
    }
    CSRflex& u = *static_cast<CSRflex*>(s->*sm_u);
    PUT_CSR(o, u, csrflex, u.*get(csrflex_bit()), -1);
    CSRflex& r = *(amg->*r_ptr)[l];
    PUT_CSR(o, r, csrflex, r.*get(csrflex_bit()), -1);
    CSRflex& p = *(amg->*p_ptr)[l];
    PUT_CSR(o, p, csrflex, p.*get(csrflex_bit()), -1);
  }
  direct::LUflex* lu = amg->*lu_ptr;
  ltri::CSRflex& L = *(lu->*get(luflex_l()));
  dutri::CSRflex& U = *(lu->*get(luflex_u()));
  PUT_CSR(o, L, lflex, L.*get(lflex_bit()), -1);
  PUT_CSR(o, U, uflex, U.*get(uflex_bit()), U.*get(uflex_dbit()));
}

template <typename T, typename DA, typename SM, typename LDC, typename PFXN>
void dump_fp_common(Out&, DA*) {}

}  // namespace

// Copies of the dfAMG dump, one per (T, smoother) combination; the field tags
// differ per class so the four instantiations are spelled out by a macro.
#define SC23_DUMP_FP(FN, DA, PFX, T, SM, SMLD, SMU, LDCLASS, LDPFX, LPFX, UPFX, \
                     LUL, LUU, CSRPFX, PREC)                                   \
  static void FN(Out& o, DA* amg, int smoother, int b_num, double omega,       \
                 double theta, CSR<T>* a0) {                                   \
    const int depth = amg->*get(PFX##_depth());                                \
    put_header(o, PREC, depth, smoother, b_num, amg->*get(PFX##_lastn()),      \
               omega, theta);                                                  \
    for (int l = 0; l < depth; ++l) {                                          \
      CSR<T>& a = (l == 0) ? *a0 : *(amg->*get(PFX##_ca()))[l - 1];            \
      PUT_CSR(o, a, CSRPFX, 0, -1);                                            \
      SM* s = (amg->*get(PFX##_sm()))[l];                                      \
      auto& L = *static_cast<LDCLASS*>(s->*get(SMLD()));                       \
      PUT_CSR(o, L, LDPFX, 0, -1);                                             \
      CSR<T>& u = *static_cast<CSR<T>*>(s->*get(SMU()));                       \
      PUT_CSR(o, u, CSRPFX, 0, -1);                                            \
      CSR<T>& r = *(amg->*get(PFX##_r()))[l];                                  \
      PUT_CSR(o, r, CSRPFX, 0, -1);                                            \
      CSR<T>& p = *(amg->*get(PFX##_p()))[l];                                  \
      PUT_CSR(o, p, CSRPFX, 0, -1);                                            \
    }                                                                          \
    direct::LU<T>* lu = amg->*get(PFX##_lu());                                 \
    auto& L = *(lu->*get(LUL()));                                              \
    auto& U = *(lu->*get(LUU()));                                              \
    PUT_CSR(o, L, LPFX, 0, -1);                                                \
    PUT_CSR(o, U, UPFX, 0, -1);                                                \
  }

SC23_DUMP_FP(dump_das, DAmgSORd, das, double, SORd, sord_ld, sord_u, LDd, ldd,
             ld_l, ud_u, lud_l, lud_u, csrd, 0)
SC23_DUMP_FP(dump_dab, DAmgBJd, dab, double, BSORd, bsord_ld, bsord_u, BJd, bjd,
             ld_l, ud_u, lud_l, lud_u, csrd, 0)
SC23_DUMP_FP(dump_fas, DAmgSORf, fas, float, SORf, sorf_ld, sorf_u, LDf, ldf,
             lf_l, uf_u, luf_l, luf_u, csrf, 1)
SC23_DUMP_FP(dump_fab, DAmgBJf, fab, float, BSORf, bsorf_ld, bsorf_u, BJf, bjf,
             lf_l, uf_u, luf_l, luf_u, csrf, 1)
SC23_DUMP_FP(dump_daj, DAmgJSd, daj, double, JSORd, jsord_ld, jsord_u, MCd, mcd,
             ld_l, ud_u, lud_l, lud_u, csrd, 0)
SC23_DUMP_FP(dump_faj, DAmgJSf, faj, float, JSORf, jsorf_ld, jsorf_u, MCf, mcf,
             lf_l, uf_u, luf_l, luf_u, csrf, 1)

// Jacobi hierarchies: the "LD" factor is the diagonal d_inv (one entry per
// row) and the "U" factor is the smoother's LU = A, as the product stores them.
#define SC23_DUMP_JAC(FN, DA, PFX, T, SM, SMLU, SMDINV, LPFX, UPFX, LUL, LUU,    \
                      CSRPFX, PREC)                                            \
  static void FN(Out& o, DA* amg, int b_num, double omega, double theta,       \
                 CSR<T>* a0) {                                                 \
    const int depth = amg->*get(PFX##_depth());                                \
    put_header(o, PREC, depth, 3, b_num, amg->*get(PFX##_lastn()), omega,      \
               theta);                                                         \
    for (int l = 0; l < depth; ++l) {                                          \
      CSR<T>& a = (l == 0) ? *a0 : *(amg->*get(PFX##_ca()))[l - 1];            \
      PUT_CSR(o, a, CSRPFX, 0, -1);                                            \
      SM* s = (amg->*get(PFX##_sm()))[l];                                      \
      CSR<T>& lu = *static_cast<CSR<T>*>(s->*get(SMLU()));                     \
      const int n = lu.*get(CSRPFX##_n());                                     \
      std::vector<int> rp(n + 1), ci(n);                                       \
      for (int i = 0; i < n; ++i) rp[i] = ci[i] = i;                           \
      rp[n] = n;                                                               \
      put_mat(o, n, n, rp.data(), ci.data(), s->*get(SMDINV()), 0, -1);        \
      PUT_CSR(o, lu, CSRPFX, 0, -1);                                           \
      CSR<T>& r = *(amg->*get(PFX##_r()))[l];                                  \
      PUT_CSR(o, r, CSRPFX, 0, -1);                                            \
      CSR<T>& p = *(amg->*get(PFX##_p()))[l];                                  \
      PUT_CSR(o, p, CSRPFX, 0, -1);                                            \
    }                                                                          \
    direct::LU<T>* lu = amg->*get(PFX##_lu());                                 \
    auto& L = *(lu->*get(LUL()));                                              \
    auto& U = *(lu->*get(LUU()));                                              \
    PUT_CSR(o, L, LPFX, 0, -1);                                                \
    PUT_CSR(o, U, UPFX, 0, -1);                                                \
  }

SC23_DUMP_JAC(dump_dac, DAmgJacd, dac, double, JACd, jacd_lu, jacd_dinv, ld_l, ud_u, lud_l,
              lud_u, csrd, 0)
SC23_DUMP_JAC(dump_fac, DAmgJacf, fac, float, JACf, jacf_lu, jacf_dinv, lf_l, uf_u, luf_l,
              luf_u, csrf, 1)

namespace {

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Order-free checksum of an int32 vector: sum over i of
// mix64(i * golden ^ (uint32)v_i), mod 2^64.
uint64_t checksum_i32(const int* v, int n) {
  uint64_t s = 0;
  for (int i = 0; i < n; ++i)
    s += mix64((uint64_t)i * 0x9E3779B97F4A7C15ULL ^ (uint64_t)(uint32_t)v[i]);
  return s;
}

struct TracePre : public Preconditioner<int> {
  Preconditioner<int>* inner;
  int n;
  std::vector<uint64_t> sums;  // (in, out) per call
  void Precondition(int* in, int* out) override {
    inner->Precondition(in, out);
    sums.push_back(checksum_i32(in, n));
    sums.push_back(checksum_i32(out, n));
  }
};

// SOR and JacobiSOR take SORParam, BlockSOR BlockSORParam, Jacobi JacobiParam.
SmootherParam* make_sp(int smoother, double omega, int b_num) {
  if (smoother == 1) return new BlockSORParam(omega, 1, b_num);
  if (smoother == 3) return new JacobiParam(omega, 1);
  return new SORParam(omega, 1);
}

template <typename T>
Preconditioner<T>* make_damg(int smoother, CSRMat* m, CSR<T>* a, AMGParam* prm) {
  if constexpr (std::is_same_v<T, double>) {
    if (smoother == 0) return new DAmgSORd(m, a, prm);
    if (smoother == 1) return new DAmgBJd(m, a, prm);
    if (smoother == 3) return new DAmgJacd(m, a, prm);
    return new DAmgJSd(m, a, prm);
  } else {
    if (smoother == 0) return new DAmgSORf(m, a, prm);
    if (smoother == 1) return new DAmgBJf(m, a, prm);
    if (smoother == 3) return new DAmgJacf(m, a, prm);
    return new DAmgJSf(m, a, prm);
  }
}

Preconditioner<int>* make_iamg(int smoother, CSRMat* m, CSRFix30* a, AMGParam* prm) {
  if (smoother == 0) return new IAmgSOR(m, a, prm);
  if (smoother == 1) return new IAmgBJ(m, a, prm);
  return new IAmgJS(m, a, prm);
}

}  // namespace

extern "C" {

// Builds the reference hierarchy for solver `precision` (0 FP64, 1 FP32,
// 2 INT) and smoother (0 SOR, 1 BlockSOR(omega, 1, b_num), 2 JacobiSOR) on the given
// row-scaled matrix; writes it to `path`.  Returns depth.
int ref_build_hier(int n, const int* rptr, const int* cind, const double* val,
                   int precision, int smoother, double omega, int b_num,
                   double theta, const char* path) {
  CSRMat* m = make_mat(n, rptr, cind, val);
  AMGParam* prm = new AMGParam(20, theta, 1, make_sp(smoother, omega, b_num));
  FILE* f = fopen(path, "wb");
  if (!f) return -1;
  Out o{f};
  int depth = 0;
  if (smoother == 3 && precision == 2) {
    fclose(f);
    return -1;  // the reference's integer Jacobi needs Fixed<bit> matrices
  }
  if (smoother == 3) {
    if (precision == 0) {
      auto* a = new CSR<double>(m);
      auto* amg = new DAmgJacd(m, a, prm);
      dump_dac(o, amg, b_num, omega, theta, a);
      depth = amg->*get(dac_depth());
    } else {
      auto* a = new CSR<float>(m);
      auto* amg = new DAmgJacf(m, a, prm);
      dump_fac(o, amg, b_num, omega, theta, a);
      depth = amg->*get(fac_depth());
    }
  } else if (precision == 2) {
    auto* ia = new CSRFix30(m);
    if (smoother == 0) {
      auto* amg = new IAmgSOR(m, ia, prm);
      dump_int<IAmgSOR, SORi, ldtri::CSRflex>(
          o, amg, 0, b_num, omega, theta, ia, get(ias_sm()), get(ias_ca()),
          get(ias_p()), get(ias_r()), get(ias_lu()), get(ias_depth()),
          get(ias_lastn()), get(sori_ld()), get(sori_u()), 0);
      depth = amg->*get(ias_depth());
      delete amg;
    } else if (smoother == 2) {
      auto* amg = new IAmgJS(m, ia, prm);
      dump_int<IAmgJS, JSORi, ldtri::MCCSRflex>(
          o, amg, 2, b_num, omega, theta, ia, get(iaj_sm()), get(iaj_ca()),
          get(iaj_p()), get(iaj_r()), get(iaj_lu()), get(iaj_depth()),
          get(iaj_lastn()), get(jsori_ld()), get(jsori_u()), 2);
      depth = amg->*get(iaj_depth());
      delete amg;
    } else {
      auto* amg = new IAmgBJ(m, ia, prm);
      dump_int<IAmgBJ, BSORi, ldtri::BJCSRflex>(
          o, amg, 1, b_num, omega, theta, ia, get(iab_sm()), get(iab_ca()),
          get(iab_p()), get(iab_r()), get(iab_lu()), get(iab_depth()),
          get(iab_lastn()), get(bsori_ld()), get(bsori_u()), 1);
      depth = amg->*get(iab_depth());
      delete amg;
    }
    delete ia;
  } else if (precision == 0) {
    auto* a = new CSR<double>(m);
    if (smoother == 0) {
      auto* amg = new DAmgSORd(m, a, prm);
      dump_das(o, amg, 0, b_num, omega, theta, a);
      depth = amg->*get(das_depth());
    } else if (smoother == 2) {
      auto* amg = new DAmgJSd(m, a, prm);
      dump_daj(o, amg, 2, b_num, omega, theta, a);
      depth = amg->*get(daj_depth());
    } else {
      auto* amg = new DAmgBJd(m, a, prm);
      dump_dab(o, amg, 1, b_num, omega, theta, a);
      depth = amg->*get(dab_depth());
    }
  } else {
    auto* a = new CSR<float>(m);
    if (smoother == 0) {
      auto* amg = new DAmgSORf(m, a, prm);
      dump_fas(o, amg, 0, b_num, omega, theta, a);
      depth = amg->*get(fas_depth());
    } else if (smoother == 2) {
      auto* amg = new DAmgJSf(m, a, prm);
      dump_faj(o, amg, 2, b_num, omega, theta, a);
      depth = amg->*get(faj_depth());
    } else {
      auto* amg = new DAmgBJf(m, a, prm);
      dump_fab(o, amg, 1, b_num, omega, theta, a);
      depth = amg->*get(fab_depth());
    }
  }
  fclose(f);
  return depth;
}

struct RefIntAmg {
  CSRMat* m;
  CSRFix30* ia;
  Preconditioner<int>* pre;
  int n;
};

void* ref_int_amg_create(int n, const int* rptr, const int* cind,
                         const double* val, int smoother, double omega,
                         int b_num, double theta) {
  if (smoother == 3) return nullptr;  // no integer Jacobi (FP only)
  auto* h = new RefIntAmg;
  h->m = make_mat(n, rptr, cind, val);
  h->ia = new CSRFix30(h->m);
  h->n = n;
  AMGParam* prm = new AMGParam(20, theta, 1, make_sp(smoother, omega, b_num));
  h->pre = make_iamg(smoother, h->m, h->ia, prm);
  return h;
}

void ref_int_amg_apply(void* hv, const int* in, int* out) {
  auto* h = static_cast<RefIntAmg*>(hv);
  std::vector<int> tmp(in, in + h->n);
  h->pre->Precondition(tmp.data(), out);
}

void ref_int_amg_destroy(void* hv) {
  auto* h = static_cast<RefIntAmg*>(hv);
  delete h->pre;
  delete h->ia;
  h->m->Free();
  delete h->m;
  delete h;
}

// The driver flow of Work/main.cpp on an already row-scaled (A, b): builds the
// requested solver, runs Restarted::Solve from x = 0 and reports iterations,
// x, the explicit residual and (solver 2) the fractional bits used plus a
// checksum pair per V-cycle.  bit < 0 selects the main.cpp:153-160 rule.
int ref_solve(int n, const int* rptr, const int* cind, const double* val,
              const double* b_in, int solver, int smoother, int m, double eps,
              int bit, double omega, int b_num, double theta, double* x,
              int* iters, int* converged, double* relres, int* bit_used,
              uint64_t* trace, int trace_cap, int* trace_len) {
  if (smoother == 3 && solver == 2) return -1;  // no integer Jacobi (FP only)
  CSRMat* mat = make_mat(n, rptr, cind, val);
  std::vector<double> b(b_in, b_in + n);
  double nrm_b = blas1::Nrm2<double>(b.data(), n);
  auto* dA = new CSR<double>(mat);
  const int outer = 1000 / m;
  AMGParam* prm = new AMGParam(20, theta, 1, make_sp(smoother, omega, b_num));
  utils::Set<double>(0, x, n);
  Converge cv(false, 0);
  *trace_len = 0;
  if (solver == 0) {
    Preconditioner<double>* M = make_damg<double>(smoother, mat, dA, prm);
    auto* inner = new FGMRES<double>(dA, m, eps, M);
    auto* s = new Restarted(dA, inner, outer, eps);
    cv = s->Solve(b.data(), nrm_b, x);
  } else if (solver == 1) {
    auto* fA = new CSR<float>(mat);
    Preconditioner<float>* M = make_damg<float>(smoother, mat, fA, prm);
    auto* inner = new FGMRES<float>(fA, m, eps, M);
    auto* s = new Restarted(dA, inner, outer, eps);
    cv = s->Solve(b.data(), nrm_b, x);
  } else {
    auto* iA = new CSRFix30(mat);
    Preconditioner<int>* iM = make_iamg(smoother, mat, iA, prm);
    if (bit < 0) {
      Preconditioner<double>* M = make_damg<double>(smoother, mat, dA, prm);
      std::vector<double> c(n);
      M->Precondition(b.data(), c.data());
      double c_max = 0;
      for (int i = 0; i < n; i++)
        if (std::fabs(c[i]) > c_max) c_max = std::fabs(c[i]);
      bit = (c_max / nrm_b < 1) ? 29
                                : 31 - ((int)std::ceil(std::log2(int(c_max / nrm_b) + 1)) + 1);
      delete M;
    }
    *bit_used = bit;
    TracePre tp;
    tp.inner = iM;
    tp.n = n;
    auto* inner = new FGMRESflex(iA, m, eps, &tp, (int8_t)bit);
    auto* s = new Restarted(dA, inner, outer, eps);
    cv = s->Solve(b.data(), nrm_b, x);
    int len = (int)tp.sums.size();
    if (len > trace_cap) len = trace_cap;
    for (int i = 0; i < len; ++i) trace[i] = tp.sums[i];
    *trace_len = len;
  }
  *iters = cv.iter;
  *converged = cv.isConverged ? 1 : 0;
  std::vector<double> r(n);
  dA->SpMV(x, r.data());
  blas1::Axpby<double>(1, b.data(), -1, r.data(), n);
  *relres = blas1::Nrm2<double>(r.data(), n) / nrm_b;
  return 0;
}

}  // extern "C"

#ifdef SC23_REF_MAIN
// Progress of the baseline run on stderr, one line per finished stage.
static double g_t0 = getTime();
static void stage(const char* what) {
  fprintf(stderr, "[ref] %s (%.1f s)\n", what, getTime() - g_t0);
  fflush(stderr);
}

// CPU baseline: <system file> <smoother 0|1|2> <solver 0|1|2> <m> [bit]
// The system file ("SC23SYS1", int32 n, nnz, rptr, cind, double val, b) holds
// the unscaled matrix and right-hand side; scaling and nrm_b follow
// Work/main.cpp:61-63.
int main(int argc, char** argv) {
  if (argc < 5) return 1;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  char magic[8];
  int n, nnz;
  if (fread(magic, 1, 8, f) != 8 || memcmp(magic, "SC23SYS1", 8)) return 3;
  if (fread(&n, 4, 1, f) != 1 || fread(&nnz, 4, 1, f) != 1) return 3;
  int* rptr = utils::SafeMalloc<int>(n + 1);
  int* cind = utils::SafeMalloc<int>(nnz);
  double* val = utils::SafeMalloc<double>(nnz);
  double* b = utils::SafeMalloc<double>(n);
  if (fread(rptr, 4, n + 1, f) != (size_t)(n + 1) ||
      fread(cind, 4, nnz, f) != (size_t)nnz ||
      fread(val, 8, nnz, f) != (size_t)nnz || fread(b, 8, n, f) != (size_t)n)
    return 3;
  fclose(f);
  stage("system read");
  CSRMat* Abase = new CSRMat(val, cind, rptr, n, n, MMShape::General, MMType::Real);
  int smoother_type = atoi(argv[2]);
  int solver_type = atoi(argv[3]);
  if (smoother_type == 3 && solver_type == 2) return 4;  // no integer Jacobi
  int m = atoi(argv[4]);
  Abase->LMax1Scaling(b);
  double nrm_b = blas1::Nrm2<double>(b, n);
  auto timer = new Timer();
  auto dA = new CSR<double>(Abase);
  const int outer = 1000 / m;
  const double epsilon = argc > 6 ? atof(argv[6]) : 1.0e-10;
  double* x = utils::SafeMalloc<double>(n);
  SmootherParam* sp = make_sp(smoother_type, 1.0, 40);
  Converge cv(false, 0);
  if (solver_type == 0) {
    Preconditioner<double>* M =
        make_damg<double>(smoother_type, Abase, dA, new AMGParam(20, 0.8, 1, sp));
    auto inner = new FGMRES<double>(dA, m, epsilon, M);
    auto s = new Restarted(dA, inner, outer, epsilon);
    utils::Set<double>(0, x, n);
    timer->Restart();
    cv = s->Solve(b, nrm_b, x);
    timer->Elapsed();
  } else if (solver_type == 1) {
    auto* fA = new CSR<float>(Abase);
    Preconditioner<float>* M =
        make_damg<float>(smoother_type, Abase, fA, new AMGParam(20, 0.8, 1, sp));
    auto inner = new FGMRES<float>(fA, m, epsilon, M);
    auto s = new Restarted(dA, inner, outer, epsilon);
    utils::Set<double>(0, x, n);
    timer->Restart();
    cv = s->Solve(b, nrm_b, x);
    timer->Elapsed();
  } else {
    auto iA = new CSRFix30(Abase);
    Preconditioner<int>* iM;
    Preconditioner<double>* M;
    M = make_damg<double>(smoother_type, Abase, dA, new AMGParam(20, 0.8, 1, sp));
    stage("FP64 AMG setup done");
    iM = make_iamg(smoother_type, Abase, iA, new AMGParam(20, 0.8, 1, sp));
    stage("integer AMG setup done");
    int bit;
    if (argc > 5 && atoi(argv[5]) > 0) {
      bit = atoi(argv[5]);
    } else {
      double* c = utils::SafeMalloc<double>(n);
      M->Precondition(b, c);
      double c_max = 0;
      for (int i = 0; i < n; i++)
        if (std::fabs(c[i]) > c_max) c_max = std::fabs(c[i]);
      bit = (c_max / nrm_b < 1) ? 29 : 31 - ((int)std::ceil(std::log2(int(c_max / nrm_b) + 1)) + 1);
      free(c);
    }
    delete M;
    stage("fractional bits chosen; solving");
    printf("[Bit] : %d\n", bit);
    auto inner = new FGMRESflex(iA, m, epsilon, iM, (int8_t)bit);
    auto s = new Restarted(dA, inner, outer, epsilon);
    utils::Set<double>(0, x, n);
    timer->Restart();
    cv = s->Solve(b, nrm_b, x);
    timer->Elapsed();
  }
  printf("[Iter] : %d\n", cv.iter);
  double* r = utils::SafeMalloc<double>(n);
  dA->SpMV(x, r);
  blas1::Axpby<double>(1, b, -1, r, n);
  double nrm_r = blas1::Nrm2<double>(r, n);
  printf("# [Res] : %e\n", nrm_r / nrm_b);
  return 0;
}
#endif
