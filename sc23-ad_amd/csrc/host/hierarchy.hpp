// AMG hierarchy construction: the host-side half of the reference's
// preconditioner constructors.
//
//  * intAMGflex::intAMGflex  (Library/preconditioner/INTAMG.hpp:149-245)
//  * dfAMG<...,T>::dfAMG     (Library/preconditioner/INTAMG.hpp:25-84)
//  * SOR / BlockSOR splits   (Library/smoother/stationary.hpp:175-264,355-439,
//                             684-796,912-1021)
//  * direct::LU / LUflex     (Library/solver/lu.hpp:11-39,71-95) with the
//                             complete sparse LU helper::matrix::ilu
//                             (Library/helper/helper_matrix.hpp:978-1060)
//
// Level l = 0..depth-1 holds the operator A_l, the smoother factors LD_l
// (lower part plus diagonal / omega, inverse diagonal stored last in each row)
// and U_l (the rest, plus (1 - 1/omega) a_ii), the restriction R_l (level l to
// l+1) and the prolongation P_l (level l+1 to l).  Level `depth` is the
// coarsest grid, solved by the L / U factors `lc` / `uc` (uc stores the
// inverse diagonal first in each row).
#pragma once

#include <vector>

#include "csr.hpp"
#include "rs_setup.hpp"

namespace sc23ad {

// kJacobiSOR: JacobiSORflex / JacobiSOR<T> (Library/smoother/stationary.hpp:
// 447-562, 564-676): rows in consecutive blocks of kJacobiSorBlock; LD keeps
// the entries left of the row's block (plus the diagonal), so the rows of a
// block are independent of each other (the multicolour LD of ldtri::MCCSR*,
// matrix/trimat/ldtri.hpp:528-702).
// kJacobi: Jacobi<CSR<T>> (stationary.hpp:15-89), FP hierarchies only:
// out += d_inv (in - A out) with d_inv = omega / a_ii.  Its LD factor holds
// d_inv as a diagonal "inverse" and its U factor is A itself.
enum class SmootherKind : int { kSOR = 0, kBlockSOR = 1, kJacobiSOR = 2, kJacobi = 3 };
constexpr int kJacobiSorBlock = 1200;  // stationary.hpp:474, 587 (fixed there)
enum class Precision : int { kFP64 = 0, kFP32 = 1, kINT = 2 };

struct AmgParams {
  int max_levels = 20;   // AMGParam::level
  double theta = 0.8;    // AMGParam::theta (AMG_THETA in Work/main.cpp:32)
  int cycle = 1;         // AMGParam::cycle (V-cycle only; unused by the solve)
  SmootherKind smoother = SmootherKind::kSOR;
  double omega = 1.0;    // SORParam / BlockSORParam omega
  int iter = 1;          // sweeps per Smooth call
  int b_num = 40;        // BlockSORParam::b_num
  int coarse_n = 2000;   // coarsening stops once a level has <= coarse_n rows
  // Per-level overrides (learned parameters); empty = the reference's rule.
  std::vector<double> level_omega;  // SOR weight of level l's smoother
  std::vector<int> level_bits;      // cap on the fractional bits of A_l, U_l and
                                    // LD_l's off-diagonal part (<= 0: none)
  // Fractional bits of the level vectors relative to the fine level's F
  // (entry l for level l = 0..depth, entry 0 is 0): level l's vectors carry
  // F + level_vbits[l] bits.  The smoother, residual and coarse solves are
  // homogeneous in the vector scale, so only the restriction and
  // prolongation shifts change (by d = level_vbits[l+1] - level_vbits[l]).
  std::vector<int> level_vbits;
  bool device_setup = false;        // interpolation and Galerkin products on the GPU
};

// The unscaled Galerkin chain every precision shares: coarse[l], interp[l] and
// restrict_[l] for l = 0..depth-1 (coarse[l] is level l+1's operator).
struct AmgChain {
  std::vector<HostCsr> coarse, interp, restrict_;
  int depth = 0;
  int warnings = 0;
};

AmgChain build_chain(const HostCsr& fine, const AmgParams& prm);

struct IntLevel {
  IntCsr a, ld, u, r, p;
};
struct IntHierarchy {
  AmgParams prm;
  int depth = 0;
  int saturated = 0;  // levels whose requested bits exceeded the flex rule's maximum
  std::vector<IntLevel> lv;  // depth entries
  IntCsr lc, uc;             // coarsest-grid factors
  int n_coarse = 0;
};

template <typename T>
struct FpLevel {
  FpCsr<T> a, ld, u, r, p;
};
template <typename T>
struct FpHierarchy {
  AmgParams prm;
  int depth = 0;
  std::vector<FpLevel<T>> lv;
  FpCsr<T> lc, uc;
  int n_coarse = 0;
};

IntHierarchy build_int_hierarchy(const HostCsr& fine, const AmgChain& chain,
                                 const AmgParams& prm);
template <typename T>
FpHierarchy<T> build_fp_hierarchy(const HostCsr& fine, const AmgChain& chain,
                                  const AmgParams& prm);

// Pieces exposed for the unit tests.
void smoother_split(const HostCsr& a, const AmgParams& prm, HostCsr& ld,
                    HostCsr& u, int level = 0);
HostCsr extract_ld_inv(const HostCsr& ld);  // helper_matrix.hpp:634-664
HostCsr extract_l(const HostCsr& a);        // helper_matrix.hpp:606-632
HostCsr extract_dinv_u(const HostCsr& a);   // helper_matrix.hpp:698-728
HostCsr complete_lu(const HostCsr& a);      // helper_matrix.hpp:978-1060
// CSRflex (spmat.hpp:96-121); cap > 0 limits the bits (saturated = 1 when
// the cap exceeds the rule's maximum and is not applied).
IntCsr to_flex(const HostCsr& a, int cap = 0, int* saturated = nullptr);
IntCsr to_fixed_csr(const HostCsr& a, int bit);  // CSR<Fixed<bit>>
IntCsr to_ld_flex(const HostCsr& ld_inv, int cap = 0,
                  int* saturated = nullptr);  // ldtri::CSRflex (ldtri.hpp:127-165)
IntCsr to_l_flex(const HostCsr& l);         // ltri::CSRflex (ltri.hpp:95-118)
IntCsr to_dinvu_flex(const HostCsr& du);    // dutri::CSRflex (dutri.hpp:105-139)

}  // namespace sc23ad
