// Host-side sparse containers and the small utilities the AMG setup shares.
//
// HostCsr is the library's analogue of the reference's CSRMat
// (Library/matrix/csrmat.hpp:13-139): row pointers, column indices and FP64
// values, rows sorted by column.  Everything in this directory runs once per
// matrix, before any solve, and is compiled with strict IEEE flags
// (-O2 -ffp-contract=off, no fast-math) because the setup's FP64 operation
// order decides the integer hierarchy bit for bit.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace sc23ad {

struct HostCsr {
  int n = 0;  // rows
  int m = 0;  // columns
  std::vector<int> rptr;
  std::vector<int> cind;
  std::vector<double> val;

  int nnz() const { return rptr.empty() ? 0 : rptr[n]; }
  void reset(int rows, int cols) {
    n = rows;
    m = cols;
    rptr.assign(static_cast<size_t>(rows) + 1, 0);
    cind.clear();
    val.clear();
  }
};

// Fixed-point sparse matrix as the solve phase consumes it.  `bit` is the
// fractional bit count of the off-diagonal (or all) values, `dbit` that of the
// stored inverse diagonal for triangular factors (-1 when unused).
struct IntCsr {
  int n = 0, m = 0;
  int bit = 0, dbit = -1;
  std::vector<int> rptr, cind;
  std::vector<int32_t> val;
  int nnz() const { return rptr.empty() ? 0 : rptr[n]; }
};

template <typename T>
struct FpCsr {
  int n = 0, m = 0;
  std::vector<int> rptr, cind;
  std::vector<T> val;
  int nnz() const { return rptr.empty() ? 0 : rptr[n]; }
};

// Status codes shared with the C ABI (mirrors include/sc23ad.h).
enum Status : int {
  kOk = 0,
  kBadParam = 1,
  kSetupFail = 2,
  kOverflow = 3,
  kStagnation = 4,
  kHipError = 5,
  kNotConverged = 6,
  kIoError = 7,
  kNoDevice = 8,
};

struct SetupError {
  int code;
  std::string what;
};

// CSRMat::LMax1Scaling (Library/matrix/csrmat.hpp:70-80): divide each row and
// the matching entry of b by the row's largest |a_ij|.
void lmax1_scaling(HostCsr& a, double* b);

// Row-wise maximum of |a_ij| (used by the coarse-level scaling of
// Library/preconditioner/INTAMG.hpp:176-211).
std::vector<double> row_abs_max(const HostCsr& a);

// Recursive Hoare partition sort keyed on `key`, co-permuting every array in
// `co`.  Same pivot choice (middle element), strict comparisons and swap
// sequence as Library/utils/sort.hpp:31-51, because its tie order decides the
// C/F splitting.
void hoare_sort(int left, int right, bool ascending, int* key,
                std::vector<int*> const& co);

// The same sort with an arbitrary key type; swap_co(a, b) exchanges positions
// a and b of every co-sorted array.
template <typename Key, typename SwapCo>
void hoare_sort_with(int left, int right, bool ascending, Key* key,
                     SwapCo&& swap_co) {
  int lo = left, hi = right;
  const Key pivot = key[(left + right) / 2];
  for (;;) {
    if (ascending) {
      while (key[lo] < pivot) ++lo;
      while (pivot < key[hi]) --hi;
    } else {
      while (key[lo] > pivot) ++lo;
      while (pivot > key[hi]) --hi;
    }
    if (lo >= hi) break;
    Key t = key[lo];
    key[lo] = key[hi];
    key[hi] = t;
    swap_co(lo, hi);
    ++lo;
    --hi;
  }
  if (left < lo - 1) hoare_sort_with(left, lo - 1, ascending, key, swap_co);
  if (hi + 1 < right) hoare_sort_with(hi + 1, right, ascending, key, swap_co);
}

// Fixed-point bit rules.
//  flex rule   (Library/matrix/spmat.hpp:108, trimat/ldtri.hpp:144-145):
//     30 if max < 1 else 32 - (ceil(log2(floor(max)+1)) + 1)
//  tri rule    (Library/matrix/trimat/ltri.hpp:106-108, dutri.hpp:120-125):
//     smallest b in [0,30] with max * 2^b >= 2^30, capped at 30
int flex_bits(double max_abs);
int tri_bits(double max_abs);

// C-style truncating conversion of v * 2^bit (Library/utils/array.hpp:43-49,
// spmat.hpp:117).  Out-of-range values follow x86 cvttsd2si (INT32_MIN).
int32_t to_fixed(double v, int bit);

}  // namespace sc23ad
