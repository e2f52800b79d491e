// AMG setup on the GPU: the classical interpolation weights and the
// row-merge Galerkin products of one Ruge-Stueben coarsening step.
//
// Both keep the host's (and the reference's) FP64 operation order exactly:
// one thread computes one output row with the same sequence of additions,
// products and divisions as the sequential code, so every value is bit for
// bit the host's.  Device code is built with FP contraction off.
#include "dev_common.hpp"
#include "launch.hpp"

namespace sc23ad::dev {

namespace {

constexpr int kSetupBlock = 256;

// get_interpolation_matrix (helper_multigrid.hpp:457-532) for F rows:
//   w_ij = -(a_ij + sum_{m in D^s_i} a_im * a_mj / sum_{k in C^s_i} a_mk)
//          / sum_{n in weak_i} a_in,
// left to right in row order.  `fine` lists the F rows; P row i starts at
// pptr[i] and holds one value per strong C neighbour, in cs order.
__global__ __launch_bounds__(kSetupBlock) void k_interp_weights(
    int nfine, const int* __restrict__ fine, const int* __restrict__ arp,
    const int* __restrict__ aci, const double* __restrict__ av,
    const int* __restrict__ csp, const int* __restrict__ csc,
    const double* __restrict__ csv, const int* __restrict__ dsp,
    const int* __restrict__ dsc, const double* __restrict__ dsv,
    const int* __restrict__ wkp, const double* __restrict__ wkv,
    const int* __restrict__ pptr, double* pval, int* status) {
  const int f = blockIdx.x * kSetupBlock + threadIdx.x;
  if (f >= nfine) return;
  const int i = fine[f];
  double deno = 0;
  for (int q = wkp[i]; q < wkp[i + 1]; ++q) deno += wkv[q];
  const int cb = csp[i], ce = csp[i + 1];
  int out = pptr[i];
  for (int q = cb; q < ce; ++q) {
    double aij = csv[q];
    const int cj = csc[q];
    for (int r = dsp[i]; r < dsp[i + 1]; ++r) {
      const int mrow = dsc[r];
      const int mb = arp[mrow], me = arp[mrow + 1];
      double amj = 0;
      for (int t = mb; t < me; ++t)
        if (aci[t] == cj) {
          amj = av[t];
          break;
        }
      bool hit = false;
      double amk = 0;
      int mp = mb, kp = cb;
      while (mp < me && kp < ce) {
        const int am = aci[mp], ck = csc[kp];
        if (am < ck) {
          ++mp;
        } else if (am == ck) {
          amk += av[mp];
          hit = true;
          ++mp;
          ++kp;
        } else {
          ++kp;
        }
      }
      if (amk == 0) {
        if (!hit) {
          atomicOr(&status[0], 1);  // the reference exits here
        } else {
          atomicAdd(&status[1], 1);  // its "1 amk = 0" warning
        }
        amk = 0.00001;
      }
      aij += dsv[r] * amj / amk;
    }
    pval[out++] = -aij / deno;
  }
}

// Row-merge product out = L * R (helper_sparse.hpp:213-285), one thread per
// output row, for rows [row0, row1).  Terms alpha * R(k, c) reach column c in
// the order of k; a column absent so far takes alpha * x, a present one
// y + alpha * x, and a value that becomes exactly zero leaves the row (a later
// term restarts it from alpha * x).  The row's columns live in an
// open-addressing table of `cap` slots at the row's scratch offset; the
// surviving entries are then written, sorted by column, to the row's output
// slice (capacity = its term count) and their number to cnt[i - row0].
__global__ __launch_bounds__(kSetupBlock) void k_rowmerge(
    int row0, int row1, const int* __restrict__ lp, const int* __restrict__ lc,
    const double* __restrict__ lv, const int* __restrict__ rp,
    const int* __restrict__ rc, const double* __restrict__ rv,
    const long long* __restrict__ toff, const long long* __restrict__ ooff,
    long long base_t, long long base_o, int* tkey, double* tval, int* ocol, double* oval,
    int* cnt) {
  const int i = row0 + blockIdx.x * kSetupBlock + threadIdx.x;
  if (i >= row1) return;
  const long long t0 = toff[i] - base_t;
  const int cap = static_cast<int>(toff[i + 1] - toff[i]);
  int* key = tkey + t0;
  double* acc = tval + t0;
  for (int s = 0; s < cap; ++s) key[s] = -1;
  // key >= 0: column present with acc; key = -2 - c: column c seen, now zero
  for (int q = lp[i]; q < lp[i + 1]; ++q) {
    const double alpha = lv[q];
    const int k = lc[q];
    for (int p = rp[k]; p < rp[k + 1]; ++p) {
      const int c = rc[p];
      unsigned h = (static_cast<unsigned>(c) * 2654435761u) % static_cast<unsigned>(cap);
      while (key[h] != -1 && key[h] != c && key[h] != -2 - c) h = h + 1 == cap ? 0 : h + 1;
      const bool present = key[h] == c;
      const double v = present ? acc[h] + alpha * rv[p] : alpha * rv[p];
      if (v == 0.0) {
        key[h] = -2 - c;
      } else {
        key[h] = c;
        acc[h] = v;
      }
    }
  }
  // compact and insertion-sort by column
  const long long o0 = ooff[i] - base_o;
  int* oc = ocol + o0;
  double* ov = oval + o0;
  int n = 0;
  for (int s = 0; s < cap; ++s) {
    const int c = key[s];
    if (c < 0) continue;
    const double v = acc[s];
    int j = n++;
    while (j > 0 && oc[j - 1] > c) {
      oc[j] = oc[j - 1];
      ov[j] = ov[j - 1];
      --j;
    }
    oc[j] = c;
    ov[j] = v;
  }
  cnt[i - row0] = n;
}

// Gathers the compacted rows into CSR arrays; optr is local to the chunk
// (optr[0] = 0 for row0).
__global__ __launch_bounds__(kSetupBlock) void k_rowmerge_gather(
    int row0, int row1, const long long* __restrict__ ooff, long long base_o,
    const int* __restrict__ ocol, const double* __restrict__ oval,
    const int* __restrict__ optr, int* cind, double* val) {
  const int i = row0 + blockIdx.x * kSetupBlock + threadIdx.x;
  if (i >= row1) return;
  const long long src = ooff[i] - base_o;
  const int b = optr[i - row0], e = optr[i - row0 + 1];
  for (int k = 0; k < e - b; ++k) {
    cind[b + k] = ocol[src + k];
    val[b + k] = oval[src + k];
  }
}

int grid_for_rows(long long rows) {
  return static_cast<int>((rows + kSetupBlock - 1) / kSetupBlock);
}

}  // namespace

hipError_t launch_interp_weights(int nfine, const int* fine, const int* arp, const int* aci,
                                 const double* av, const int* csp, const int* csc,
                                 const double* csv, const int* dsp, const int* dsc,
                                 const double* dsv, const int* wkp, const double* wkv,
                                 const int* pptr, double* pval, int* status, hipStream_t s) {
  if (nfine <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_interp_weights, dim3(grid_for_rows(nfine)), dim3(kSetupBlock), 0, s,
                     nfine, fine, arp, aci, av, csp, csc, csv, dsp, dsc, dsv, wkp, wkv, pptr,
                     pval, status);
  return hipGetLastError();
}

hipError_t launch_rowmerge(int row0, int row1, const int* lp, const int* lc, const double* lv,
                           const int* rp, const int* rc, const double* rv,
                           const long long* toff, const long long* ooff, long long base_t,
                           long long base_o, int* tkey, double* tval, int* ocol, double* oval,
                           int* cnt, hipStream_t s) {
  if (row1 <= row0) return hipSuccess;
  hipLaunchKernelGGL(k_rowmerge, dim3(grid_for_rows(row1 - row0)), dim3(kSetupBlock), 0, s,
                     row0, row1, lp, lc, lv, rp, rc, rv, toff, ooff, base_t, base_o, tkey, tval,
                     ocol, oval, cnt);
  return hipGetLastError();
}

hipError_t launch_rowmerge_gather(int row0, int row1, const long long* ooff, long long base_o,
                                  const int* ocol, const double* oval, const int* optr,
                                  int* cind, double* val, hipStream_t s) {
  if (row1 <= row0) return hipSuccess;
  hipLaunchKernelGGL(k_rowmerge_gather, dim3(grid_for_rows(row1 - row0)), dim3(kSetupBlock), 0,
                     s, row0, row1, ooff, base_o, ocol, oval, optr, cind, val);
  return hipGetLastError();
}

}  // namespace sc23ad::dev
