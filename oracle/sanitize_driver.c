/*
 * Test infrastructure: drives the oracle (oracle.c) under AddressSanitizer
 * and UndefinedBehaviorSanitizer.  The reference relies on signed wrap-around
 * in its integer kernels; the restatement does all such arithmetic in
 * unsigned types, so a clean UBSan run shows that no operation depends on
 * undefined behaviour, and a clean ASan run that every index stays in bounds.
 *
 *   sanitize_driver <hier_int.bin> <hier_fp64.bin> <system.bin> <bit>
 *
 * system.bin: int32 n, nnz; rptr[n+1], cind[nnz] int32; val[nnz], b[n] f64
 * (the row-scaled system the hierarchies were built from).  Runs one V-cycle
 * of each hierarchy, every level operation of the integer one, and an
 * integer solve with <bit> fractional bits and an FP64 solve.  Prints both
 * solves' statuses and iteration counts; exits 0 when the V-cycles and level
 * operations returned OR_OK (a solve may end in OR_ERR_OVERFLOW by design).
 */
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static void* xmalloc(size_t b) {
  void* p = malloc(b ? b : 1);
  if (!p) exit(3);
  return p;
}

int main(int argc, char** argv) {
  if (argc < 5) return 2;
  or_hier *hi = NULL, *hd = NULL;
  if (or_hier_load(argv[1], &hi) != OR_OK || or_hier_load(argv[2], &hd) != OR_OK) return 4;
  FILE* f = fopen(argv[3], "rb");
  if (!f) return 5;
  int n = 0, nnz = 0;
  if (fread(&n, 4, 1, f) != 1 || fread(&nnz, 4, 1, f) != 1) return 5;
  int* rptr = xmalloc(sizeof(int) * (size_t)(n + 1));
  int* cind = xmalloc(sizeof(int) * (size_t)nnz);
  double* val = xmalloc(sizeof(double) * (size_t)nnz);
  double* b = xmalloc(sizeof(double) * (size_t)n);
  if (fread(rptr, 4, (size_t)n + 1, f) != (size_t)n + 1 ||
      fread(cind, 4, (size_t)nnz, f) != (size_t)nnz ||
      fread(val, 8, (size_t)nnz, f) != (size_t)nnz || fread(b, 8, (size_t)n, f) != (size_t)n)
    return 5;
  fclose(f);
  const int bit = atoi(argv[4]);
  int32_t* vi = xmalloc(sizeof(int32_t) * (size_t)n);
  int32_t* vo = xmalloc(sizeof(int32_t) * (size_t)n);
  double* di = xmalloc(sizeof(double) * (size_t)n);
  double* dout = xmalloc(sizeof(double) * (size_t)n);
  uint64_t s = 88172645463325252ULL;
  for (int i = 0; i < n; ++i) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    vi[i] = (int32_t)(int64_t)(s % (1u << 30)) - (1 << 29);
    di[i] = (double)vi[i] / (double)(1 << 29);
  }
  int rc = 0;
  rc |= or_vcycle(hi, vi, vo);
  rc |= or_vcycle(hd, di, dout);
  for (int l = 0; l <= hi->depth; ++l) {
    const int ln = or_hier_level_n(hi, l);
    int32_t* a = xmalloc(sizeof(int32_t) * (size_t)(ln > 0 ? ln : 1));
    int32_t* o = xmalloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < ln; ++i) a[i] = vi[i % n];
    if (l < hi->depth) {
      for (int op = 0; op <= 6; ++op) {
        if (op == 3) continue;  /* P maps level l+1 to l: below */
        for (int i = 0; i < n; ++i) o[i] = 0;
        rc |= or_level_op(hi, l, op, a, o);
      }
      rc |= or_level_op(hi, l, 3, a, o);  /* a covers level l+1 */
    } else {
      rc |= or_level_op(hi, l, 7, a, o);
    }
    free(a);
    free(o);
  }
  double* x = xmalloc(sizeof(double) * (size_t)n);
  or_result ri, rd;
  double nrm = or_nrm2_f64(b, n, OR_RED_SEQ);
  for (int i = 0; i < n; ++i) x[i] = 0;
  or_solve(hi, n, rptr, cind, val, b, nrm, 30, 1e-8, bit, 3, OR_RED_SEQ, x, NULL, 0, &ri);
  for (int i = 0; i < n; ++i) x[i] = 0;
  or_solve(hd, n, rptr, cind, val, b, nrm, 30, 1e-8, 0, 3, OR_RED_GPU, x, NULL, 0, &rd);
  printf("level ops %d; int solve status %d iters %d; fp64 solve status %d iters %d\n", rc,
         ri.status, ri.iters, rd.status, rd.iters);
  or_hier_free(hi);
  or_hier_free(hd);
  free(rptr); free(cind); free(val); free(b); free(vi); free(vo); free(di); free(dout); free(x);
  return rc == 0 ? 0 : 1;
}
