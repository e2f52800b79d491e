"""Writes tests/golden/reference_convdiff_batch.json: what the reference,
compiled in this container (oracle/_ref/libref.so), does on a sample of the
batched convection-diffusion workload (2D 5-point first-order upwind, side,
Peclet number and flow angle drawn from seed 3, right-hand side k of system i
from x* seed 1000 (i + 1) + k, row-scaled; BlockSOR(40), int-AMG-FGMRES(20),
eps 1e-8, F by the driver's rule).

About a tenth of that workload's integer solves stop with OVERFLOW in the
product and in the oracle: the int64 sum of squares of the Arnoldi vector
wraps and lsqrt(<= 0) = 0 is then a divisor (gmres.hpp:243-244), where the
reference's own process dies.  The sample holds such right-hand sides and
converging ones of the same systems; for each it records the reference's
outcome (converged: iterations and the hash of x, in the reference's
sequential order; crashed: the signal) and the reference's FP64 solver 0 on
the same system.

Run:  python tests/golden/make_batch_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "sc23-ad_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import sc23ad as S  # noqa: E402
import oracle as O  # noqa: E402
from make_golden import _ref_solve_isolated, sha  # noqa: E402

# (system, right-hand side): the first six overflowing ones of the workload in
# the oracle and a converging right-hand side of three of the same systems
SAMPLE = [(1, 1), (1, 3), (3, 0), (4, 0), (6, 2), (13, 1), (1, 0), (3, 1), (4, 1)]
M, EPS = 20, 1e-8


def system(i):
    ns, pes, angs = S.convdiff_draws(256, 3)
    a = S.CsrMat.generate(S.GEN_CONVDIFF2D, int(ns[i]), float(pes[i]), float(angs[i]))
    rptr, cind, val = a.arrays()
    rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
    return a, rmax, (int(ns[i]), float(pes[i]), float(angs[i]))


def rhs(a, rmax, i, k):
    return a.matvec(S.xstar(a.n, 1000 * (i + 1) + k)) / rmax


def main() -> None:
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so is missing: run make -C oracle ref")
    out = {"m": M, "eps": EPS, "smoother": 1, "cases": []}
    for i, k in SAMPLE:
        a, rmax, draw = system(i)
        b = rhs(a, rmax, i, k)
        a.lmax1_scaling(np.ones(a.n))
        arrs = a.arrays()
        rec = {"system": i, "rhs": k, "n": draw[0], "peclet": draw[1], "angle": draw[2],
               "a_sha": sha(*arrs), "b_sha": sha(b)}
        r = _ref_solve_isolated(arrs, b, 2, 1, m=M, eps=EPS, trace_cap=0)
        rec["int"] = r if "crashed" in r else {
            "iters": r["iters"], "converged": r["converged"], "bit": r["bit"],
            "relres": r["relres"], "x_sha": sha(r["x"])}
        r0 = _ref_solve_isolated(arrs, b, 0, 1, m=M, eps=EPS, trace_cap=0)
        rec["fp64"] = r0 if "crashed" in r0 else {
            "iters": r0["iters"], "converged": r0["converged"], "relres": r0["relres"]}
        out["cases"].append(rec)
        print(i, k, draw, "int", rec["int"], "fp64", rec["fp64"], flush=True)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "reference_convdiff_batch.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
