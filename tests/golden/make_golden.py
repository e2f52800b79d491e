"""Writes tests/golden/reference.json from the reference compiled in this
container (oracle/_ref/libref.so, built by `make -C oracle ref` from the
reference headers where they lie).

For each small seeded system it records what the reference's own code
computes: the SHA-256 of every hierarchy file (per precision and smoother),
and for Restarted(FGMRES) with solvers 0/1/2 the iteration count, the
convergence flag, the explicit residual, the SHA-256 of x and, for the
integer solver, the checksum of every V-cycle input and output.  The systems
themselves come from this project's generators; their SHA-256 is recorded
so that a change of generator shows up as a fixture mismatch rather than as
a parity failure.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "sc23-ad_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import sc23ad as S  # noqa: E402
import oracle as O  # noqa: E402

# name, generator, params, seed, restart m
CASES = [
    ("poisson2d_32", S.GEN_POISSON2D, (32,), 0, 30),     # n <= 2000: coarse LU only
    ("poisson2d_48", S.GEN_POISSON2D, (48,), 0, 30),     # one coarse level (zeroed R)
    ("poisson2d_64", S.GEN_POISSON2D, (64,), 0, 30),
    ("laplace3d_16", S.GEN_LAPLACE3D, (16,), 1, 30),
    ("laplace3d_20", S.GEN_LAPLACE3D, (20,), 1, 10),
    ("convdiff2d_60", S.GEN_CONVDIFF2D, (60, 37.0, 1.2), 3, 20),
    ("diffusion27_16", S.GEN_DIFFUSION27, (16, 4, 1.0, 1.0, 1e-3, 2), 2, 30),
    ("inclusions3d_20", S.GEN_INCLUSIONS3D, (20, 6, 2.0, 4.0, 1e6, 4), 4, 30),
]
EPS = 1e-10


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def file_sha(path: str) -> str:
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main() -> None:
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so is missing: run make -C oracle ref")
    out = {"eps": EPS, "cases": []}
    tmp = tempfile.mkdtemp()
    for name, kind, params, seed, m in CASES:
        sy = S.make_system(kind, params, seed)
        arrs = sy.a.arrays()
        rec = {"name": name, "kind": kind, "params": list(params), "seed": seed, "m": m,
               "n": sy.a.n, "nnz": sy.a.nnz, "a_sha": sha(*arrs), "b_sha": sha(sy.b),
               "nrm_b": sy.nrm_b, "hier": {}, "solves": {}}
        for prec in (0, 1, 2):
            for sm in (0, 1, 2, 3):
                if sm == 3 and prec == 2:
                    continue  # Jacobi: FP precisions only
                path = os.path.join(tmp, f"h_{prec}_{sm}.bin")
                depth = O.ref_build_hier(arrs, prec, sm, path)
                rec["hier"][f"{prec}_{sm}"] = file_sha(path)
                rec["depth"] = depth
        for sm in (0, 1, 2, 3):
            for solver in ((0, 1) if sm == 3 else (2, 0, 1)):
                r = O.ref_solve(arrs, sy.b, solver, sm, m=m, eps=EPS,
                                trace_cap=8192 if solver == 2 else 0)
                s = {"iters": r["iters"], "converged": r["converged"],
                     "relres": r["relres"], "x_sha": sha(r["x"])}
                if solver == 2:
                    s["bit"] = r["bit"]
                    s["trace"] = [format(int(v), "016x") for v in r["trace"]]
                rec["solves"][f"{solver}_{sm}"] = s
        out["cases"].append(rec)
        print(name, "n", sy.a.n, "depth", rec["depth"],
              {k: (v["iters"], v.get("bit")) for k, v in rec["solves"].items()}, flush=True)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main()
