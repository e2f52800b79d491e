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

Run:  python tests/golden/make_golden.py            (reference.json)
      python tests/golden/make_golden.py --omega    (reference_omega.json)
      python tests/golden/make_golden.py --large    (reference_laplace3d_128.json)
      python tests/golden/make_golden.py --inclusions K SOLVER
                                       (reference_inclusions160_SOLVER_K.json)

--omega records the same quantities for SOR / BlockSOR weights other than 1
(the LD/U split of stationary.hpp:375-396 and :941-978 then carries a
non-zero (1 - 1/omega) a_ii entry in U and a_ii / omega in LD).
--large records the full-size case of the plain GPU workload: the 128^3
7-point Laplacian, x* from seed 1, BlockSOR(40), FGMRES(30), eps 1e-8, F
from the driver's rule: the integer hierarchy's hash, the V-cycle checksums
of every Arnoldi step and the hash of x (about ten minutes on one core).
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


def _solve_rec(r, solver):
    s = {"iters": r["iters"], "converged": r["converged"], "relres": r["relres"],
         "x_sha": sha(r["x"])}
    if solver == 2:
        s["bit"] = r["bit"]
        s["trace"] = [format(int(v), "016x") for v in r["trace"]]
    return s


# name, generator, params, seed, restart m (systems of the omega fixtures)
OMEGA_CASES = [CASES[2], CASES[3], CASES[5]]
OMEGAS = (0.8, 1.3, 1.7)


def _ref_solve_isolated(*args, **kw):
    """ref_solve in a forked child: with some weights the reference's integer
    solver overflows (signed wrap in blas1::Add / Nrm2, then lsqrt(<= 0) = 0
    and a division by zero at gmres.hpp:244) and the process dies.  Returns
    the result, or {"crashed": signal number} when the child was killed."""
    import multiprocessing as mp
    ctx = mp.get_context("fork")
    rd, wr = ctx.Pipe(duplex=False)

    def child():
        r = O.ref_solve(*args, **kw)
        wr.send({k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in r.items()})
        wr.close()

    p = ctx.Process(target=child)
    p.start()
    wr.close()
    try:
        r = rd.recv()
    except EOFError:
        r = None
    p.join()
    if r is None or p.exitcode != 0:
        return {"crashed": -p.exitcode if p.exitcode and p.exitcode < 0 else p.exitcode}
    r["x"] = np.asarray(r["x"])
    r["trace"] = np.asarray(r["trace"], np.uint64)
    return r


def main_omega() -> None:
    out = {"eps": EPS, "omegas": list(OMEGAS), "cases": []}
    tmp = tempfile.mkdtemp()
    for name, kind, params, seed, m in OMEGA_CASES:
        sy = S.make_system(kind, params, seed)
        arrs = sy.a.arrays()
        rec = {"name": name, "kind": kind, "params": list(params), "seed": seed, "m": m,
               "a_sha": sha(*arrs), "hier": {}, "solves": {}}
        for om in OMEGAS:
            for sm in (0, 1):
                for prec in (0, 2):
                    path = os.path.join(tmp, f"h_{prec}_{sm}.bin")
                    O.ref_build_hier(arrs, prec, sm, path, omega=om)
                    rec["hier"][f"{om}_{prec}_{sm}"] = file_sha(path)
                for solver in (2, 0):
                    r = _ref_solve_isolated(arrs, sy.b, solver, sm, m=m, eps=EPS, omega=om,
                                            trace_cap=8192 if solver == 2 else 0)
                    rec["solves"][f"{om}_{solver}_{sm}"] = (
                        r if "crashed" in r else _solve_rec(r, solver))
        out["cases"].append(rec)
        print(name, {k: (v.get("iters"), v.get("bit"), v.get("crashed"))
                     for k, v in rec["solves"].items()}, flush=True)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_omega.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


def main_large() -> None:
    name, kind, params, seed, m, eps, sm = "laplace3d_128", S.GEN_LAPLACE3D, (128,), 1, 30, \
        1e-8, 1
    sy = S.make_system(kind, params, seed)
    arrs = sy.a.arrays()
    tmp = tempfile.mkdtemp()
    path = os.path.join(tmp, "h.bin")
    O.ref_build_hier(arrs, 2, sm, path)
    hsha = file_sha(path)
    os.unlink(path)
    print(name, "hierarchy", hsha, flush=True)
    r = O.ref_solve(arrs, sy.b, 2, sm, m=m, eps=eps, trace_cap=8192)
    rec = {"name": name, "kind": kind, "params": list(params), "seed": seed, "m": m,
           "eps": eps, "smoother": sm, "n": sy.a.n, "nnz": sy.a.nnz, "a_sha": sha(*arrs),
           "b_sha": sha(sy.b), "nrm_b": sy.nrm_b, "hier_int": hsha,
           "solve": _solve_rec(r, 2)}
    print(name, "iterations", r["iters"], "bit", r["bit"], "relres", r["relres"], flush=True)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "reference_laplace3d_128.json")
    with open(dst, "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote", dst)


def main_inclusions(k: int, solver: int) -> None:
    """Right-hand side k of the 160^3 inclusions workload (3D 7-point
    diffusion, k = 1 with 64 spheres of radius U[4, 12] cells and k = 1e6,
    seed 4; b = A x*, x* from seed 400 + k, row-scaled), BlockSOR(40),
    FGMRES(30), eps 1e-8, F by the driver's rule: the reference's outcome
    (iterations, x hash, V-cycle checksums; or the signal that ended it) for
    solver 2, or iterations and residual for solver 0.  Writes
    reference_inclusions160_<solver>_<k>.json (about ten minutes each)."""
    a = S.CsrMat.generate(S.GEN_INCLUSIONS3D, 160, 64, 4.0, 12.0, 1e6, 4)
    rptr, cind, val = a.arrays()
    rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
    b = a.matvec(S.xstar(a.n, 400 + k)) / rmax
    a.lmax1_scaling(np.ones(a.n))
    arrs = a.arrays()
    r = _ref_solve_isolated(arrs, b, solver, 1, m=30, eps=1e-8,
                            trace_cap=8192 if solver == 2 else 0)
    rec = {"name": "inclusions3d_160", "rhs": k, "solver": solver, "m": 30, "eps": 1e-8,
           "smoother": 1, "n": a.n, "nnz": a.nnz, "a_sha": sha(*arrs), "b_sha": sha(b),
           "solve": r if "crashed" in r else _solve_rec(r, solver)}
    print("inclusions160", k, solver, {kk: v for kk, v in rec["solve"].items() if kk != "trace"},
          flush=True)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       f"reference_inclusions160_{solver}_{k}.json")
    with open(dst, "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote", dst)


def main() -> None:
    if not O.ref_available():
        sys.exit("oracle/_ref/libref.so is missing: run make -C oracle ref")
    if "--inclusions" in sys.argv:
        at = sys.argv.index("--inclusions")
        main_inclusions(int(sys.argv[at + 1]), int(sys.argv[at + 2]))
        return
    if "--omega" in sys.argv:
        main_omega()
        return
    if "--large" in sys.argv:
        main_large()
        return
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
