#!/usr/bin/env python3
"""Benchmark of the solve phase of int-AMG-FGMRES(m) on one MI355X.

Workload (default): the 3D 7-point Laplacian on a 128^3 grid (2,097,152 rows,
14,581,760 nonzeros, values 6 / -1, Dirichlet), x* ~ U[-1,1] from seed 1,
b = A x*, both row-scaled as the reference driver does; classical AMG
(theta 0.8) with the BlockSOR(40) smoother of the reference's multi.exe
(--smoother 0 selects seq.exe's SOR ordering), Restarted(FGMRES(30)) in
integer arithmetic (solver 2) with the driver's fractional-bit rule, stopped at
a relative residual of 1e-8.

One step = one full solve from x = 0 with b and x already resident in HBM.
The line printed on stdout is the benchmark record; progress goes to stderr.
Data are synthetic (generated), never downloaded.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "sc23-ad_amd"))

METRIC = ("solve time to 1e-8 rel. residual; int-SpMV HBM GB/s; "
          "learned/batched vs plain")
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_int_bytes(n: int, m: int, nnz: int) -> int:
    """Algorithmic HBM bytes of one fixed-point CSR SpMV: int32 values and
    column indices, row pointers, the gathered vector read once, the result
    written once."""
    return 8 * nnz + 4 * (n + 1) + 4 * m + 4 * n


def sptrsv_int_bytes(n: int, nnz_ld: int) -> int:
    """Algorithmic HBM bytes of one integer LD solve: int32 values and column
    indices of LD, row pointers, the right-hand side read once and the
    solution written once."""
    return 8 * nnz_ld + 4 * (n + 1) + 8 * n


SMOOTHER_NAMES = {0: "SOR", 1: "BlockSOR(40)", 2: "JacobiSOR(1200)"}


def _hier_sha(h) -> str:
    import hashlib
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    try:
        h.export(path)
        with open(path, "rb") as f:
            return hashlib.sha256(f.read()).hexdigest()
    finally:
        os.unlink(path)


def ld_nnz(rptr, cind, smoother: int, b_num: int = 40) -> int:
    """Entries of the smoother's LD factor (diagonal included): col <= i for
    SOR; left <= col <= i inside each of the b_num row blocks of ceil(N / b_num)
    rows for BlockSOR (smoother/stationary.hpp:934-978); col == i or col below
    the row's 1200-row block for JacobiSOR (stationary.hpp:586-631)."""
    n = len(rptr) - 1
    row = np.repeat(np.arange(n, dtype=np.int64), np.diff(rptr))
    col = np.asarray(cind, np.int64)
    if smoother == 0:
        return int(np.count_nonzero(col <= row))
    if smoother == 2:
        return int(np.count_nonzero((col == row) | (col < (row // 1200) * 1200)))
    size = -(-n // b_num)
    left = (row // size) * size
    return int(np.count_nonzero((col <= row) & (col >= left)))


def read_traffic(path: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each scaled by the factor
    the same counter gives for the library's copy kernel of known size)."""
    if not os.path.exists(path):
        return None
    txt = open(path).read()

    def num(key):
        m = re.search(r"^" + key + r":\s*([0-9.eE+-]+)", txt, re.M)
        return float(m.group(1)) if m else None

    f, w = num("FETCH_SIZE_KiB_per_dispatch"), num("WRITE_SIZE_KiB_per_dispatch")
    if f is None or w is None:
        return None
    cf = num("fetch_calibration_factor") or 1.0
    cw = num("write_calibration_factor") or 1.0
    return (f * cf + w * cw) * 1024.0


def calibrate_copy(mib: int, reps: int) -> None:
    """Copy kernel of known size for the traffic-counter calibration: each
    launch reads and writes mib MiB once."""
    import sc23ad as S
    nbytes = mib << 20
    src = S.DevArray(nbytes=nbytes)
    dst = S.DevArray(nbytes=nbytes)
    S.check(S.lib.sc23ad_dev_copy(dst.p, src.p, nbytes))
    for _ in range(reps):
        S.check(S.lib.sc23ad_dev_copy(dst.p, src.p, nbytes))
    S.check(S.lib.sc23ad_dev_sync())
    print(json.dumps({"copy_kernel": "k_copy_bytes", "bytes_read_per_launch": nbytes,
                      "bytes_written_per_launch": nbytes, "launches": reps + 1}), flush=True)


def write_system_file(path, a, b_unscaled):
    rptr, cind, val = a
    with open(path, "wb") as fh:
        fh.write(b"SC23SYS1")
        fh.write(np.asarray([len(rptr) - 1, len(cind)], np.int32).tobytes())
        fh.write(np.ascontiguousarray(rptr, np.int32).tobytes())
        fh.write(np.ascontiguousarray(cind, np.int32).tobytes())
        fh.write(np.ascontiguousarray(val, np.float64).tobytes())
        fh.write(np.ascontiguousarray(b_unscaled, np.float64).tobytes())


def cpu_baseline(args, unscaled_arrays, b_unscaled, smoother):
    """The reference (oracle/_ref, built from the reference sources with the
    reference Makefile's GNU flags) timed on this host: Restarted::Solve only,
    as Work/main.cpp times it.  Falls back to the oracle's CPU restatement
    (single thread) when the reference build is absent."""
    import oracle.oracle as O  # noqa: F401  (checker; cpu_baseline leg only)
    omp = smoother in (1, 2)   # BlockSOR and JacobiSOR are the OpenMP paths
    exe = O.MULTI_REF if omp else O.SEQ_REF
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    threads = min(threads, 16)
    if os.path.exists(exe):
        sysf = os.path.join("/tmp", f"sc23ad_sys_{os.getpid()}.bin")
        write_system_file(sysf, unscaled_arrays, b_unscaled)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(threads if omp else 1)
        cmd = [exe, sysf, str(smoother), "2", str(args.m), "0", repr(args.rtol)]
        log("cpu baseline:", " ".join(cmd))
        t0 = time.time()
        # The baseline prints one stderr line per finished stage; relay them.
        outf = sysf + ".out"
        with open(outf, "w") as fo:
            proc = subprocess.Popen(cmd, stdout=fo, stderr=subprocess.PIPE, text=True,
                                    env=env)
            for line in proc.stderr:
                log(line.rstrip())
            try:
                rc = proc.wait(timeout=max(1.0, args.cpu_timeout - (time.time() - t0)))
            except subprocess.TimeoutExpired:
                proc.kill()
                log("cpu baseline timed out")
                return None
            finally:
                os.unlink(sysf)
        with open(outf) as fo:
            stdout = fo.read()
        os.unlink(outf)
        wall = time.time() - t0
        tm = re.search(r"^\[Time\] : ([0-9.]+)", stdout, re.M)
        it = re.search(r"^\[Iter\] : (\d+)", stdout, re.M)
        rs = re.search(r"^# \[Res\] : ([0-9.eE+-]+)", stdout, re.M)
        if rc != 0 or not tm:
            log("cpu baseline failed", rc)
            return None
        return {"value": float(tm.group(1)), "unit": "s",
                "cores": threads if omp else 1, "kind": "reference",
                "sample": (f"full Restarted::Solve of the same system "
                           f"({'multi.exe' if omp else 'seq.exe'}, {SMOOTHER_NAMES[smoother]},"
                           f" solver 2, m={args.m}, eps={args.rtol}); "
                           f"iterations {it.group(1) if it else '?'}, explicit relres "
                           f"{rs.group(1) if rs else '?'}; process wall incl. setup {wall:.1f} s"),
                "iterations": int(it.group(1)) if it else None}
    return None


def run_batch(args):
    """Batched uneven systems: `--systems` 2D convection-diffusion systems
    (5-point first-order upwind; side uniform in [32, 256], Peclet in [1, 100],
    flow angle in [0, 2 pi) drawn from seed 3), 4 right-hand sides each
    (b = A x*, x* ~ U[-1,1]), int-AMG-FGMRES(20) to rtol.  Reports systems per
    second for all solves issued at once (each on its own stream) against the
    same solves one at a time, and checks that both give identical x."""
    import sc23ad as S
    K, R = args.systems, 4
    ns, pes, angs = S.convdiff_draws(K, 3)
    t0 = time.time()
    prm = S.amg_param(args.smoother)
    solvers, rhs, nrms, amgs = [], [], [], []
    rows = 0
    for i in range(K):
        a = S.CsrMat.generate(S.GEN_CONVDIFF2D, int(ns[i]), float(pes[i]), float(angs[i]))
        rptr, cind, val = a.arrays()
        rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
        bs = [a.matvec(S.xstar(a.n, 1000 * (i + 1) + k)) for k in range(R)]
        a.lmax1_scaling(bs[0].copy())
        bs = [b / rmax for b in bs]
        hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], prm)
        am, ad = S.Amg(hi), S.Amg(hd)
        for b in bs:
            nrm = float(np.sqrt(np.add.accumulate(b * b)[-1]))
            bit = ad.default_frac_bits(b, nrm)
            solvers.append(S.Solver(S.INT, a, am, m=args.m, eps=args.rtol, bit=bit))
            rhs.append(b)
            nrms.append(nrm)
        amgs.append((am, ad, a))
        rows += a.n
        if (i + 1) % 32 == 0:
            log(f"batch setup: {i + 1}/{K} systems ({time.time() - t0:.1f} s)")
    log(f"batch: {K} systems, {rows} rows, {len(solvers)} solves; setup {time.time() - t0:.1f} s")
    # warm-up (first launches, code objects)
    S.solve_batch(solvers[:16], rhs[:16], nrms[:16], threads=args.threads)
    t = time.perf_counter()
    xs_seq, cvs = [], []
    for s_, b, nb in zip(solvers, rhs, nrms):
        x, cv = s_.solve(b, nb)
        xs_seq.append(x)
        cvs.append(cv)
    t_seq = time.perf_counter() - t
    log(f"one at a time: {t_seq:.2f} s")
    t = time.perf_counter()
    xs_bat, cvb = S.solve_batch(solvers, rhs, nrms, threads=args.threads)
    t_bat = time.perf_counter() - t
    log(f"all at once: {t_bat:.2f} s")
    same = all(np.array_equal(p.view(np.uint64), q.view(np.uint64)) for p, q in zip(xs_seq, xs_bat))
    conv = sum(c.is_converged for c in cvb)
    iters = [c.iter for c in cvb]
    rec = {
        "metric": "batched uneven systems: systems per second (all at once vs one at a time)",
        "value": K / t_bat, "unit": "systems/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_bat * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32/int64 (FP64 outer refinement)",
        "data": "synthetic (2D convection-diffusion family, seed 3)",
        "config": {"workload": f"{K} 2D convection-diffusion systems (side 32..256), {R} RHS each, "
                               f"int-AMG-FGMRES({args.m}), rtol {args.rtol}",
                   "systems": K, "rhs_per_system": R, "rows_total": rows,
                   "threads": args.threads or 16},
        "one_at_a_time_systems_per_s": K / t_seq,
        "speedup_vs_one_at_a_time": t_seq / t_bat,
        "identical_results": bool(same), "converged": int(conv), "solves": len(cvb),
        "iterations_mean": float(np.mean(iters)), "iterations_max": int(np.max(iters)),
    }
    print(json.dumps(rec), flush=True)


def run_learned(args):
    """Learned per-level omega / bit caps against the reference's defaults on
    the 3D 27-point variable-coefficient diffusion system (side --n, default
    96: coefficient exp(2 g) on 8^3-cell blocks, anisotropy 1:1:1e-3 in z,
    seed 2), int-AMG-FGMRES(30) to rtol: seconds per solve for both, with b and
    x resident in HBM, and the iterations; the learned setup falls back to
    the defaults when its solve does not converge."""
    import sc23ad as S
    from sc23ad import learned as LR
    n = args.n
    t0 = time.time()
    sy = S.make_system(S.GEN_DIFFUSION27, (n, 8, 1.0, 1.0, 1e-3, 2), 2)
    prm = S.amg_param(args.smoother)
    bi, bd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], prm)
    log(f"diffusion27 {n}^3: N={sy.a.n}, depth {bi.info()[0]} ({time.time() - t0:.1f} s)")
    mlp = LR.load_mlp()
    ls = LR.learned_setup(bi, bd, mlp)
    log(f"learned omega {np.round(ls.omega, 3).tolist()} bits {ls.bits.tolist()} "
        f"saturated {ls.saturated}")
    b_dev = S.DevArray(sy.b)
    x_dev = S.DevArray(np.zeros(sy.a.n))
    zeros = np.zeros(sy.a.n)

    def timed(hint, hfp):
        bit = S.Amg(hfp).default_frac_bits(sy.b, sy.nrm_b)
        sol = S.Solver(S.INT, sy.a, S.Amg(hint), m=args.m, eps=args.rtol, bit=bit)
        for _ in range(args.warmup):
            x_dev.set(zeros)
            sol.solve_device(b_dev, sy.nrm_b, x_dev)
        ts = []
        for _ in range(args.steps):
            x_dev.set(zeros)
            S.check(S.lib.sc23ad_dev_sync())
            t = time.perf_counter()
            cv = sol.solve_device(b_dev, sy.nrm_b, x_dev)
            S.check(S.lib.sc23ad_dev_sync())
            ts.append(time.perf_counter() - t)
        rel = sol.relres(sy.b, sy.nrm_b, x_dev.get())
        return float(np.mean(ts)), cv, rel, bit

    t_def, cv_def, rel_def, bit_def = timed(bi, bd)
    log(f"default: {t_def * 1e3:.1f} ms, {cv_def.iter} iterations, relres {rel_def:.3e}")
    t_l, cv_l, rel_l, bit_l = timed(ls.hier_int, ls.hier_fp64)
    fallback = bool(ls.saturated > 0)
    if not (cv_l.is_converged and cv_l.status == S.OK):
        fallback = True
        t_l, cv_l, rel_l, bit_l = t_def, cv_def, rel_def, bit_def
    log(f"learned: {t_l * 1e3:.1f} ms, {cv_l.iter} iterations, relres {rel_l:.3e}, "
        f"fallback {fallback}")
    rec = {
        "metric": "learned per-level omega/bits vs reference defaults: solve time to 1e-8",
        "value": t_l, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_l * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32/int64 (FP64 outer refinement)",
        "data": "synthetic (27-point variable-coefficient diffusion, seed 2)",
        "config": {"workload": f"int-AMG-FGMRES({args.m}), 3D 27-point diffusion {n}^3 "
                               f"(N={sy.a.n}), anisotropy 1:1:1e-3, rtol {args.rtol}"},
        "default_s": t_def, "default_iterations": cv_def.iter, "default_relres": rel_def,
        "learned_iterations": cv_l.iter, "learned_relres": rel_l,
        "learned_omega": [float(v) for v in ls.omega], "learned_bits": [int(v) for v in ls.bits],
        "fallback_used": fallback, "gain_vs_default": t_def / t_l,
    }
    print(json.dumps(rec), flush=True)


def run_scale(args):
    """Learned + batched at scale: the 3D 7-point diffusion system on a 160^3
    grid (4,096,000 rows; k = 1 with 64 spherical inclusions of radius
    U[4, 12] cells and k = 1e6, seed 4), 16 right-hand sides b = A x*
    (x* ~ U[-1,1], seeds 400 + k), row-scaled as the reference driver does.
    Learned + batched: int32 fixed point with the network's per-level omega /
    bit caps, all 16 solves issued at once (one stream each), any solve that
    does not converge cleanly repeated with the default hierarchy.  Plain:
    FP64 (solver 0) and FP32-mixed (solver 1) AMG-FGMRES on the same GPU, one
    right-hand side after another; also the default integer solver one at a
    time and batched, which separate the two effects.  Time to solution for
    all right-hand sides, host arrays in and out for every method; each
    answer is judged by its explicit FP64 residual."""
    import sc23ad as S
    from sc23ad import learned as LR
    n = args.n if "--n" in sys.argv else 160
    R = args.rhs
    # The host setup of a 4M-row system takes minutes: let it report its stages.
    os.environ.setdefault("SC23AD_VERBOSE", "1")
    t0 = time.time()
    a = S.CsrMat.generate(S.GEN_INCLUSIONS3D, n, 64, 4.0, 12.0, 1e6, 4)
    rptr, cind, val = a.arrays()
    rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
    bs = [a.matvec(S.xstar(a.n, 400 + k)) for k in range(R)]
    a.lmax1_scaling(bs[0].copy())
    bs = [b / rmax for b in bs]
    nrms = [float(np.sqrt(np.add.accumulate(b * b)[-1])) for b in bs]
    prm = S.amg_param(args.smoother)
    hi, hd, hf = S.Hierarchy.build_multi(a, [S.INT, S.FP64, S.FP32], prm)
    log(f"inclusions {n}^3: N={a.n}, nnz={a.nnz}, depth {hi.info()[0]} "
        f"({time.time() - t0:.1f} s)")
    ls = LR.learned_setup(hi, hd, LR.load_mlp())
    log(f"learned omega {np.round(ls.omega, 3).tolist()} bits {ls.bits.tolist()} "
        f"saturated {ls.saturated} ({time.time() - t0:.1f} s)")
    amg_i, amg_d, amg_f = S.Amg(hi), S.Amg(hd), S.Amg(hf)
    amg_l, amg_ld = S.Amg(ls.hier_int), S.Amg(ls.hier_fp64)
    bits_d = [amg_d.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
    bits_l = [amg_ld.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
    sols_i = [S.Solver(S.INT, a, amg_i, m=args.m, eps=args.rtol, bit=bt) for bt in bits_d]
    sols_l = [S.Solver(S.INT, a, amg_l, m=args.m, eps=args.rtol, bit=bt) for bt in bits_l]
    sol_d = S.Solver(S.FP64, a, amg_d, m=args.m, eps=args.rtol)
    sol_f = S.Solver(S.FP32, a, amg_f, m=args.m, eps=args.rtol)
    log(f"setup done ({time.time() - t0:.1f} s); warm-up")
    for sol in (sols_i[0], sols_l[0], sol_d, sol_f):
        sol.solve(bs[0], nrms[0])

    def one_by_one(sols):
        t = time.perf_counter()
        out = [sol.solve(b, nb) for sol, b, nb in zip(sols, bs, nrms)]
        return time.perf_counter() - t, [x for x, _ in out], [c for _, c in out]

    def judge(sol, xs):
        return [sol.relres(b, nb, x) for b, nb, x in zip(bs, nrms, xs)]

    res = {}
    t, xs, cvs = one_by_one([sol_d] * R)
    res["fp64"] = (t, cvs, judge(sol_d, xs))
    log(f"plain FP64, one at a time: {t:.3f} s")
    t, xs, cvs = one_by_one([sol_f] * R)
    res["fp32"] = (t, cvs, judge(sol_f, xs))
    log(f"plain FP32-mixed, one at a time: {t:.3f} s")
    t, xs, cvs = one_by_one(sols_i)
    res["int"] = (t, cvs, judge(sols_i[0], xs))
    log(f"integer defaults, one at a time: {t:.3f} s")
    t = time.perf_counter()
    xs, cvs = S.solve_batch(sols_i, bs, nrms, threads=args.threads)
    res["int_batched"] = (time.perf_counter() - t, cvs, judge(sols_i[0], xs))
    log(f"integer defaults, batched: {res['int_batched'][0]:.3f} s")
    t = time.perf_counter()
    xs, cvs = S.solve_batch(sols_l, bs, nrms, threads=args.threads)
    redo = [k for k, c in enumerate(cvs) if not (c.is_converged and c.status == S.OK)]
    if redo:
        xr, cr = S.solve_batch([sols_i[k] for k in redo], [bs[k] for k in redo],
                               [nrms[k] for k in redo], threads=args.threads)
        for k, x_, c_ in zip(redo, xr, cr):
            xs[k], cvs[k] = x_, c_
    t_lb = time.perf_counter() - t
    res["learned_batched"] = (t_lb, cvs, judge(sols_l[0], xs))
    log(f"learned + batched: {t_lb:.3f} s (fallbacks {len(redo)})")

    def summary(key):
        t, cvs, rel = res[key]
        return {"s": t, "iterations": [int(c.iter) for c in cvs],
                "converged": int(sum(c.is_converged for c in cvs)),
                "explicit_relres_max": float(max(rel))}
    rec = {
        "metric": "learned + batched vs plain FP64 / FP32-mixed: time to solution for "
                  f"{R} right-hand sides to rtol {args.rtol}",
        "value": t_lb, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_lb * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32/int64 (FP64 outer refinement)",
        "data": "synthetic (3D diffusion with 64 spherical inclusions, seed 4)",
        "config": {"workload": f"{R} RHS, 3D 7-point diffusion {n}^3 (N={a.n}, nnz={a.nnz}) "
                               f"with 64 inclusions k=1e6, FGMRES({args.m}), rtol {args.rtol}",
                   "rhs": R, "smoother": SMOOTHER_NAMES[args.smoother],
                   "threads": args.threads or 16},
        "speedup_vs_fp64": res["fp64"][0] / t_lb,
        "speedup_vs_fp32": res["fp32"][0] / t_lb,
        "speedup_vs_int_one_at_a_time": res["int"][0] / t_lb,
        "fallbacks": len(redo), "learned_saturated_levels": int(ls.saturated),
        "learned_omega": [float(v) for v in ls.omega], "learned_bits": [int(v) for v in ls.bits],
        **{k: summary(k) for k in res},
    }
    print(json.dumps(rec), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128, help="grid side of the 3D 7-point system")
    ap.add_argument("--smoother", type=int, default=1, choices=(0, 1, 2),
                    help="0: SOR (seq.exe ordering), 1: BlockSOR(40) (multi.exe), "
                         "2: JacobiSOR (1200-row blocks)")
    ap.add_argument("--m", type=int, default=30)
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-fp", action="store_true", help="skip the FP64/FP32 solver runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-timeout", type=float, default=400.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel (MIB MiB per launch) for counter calibration")
    ap.add_argument("--workload", default="laplace3d",
                    choices=("laplace3d", "poisson2d", "batch", "learned", "scale"),
                    help="laplace3d: the 3D 7-point solve (default); poisson2d: the 2D "
                         "5-point 256^2 case with SOR (the reference's CPU-sized case); "
                         "batch: uneven 2D "
                         "convection-diffusion systems solved at once; learned: learned "
                         "per-level parameters on the 27-point diffusion system (--n 96); "
                         "scale: learned + batched right-hand sides on the 160^3 "
                         "inclusions system against the plain FP64 / FP32 solvers")
    ap.add_argument("--rhs", type=int, default=16, help="scale: right-hand sides")
    ap.add_argument("--systems", type=int, default=256, help="batch: number of systems")
    ap.add_argument("--threads", type=int, default=16, help="batch: host threads")
    args = ap.parse_args()
    if args.workload == "batch" and args.m == 30 and "--m" not in sys.argv:
        args.m = 20

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(requested --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import sc23ad as S

    if S.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, 9)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "learned":
        if "--n" not in sys.argv:
            args.n = 96
        run_learned(args)
        return
    if args.workload == "scale":
        run_scale(args)
        return

    n = args.n
    if args.workload == "poisson2d":
        # the reference's own CPU-sized case: 2D 5-point 256^2, x* seed 0, SOR
        n = args.n if "--n" in sys.argv else 256
        if "--smoother" not in sys.argv:
            args.smoother = 0
        gen, seed, tag = S.GEN_POISSON2D, 0, "poisson2d"
        desc = f"2D 5-point Poisson {n}^2"
    else:
        gen, seed, tag = S.GEN_LAPLACE3D, 1, "laplace3d"
        desc = f"3D 7-point Laplacian {n}^3"
    t0 = time.time()
    sy = S.make_system(gen, (n,), seed)
    N, _, nnz = sy.a.info()
    log(f"system: {desc}, N={N}, nnz={nnz} ({time.time() - t0:.1f} s)")
    # unscaled copy for the CPU baseline (the reference scales it itself)
    unscaled = S.CsrMat.generate(gen, n).arrays()

    t0 = time.time()
    prm = S.amg_param(args.smoother)
    precs = [S.INT, S.FP64] + ([] if args.no_fp else [S.FP32])
    hs = S.Hierarchy.build_multi(sy.a, precs, prm)
    hint, h64 = hs[0], hs[1]
    setup_host = time.time() - t0
    depth, level_n, level_nnz = hint.info()
    log(f"setup: depth {depth}, level rows {list(level_n)} ({setup_host:.1f} s)")
    # the same setup with the interpolation and Galerkin products on the GPU
    t0 = time.time()
    hs_dev = S.Hierarchy.build_multi(sy.a, precs, prm, device_setup=True)
    setup_dev = time.time() - t0
    same_setup = all(_hier_sha(h) == _hier_sha(d) for h, d in zip(hs, hs_dev))
    del hs_dev
    log(f"setup on the GPU: {setup_dev:.1f} s, identical hierarchies: {same_setup}")

    amg64 = S.Amg(h64)
    bit = amg64.default_frac_bits(sy.b, sy.nrm_b)
    amgi = S.Amg(hint)
    log(f"fractional bits (Work/main.cpp rule): {bit}")

    b_dev = S.DevArray(sy.b)
    x_dev = S.DevArray(np.zeros(N))
    zeros = np.zeros(N)

    def run(solver, steps, warmup, label):
        for _ in range(warmup):
            x_dev.set(zeros)
            solver.solve_device(b_dev, sy.nrm_b, x_dev)
        S.check(S.lib.sc23ad_dev_sync())
        times, cv = [], None
        for _ in range(steps):
            x_dev.set(zeros)
            S.check(S.lib.sc23ad_dev_sync())
            t = time.perf_counter()
            cv = solver.solve_device(b_dev, sy.nrm_b, x_dev)
            S.check(S.lib.sc23ad_dev_sync())
            times.append(time.perf_counter() - t)
        x = x_dev.get()
        rel = solver.relres(sy.b, sy.nrm_b, x)
        log(f"{label}: {np.mean(times) * 1e3:.1f} ms/solve (device {cv.solve_ms:.1f} ms), "
            f"iter {cv.iter}, converged {cv.is_converged}, explicit relres {rel:.3e}")
        return float(np.mean(times)), cv, rel

    soli = S.Solver(S.INT, sy.a, amgi, m=args.m, eps=args.rtol, bit=bit)
    soli.time_spmv(True)
    t_int, cv_int, rel_int = run(soli, args.steps, args.warmup, "solver 2 (int)")
    launches, spmv_ms = soli.spmv_stats()
    spmv_bytes = spmv_int_bytes(N, N, nnz)
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else None
    trsv_launches, trsv_ms = soli.trsv_stats()
    # the whole V-cycle: 20 back-to-back cycles on device vectors
    vin = np.random.default_rng(5).integers(-(2**29), 2**29, N).astype(np.int32)
    vc_ms, _ = amgi.time_vcycles(vin, 20)
    vc_bytes = amgi.vcycle_bytes()
    vc_achieved = vc_bytes / (vc_ms * 1e-3) / 1e9 if vc_ms > 0 else None
    log(f"V-cycle: {vc_ms:.3f} ms, {vc_bytes / 1e6:.1f} MB algorithmic, "
        f"{vc_achieved:.0f} GB/s")
    rptr, cind, _ = sy.a.arrays()
    nnz_ld = ld_nnz(rptr, cind, args.smoother)
    trsv_bytes = sptrsv_int_bytes(N, nnz_ld)
    trsv_achieved = trsv_bytes / (trsv_ms * 1e-3) / 1e9 if trsv_ms > 0 else None

    extra = {}
    if not args.no_fp:
        sol0 = S.Solver(S.FP64, sy.a, amg64, m=args.m, eps=args.rtol)
        t0s, cv0, rel0 = run(sol0, 1, 1, "solver 0 (fp64)")
        amg32 = S.Amg(hs[2])
        sol1 = S.Solver(S.FP32, sy.a, amg32, m=args.m, eps=args.rtol)
        t1s, cv1, rel1 = run(sol1, 1, 1, "solver 1 (fp32)")
        extra = {"solver0_fp64_s": t0s, "solver0_iter": cv0.iter, "solver0_relres": rel0,
                 "solver1_fp32_s": t1s, "solver1_iter": cv1.iter, "solver1_relres": rel1}

    traffic = read_traffic(os.path.join(HERE, "profiles",
                                        f"spmv_int_level0_{tag}_{n}_pmc.txt"))
    trsv_traffic = read_traffic(os.path.join(HERE, "profiles",
                                             f"sptrsv_int_level0_{tag}_{n}_pmc.txt"))
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args, unscaled, sy.b_unscaled, args.smoother)
        except Exception as e:  # the baseline is reported, never required
            log("cpu baseline error:", e)

    rec = {
        "metric": METRIC,
        "value": t_int,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_int * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32/int64 (FP64 outer refinement)",
        "data": f"synthetic (generated {desc}, x* ~ U[-1,1] seed {seed})",
        "config": {"workload": f"int-AMG-FGMRES({args.m}) solve, {desc} "
                               f"(N={N}, nnz={nnz}), "
                               f"{SMOOTHER_NAMES[args.smoother]} smoother, "
                               f"rtol {args.rtol}",
                   "solver": 2, "restart_m": args.m, "frac_bits": bit,
                   "smoother": SMOOTHER_NAMES[args.smoother],
                   "amg_levels": [int(v) for v in level_n]},
        "iterations": cv_int.iter,
        "converged": bool(cv_int.is_converged),
        "explicit_relres": rel_int,
        "roofline": {"bound": "hbm",
                     "kernel": "fixed-point SpMV of A0 (Arnoldi step; SELL-64 with 16-bit "
                               "column offsets, SELL-64 or CSR, per-matrix autotuned)",
                     "timing": "hipEvents bound to each launch's dispatch (hipExtLaunchKernel)",
                     "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBPS) if achieved else None,
                     "traffic": traffic, "algorithmic_bytes": spmv_bytes,
                     "launches_timed": launches, "mean_ms": spmv_ms},
        "roofline_sptrsv": {"bound": "hbm",
                            "kernel": "integer LD SpTRSV, level 0 (smoother; row-parallel, "
                                      "one workgroup per BlockSOR chain)",
                            "timing": "hipEvents bound to each launch's dispatch "
                                      "(hipExtLaunchKernel)",
                            "achieved": trsv_achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                            "frac": (trsv_achieved / HBM_PEAK_GBPS) if trsv_achieved else None,
                            "traffic": trsv_traffic, "algorithmic_bytes": trsv_bytes,
                            "nnz_ld": nnz_ld, "launches_timed": trsv_launches,
                            "mean_ms": trsv_ms},
        "roofline_vcycle": {"bound": "hbm", "kernel": "one integer V-cycle (all its kernels)",
                            "timing": "hipEvents around 20 back-to-back V-cycles",
                            "achieved": vc_achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                            "frac": (vc_achieved / HBM_PEAK_GBPS) if vc_achieved else None,
                            "algorithmic_bytes": vc_bytes, "mean_ms": vc_ms},
        "cpu_baseline": cpu,
        "setup_s": {"host": setup_host, "gpu_interp_galerkin": setup_dev,
                    "identical": bool(same_setup), "precisions": len(precs)},
    }
    rec.update(extra)
    if cpu and cpu.get("value"):
        rec["speedup_vs_cpu"] = cpu["value"] / t_int
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
