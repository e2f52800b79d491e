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
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "sc23-ad_amd"))

METRIC = ("solve time to 1e-8 rel. residual; int-SpMV HBM GB/s; "
          "learned/batched vs plain")
HBM_PEAK_GBPS = 8000.0
T_START = time.time()
# The batched workloads run one stream per host thread; with HIP's default of
# 4 hardware queues per process only 4 of them run at once.  Set before the
# library touches the GPU (the value is read when HIP initialises).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")


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
    print(json.dumps({"copy_kernel": "k_copy_16", "bytes_read_per_launch": nbytes,
                      "bytes_written_per_launch": nbytes, "launches": reps + 1}), flush=True)


def copy_ceiling(gib: int = 2, reps: int = 10):
    """HBM copy ceiling of this GPU: the library's streaming copy (16-byte
    loads and stores) of `gib` GiB, `reps` times back to back; GB/s counts
    the bytes read and written."""
    import sc23ad as S
    nbytes = gib << 30
    src = S.DevArray(nbytes=nbytes)
    dst = S.DevArray(nbytes=nbytes)
    S.check(S.lib.sc23ad_dev_copy(dst.p, src.p, nbytes))
    S.check(S.lib.sc23ad_dev_sync())
    t = time.perf_counter()
    for _ in range(reps):
        S.check(S.lib.sc23ad_dev_copy(dst.p, src.p, nbytes))
    S.check(S.lib.sc23ad_dev_sync())
    dt = (time.perf_counter() - t) / reps
    return {"kernel": "k_copy_16 (16-byte loads and stores, one word per thread)",
            "bytes_per_copy": 2 * nbytes, "ms_per_copy": dt * 1e3,
            "gbps": 2 * nbytes / dt / 1e9}


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
    # every core of this process's CPU share (OMP_NUM_THREADS where the host
    # sets it, else the cores this process may run on)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        len(os.sched_getaffinity(0)) or os.cpu_count() or 1
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
        try:
            import psutil
            phys = psutil.cpu_count(logical=False)
        except Exception:  # the count is reported, never required
            phys = None
        return {"value": float(tm.group(1)), "unit": "s",
                "cores": threads if omp else 1, "kind": "reference",
                "host_physical_cores": phys, "host_logical_cpus": os.cpu_count(),
                "sample": (f"full Restarted::Solve of the same system "
                           f"({'multi.exe' if omp else 'seq.exe'}, {SMOOTHER_NAMES[smoother]},"
                           f" solver 2, m={args.m}, eps={args.rtol}); "
                           f"iterations {it.group(1) if it else '?'}, explicit relres "
                           f"{rs.group(1) if rs else '?'}; process wall incl. setup {wall:.1f} s"),
                "iterations": int(it.group(1)) if it else None}
    return None


def _status_counts(cvs):
    names = {0: "converged", 3: "overflow", 4: "stagnation", 6: "not_converged"}
    out = {}
    for c in cvs:
        k = names.get(int(c.status), f"status_{int(c.status)}")
        out[k] = out.get(k, 0) + 1
    return out


def _conv_stats(cvs):
    it = [int(c.iter) for c in cvs if c.is_converged]
    return {"converged": len(it), "solves": len(cvs), "statuses": _status_counts(cvs),
            "iterations_mean": float(np.mean(it)) if it else None,
            "iterations_max": int(np.max(it)) if it else None}


def run_batch(args):
    """Batched uneven systems: `--systems` 2D convection-diffusion systems
    (5-point first-order upwind; side uniform in [32, 256], Peclet in [1, 100],
    flow angle in [0, 2 pi) drawn from seed 3), 4 right-hand sides each
    (b = A x*, x* ~ U[-1,1] from seed 1000 (i + 1) + k, row-scaled),
    int-AMG-FGMRES(20) to rtol with F by the driver's rule.

    Batched: one block solver per system (its 4 right-hand sides share every
    launch of the step), all systems at once (one stream each, `--threads`
    host threads).  Plain: the same integer solves one right-hand side at a
    time, and the FP64 solver 0 one at a time.  Rates count converged solves
    only (systems/s = converged solves / 4 / seconds); about a tenth of this
    family's integer solves stop with OVERFLOW, as the reference's do (its
    process dies at gmres.hpp:244 on the same inputs)."""
    import sc23ad as S
    K, R = args.systems, 4
    ns, pes, angs = S.convdiff_draws(K, 3)
    t0 = time.time()
    prm = S.amg_param(args.smoother)
    syss, rows = [], 0
    for i in range(K):
        a = S.CsrMat.generate(S.GEN_CONVDIFF2D, int(ns[i]), float(pes[i]), float(angs[i]))
        rptr, cind, val = a.arrays()
        rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
        bs = [a.matvec(S.xstar(a.n, 1000 * (i + 1) + k)) / rmax for k in range(R)]
        a.lmax1_scaling(np.ones(a.n))
        hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], prm)
        am, ad = S.Amg(hi), S.Amg(hd)
        nrms = [float(np.sqrt(np.add.accumulate(b * b)[-1])) for b in bs]
        bits = [ad.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
        syss.append(dict(a=a, am=am, ad=ad, bs=bs, nrms=nrms, bits=bits))
        rows += a.n
        if (i + 1) % 64 == 0:
            log(f"batch setup: {i + 1}/{K} systems ({time.time() - t0:.1f} s)")
    setup_s = time.time() - t0
    singles = [[S.Solver(S.INT, s["a"], s["am"], m=args.m, eps=args.rtol, bit=bt)
                for bt in s["bits"]] for s in syss]
    fp64 = [S.Solver(S.FP64, s["a"], s["ad"], m=args.m, eps=args.rtol) for s in syss]
    blocks = [S.BlockSolver(s["a"], s["am"], s["bits"], m=args.m, eps=args.rtol) for s in syss]
    log(f"batch: {K} systems, {rows} rows, {K * R} right-hand sides; setup {setup_s:.1f} s")
    # warm-up (first launches of every kernel)
    for q in range(min(4, K)):
        singles[q][0].solve(syss[q]["bs"][0], syss[q]["nrms"][0])
        fp64[q].solve(syss[q]["bs"][0], syss[q]["nrms"][0])
        blocks[q].solve(syss[q]["bs"], syss[q]["nrms"])
    t = time.perf_counter()
    xs_seq, cv_seq = [], []
    for s, sol in zip(syss, singles):
        for k in range(R):
            x, cv = sol[k].solve(s["bs"][k], s["nrms"][k])
            xs_seq.append(x)
            cv_seq.append(cv)
    t_seq = time.perf_counter() - t
    log(f"integer, one right-hand side at a time: {t_seq:.2f} s")
    t = time.perf_counter()
    xs_blk, cv_blk = S.block_solve_batch(blocks, [s["bs"] for s in syss],
                                         [s["nrms"] for s in syss], threads=args.threads)
    t_bat = time.perf_counter() - t
    log(f"integer, batched (block solver per system, all systems at once): {t_bat:.2f} s")
    cv_bat = [c for cs in cv_blk for c in cs]
    xb = [x for xs in xs_blk for x in xs]
    same = all(np.array_equal(p.view(np.uint64), q.view(np.uint64)) for p, q in zip(xs_seq, xb))
    same = same and [c.status for c in cv_seq] == [c.status for c in cv_bat]
    t = time.perf_counter()
    cv_fp = []
    for s, sol in zip(syss, fp64):
        for k in range(R):
            cv_fp.append(sol.solve(s["bs"][k], s["nrms"][k])[1])
    t_fp = time.perf_counter() - t
    log(f"FP64 (solver 0), one at a time: {t_fp:.2f} s")
    st_bat, st_seq, st_fp = _conv_stats(cv_bat), _conv_stats(cv_seq), _conv_stats(cv_fp)
    rate = lambda st, tt: st["converged"] / R / tt
    return {
        "metric": "batched uneven systems: systems per second over converged solves",
        "value": rate(st_bat, t_bat), "unit": "systems/s", "higher_is_better": True,
        "config": {"workload": f"{K} 2D convection-diffusion systems (side 32..256, Peclet "
                               f"1..100), {R} RHS each, int-AMG-FGMRES({args.m}), "
                               f"rtol {args.rtol}, BlockSOR(40)",
                   "systems": K, "rhs_per_system": R, "rows_total": rows,
                   "threads": args.threads},
        "batched_s": t_bat, "int_one_at_a_time_s": t_seq, "fp64_one_at_a_time_s": t_fp,
        "batched_systems_per_s": rate(st_bat, t_bat),
        "int_one_at_a_time_systems_per_s": rate(st_seq, t_seq),
        "fp64_one_at_a_time_systems_per_s": rate(st_fp, t_fp),
        "speedup_vs_int_one_at_a_time": t_seq / t_bat,
        "speedup_vs_fp64_one_at_a_time": t_fp / t_bat,
        "batched": st_bat, "int_one_at_a_time": st_seq, "fp64_one_at_a_time": st_fp,
        "identical_to_one_at_a_time": bool(same), "setup_s": setup_s,
    }


def run_learned(args):
    """Learned per-level parameters against the reference's defaults on the
    3D 27-point variable-coefficient diffusion system (side --n, default 96:
    coefficient exp(2 g) on 8^3-cell blocks, anisotropy 1:1:1e-3 in z, seed 2),
    int-AMG-FGMRES(30) to rtol: seconds per solve for both, with b and x
    resident in HBM, and the iterations.  The learned solve stops after
    STALL_STEPS steps without progress and then the defaults solve it; its
    time then includes both."""
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
    log(f"learned omega {np.round(ls.omega, 3).tolist()} vector bits {ls.vbits.tolist()} "
        f"saturated {ls.saturated}, levels outside the training range {ls.abstained}")
    b_dev = S.DevArray(sy.b)
    x_dev = S.DevArray(np.zeros(sy.a.n))
    zeros = np.zeros(sy.a.n)

    def make(hint, hfp, stall):
        bit = S.Amg(hfp).default_frac_bits(sy.b, sy.nrm_b)
        sol = S.Solver(S.INT, sy.a, S.Amg(hint), m=args.m, eps=args.rtol, bit=bit)
        sol.set_stall(stall)
        return sol

    def timed(sol):
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
        return float(np.mean(ts)), cv, rel

    sol_d = make(bi, bd, 0)
    t_def, cv_def, rel_def = timed(sol_d)
    log(f"default: {t_def * 1e3:.1f} ms, {cv_def.iter} iterations, relres {rel_def:.3e}")
    sol_l = make(ls.hier_int, ls.hier_fp64,
                 LR.STALL_STEPS if args.stall is None else args.stall)
    t_l, cv_l, rel_l = timed(sol_l)
    fallback = not (cv_l.is_converged and cv_l.status == S.OK and rel_l <= args.rtol)
    it_l = cv_l.iter
    if fallback:  # the learned attempt's time plus the default solve
        t_l, it_l, rel_l = t_l + t_def, cv_def.iter, rel_def
    log(f"learned: {t_l * 1e3:.1f} ms, {it_l} iterations, relres {rel_l:.3e}, "
        f"fallback {fallback}")
    return {
        "metric": "learned per-level omega / vector bits vs reference defaults: solve time "
                  "to rtol",
        "value": t_l, "unit": "s", "higher_is_better": False,
        "config": {"workload": f"int-AMG-FGMRES({args.m}), 3D 27-point diffusion {n}^3 "
                               f"(N={sy.a.n}), anisotropy 1:1:1e-3, rtol {args.rtol}"},
        "default_s": t_def, "default_iterations": cv_def.iter, "default_relres": rel_def,
        "learned_s": t_l, "learned_iterations": it_l, "learned_relres": rel_l,
        "learned_omega": [float(v) for v in ls.omega],
        "learned_vbits": [int(v) for v in ls.vbits], "saturated": int(ls.saturated),
        "abstained_levels": int(ls.abstained),
        "fallback_used": bool(fallback), "learned_vs_default": t_def / t_l,
    }


def run_vcycles(args):
    """k-RHS block V-cycle against k single V-cycles on device vectors
    (INTAMG.hpp:269-293 on several right-hand sides per launch), at 128^3
    (3D 7-point Laplacian) and on the 160^3 inclusions system."""
    import sc23ad as S
    out = {}
    cases = [("laplace3d_128", S.GEN_LAPLACE3D, (128,)),
             ("inclusions3d_160", S.GEN_INCLUSIONS3D, (160, 64, 4.0, 12.0, 1e6, 4))]
    for tag, gen, params in cases:
        a = S.CsrMat.generate(gen, *params)
        a.lmax1_scaling(np.ones(a.n))
        hi = S.Hierarchy(a, S.INT, S.amg_param(args.smoother))
        amg = S.Amg(hi)
        rng = np.random.default_rng(7)
        v = rng.integers(-(2**29), 2**29, (16, a.n)).astype(np.int32)
        single_ms, _ = amg.time_vcycles(v[0], 20)
        r = {"n": a.n, "native": S.block_native(amg), "single_ms": single_ms}
        for k in (4, 8, 16):
            ms = S.time_vcycles_block(amg, v[:k], 20)
            r[f"block{k}_ms"] = ms
            r[f"block{k}_per_rhs_vs_single"] = single_ms * k / ms
        log(f"V-cycles {tag}: single {single_ms:.3f} ms; " +
            ", ".join(f"k={k} {r[f'block{k}_ms']:.3f} ms ({r[f'block{k}_per_rhs_vs_single']:.2f}x)"
                      for k in (4, 8, 16)))
        out[tag] = r
    return {"metric": "block V-cycle throughput over single V-cycles, 16 right-hand sides "
                      "(3D 7-point Laplacian 128^3)",
            "value": out["laplace3d_128"]["block16_per_rhs_vs_single"], "unit": "x",
            "higher_is_better": True, "config": {"workload": "vcycles"}, "cases": out}


def run_scale(args):
    """Learned + batched at scale: the 3D 7-point diffusion system on a 160^3
    grid (4,096,000 rows; k = 1 with 64 spherical inclusions of radius
    U[4, 12] cells and k = 1e6, seed 4), --rhs right-hand sides b = A x*
    (x* ~ U[-1,1], seeds 400 + k), row-scaled as the reference driver does;
    BlockSOR(40), FGMRES(30), rtol.

    Plain GPU forms: FP64 (solver 0) and FP32-mixed (solver 1) one
    right-hand side after another, and the integer solver with the default
    parameters one after another.  Batched: the integer solver on all
    right-hand sides in one block solver (every V-cycle, SpMV and
    Gram-Schmidt pass of a step shared).  Learned + batched: the same block
    solver on the hierarchy with the network's per-level weights and vector
    bits; a right-hand side whose learned solve does not converge cleanly is
    solved again with the defaults (counted in its time).  b, x in HBM for
    every method.  A right-hand side counts as solved only if its explicit
    FP64 residual is at most rtol (the same accuracy for every method)."""
    import sc23ad as S
    from sc23ad import learned as LR
    n = args.n if "--n" in sys.argv else 160
    R = args.rhs
    t0 = time.time()
    a = S.CsrMat.generate(S.GEN_INCLUSIONS3D, n, 64, 4.0, 12.0, 1e6, 4)
    rptr, cind, val = a.arrays()
    rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
    bs = [a.matvec(S.xstar(a.n, 400 + k)) / rmax for k in range(R)]
    a.lmax1_scaling(np.ones(a.n))
    nrms = [float(np.sqrt(np.add.accumulate(b * b)[-1])) for b in bs]
    prm = S.amg_param(args.smoother)
    hi, hd, hf = S.Hierarchy.build_multi(a, [S.INT, S.FP64, S.FP32], prm)
    log(f"inclusions {n}^3: N={a.n}, nnz={a.nnz}, depth {hi.info()[0]} "
        f"({time.time() - t0:.1f} s)")
    ls = LR.learned_setup(hi, hd, LR.load_mlp())
    log(f"learned omega {np.round(ls.omega, 3).tolist()} vector bits {ls.vbits.tolist()} "
        f"saturated {ls.saturated}, levels outside the training range {ls.abstained} "
        f"({time.time() - t0:.1f} s)")
    amg_i, amg_d, amg_f = S.Amg(hi), S.Amg(hd), S.Amg(hf)
    amg_l, amg_ld = S.Amg(ls.hier_int), S.Amg(ls.hier_fp64)
    bits_d = [amg_d.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
    bits_l = [amg_ld.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
    sol_d = S.Solver(S.FP64, a, amg_d, m=args.m, eps=args.rtol)
    sol_f = S.Solver(S.FP32, a, amg_f, m=args.m, eps=args.rtol)
    blk_i = S.BlockSolver(a, amg_i, bits_d, m=args.m, eps=args.rtol)
    blk_l = S.BlockSolver(a, amg_l, bits_l, m=args.m, eps=args.rtol)
    # the integer solves of this system stay on plateaus for hundreds of
    # steps with the defaults too: no stall stop unless asked for
    blk_l.set_stall(0 if args.stall is None else args.stall)
    setup_s = time.time() - t0
    log(f"setup done ({setup_s:.1f} s)")
    B = np.ascontiguousarray(np.stack(bs))
    b_dev = S.DevArray(B.reshape(-1))
    x_dev = S.DevArray(np.zeros(R * a.n))
    zeros = np.zeros(R * a.n)
    nb = np.asarray(nrms)
    sol_d.solve(bs[0], nrms[0])  # warm-up
    blk_i.solve(bs[:1] * R, nrms[:1] * R)

    def explicit_check(xs):
        return [sol_d.relres(b, nbv, x) for b, nbv, x in zip(bs, nrms, xs)]

    def one_by_one(sol, label):
        x_dev.set(zeros)  # every right-hand side's solve starts from x = 0 in its own view
        S.check(S.lib.sc23ad_dev_sync())
        t = time.perf_counter()
        xs, cvs = [], []
        for k in range(R):
            cv = sol.solve_device(b_dev.view(k * a.n, a.n), nrms[k], x_dev.view(k * a.n, a.n))
            cvs.append(cv)
        S.check(S.lib.sc23ad_dev_sync())
        tt = time.perf_counter() - t
        xs = list(x_dev.get().reshape(R, a.n))
        log(f"{label}: {tt:.3f} s, iterations {[c.iter for c in cvs]}")
        return tt, cvs, explicit_check(xs)

    def block(blk, label):
        x_dev.set(zeros)
        S.check(S.lib.sc23ad_dev_sync())
        t = time.perf_counter()
        cvs = blk.solve_device(b_dev, nb, x_dev)
        S.check(S.lib.sc23ad_dev_sync())
        tt = time.perf_counter() - t
        xs = list(x_dev.get().reshape(R, a.n))
        log(f"{label}: {tt:.3f} s, iterations {[c.iter for c in cvs]}, "
            f"statuses {[c.status for c in cvs]}")
        return tt, cvs, xs

    res = {}
    res["fp64"] = one_by_one(sol_d, "plain FP64, one at a time")
    res["fp32"] = one_by_one(sol_f, "plain FP32-mixed, one at a time")
    if not args.no_int_single:
        sols_i = [S.Solver(S.INT, a, amg_i, m=args.m, eps=args.rtol, bit=bt) for bt in bits_d]
        x_dev.set(zeros)
        S.check(S.lib.sc23ad_dev_sync())
        t = time.perf_counter()
        cvs = []
        for k in range(R):
            cvs.append(sols_i[k].solve_device(b_dev.view(k * a.n, a.n), nrms[k],
                                              x_dev.view(k * a.n, a.n)))
        S.check(S.lib.sc23ad_dev_sync())
        tt = time.perf_counter() - t
        res["int"] = (tt, cvs, explicit_check(list(x_dev.get().reshape(R, a.n))))
        log(f"integer defaults, one at a time: {tt:.3f} s")
        del sols_i
    tt, cvs, xs = block(blk_i, "integer defaults, batched")
    res["int_batched"] = (tt, cvs, explicit_check(xs))
    tt, cvs, xs = block(blk_l, "learned + batched")
    rel = explicit_check(xs)
    # Fallback at scale: the right-hand sides whose learned solve overflowed
    # or stalled are solved again with the defaults.  A learned solve that
    # ran out of restarts, or converged (the reference's implicit test) with
    # an explicit residual above rtol, is kept: the defaults' own batched
    # solve above shows the same outcomes on this system, so a second solve
    # would only repeat them (solved_at_rtol reports the accuracy).
    redo = [k for k in range(R) if cvs[k].status in (S.OVERFLOW, S.STAGNATION)]
    t_redo = 0.0
    if redo:
        blk_f = S.BlockSolver(a, amg_i, [bits_d[k] for k in redo], m=args.m, eps=args.rtol)
        t = time.perf_counter()
        xr, cr = blk_f.solve([bs[k] for k in redo], [nrms[k] for k in redo])
        t_redo = time.perf_counter() - t
        for q, k in enumerate(redo):
            xs[k], cvs[k] = xr[q], cr[q]
        rel = explicit_check(xs)
    res["learned_batched"] = (tt + t_redo, cvs, rel)
    log(f"learned + batched: {tt + t_redo:.3f} s (fallbacks {len(redo)})")

    def summary(key):
        t, cvs, rel = res[key]
        solved = sum(1 for c, r in zip(cvs, rel) if c.is_converged and r <= args.rtol)
        return {"s": t, "iterations": [int(c.iter) for c in cvs],
                "statuses": _status_counts(cvs), "solved_at_rtol": solved,
                "explicit_relres_max": float(max(rel))}
    out = {k: summary(k) for k in res}
    t_lb = res["learned_batched"][0]
    return {
        "metric": "learned + batched vs plain FP64 / FP32-mixed / integer: time to solution "
                  f"for {R} right-hand sides to rtol {args.rtol}",
        "value": t_lb, "unit": "s", "higher_is_better": False,
        "config": {"workload": f"{R} RHS, 3D 7-point diffusion {n}^3 (N={a.n}, nnz={a.nnz}) "
                               f"with 64 inclusions k=1e6, FGMRES({args.m}), rtol {args.rtol}",
                   "rhs": R, "smoother": SMOOTHER_NAMES[args.smoother]},
        "speedup_vs_fp64": res["fp64"][0] / t_lb, "speedup_vs_fp32": res["fp32"][0] / t_lb,
        "speedup_vs_int_one_at_a_time": (res["int"][0] / t_lb) if "int" in res else None,
        "speedup_batched_vs_int_one_at_a_time":
            (res["int"][0] / res["int_batched"][0]) if "int" in res else None,
        "all_solved_at_rtol": {k: v["solved_at_rtol"] == R for k, v in out.items()},
        "fallbacks": len(redo), "learned_saturated_levels": int(ls.saturated),
        "learned_abstained_levels": int(ls.abstained),
        "learned_omega": [float(v) for v in ls.omega],
        "learned_vbits": [int(v) for v in ls.vbits], "setup_s": setup_s,
        **out,
    }


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
                    choices=("laplace3d", "poisson2d", "batch", "learned", "scale", "vcycles"),
                    help="laplace3d: the 3D 7-point solve (default); poisson2d: the 2D "
                         "5-point 256^2 case with SOR (the reference's CPU-sized case); "
                         "batch: uneven 2D "
                         "convection-diffusion systems solved at once; learned: learned "
                         "per-level parameters on the 27-point diffusion system (--n 96); "
                         "scale: learned + batched right-hand sides on the 160^3 "
                         "inclusions system against the plain FP64 / FP32 solvers; "
                         "vcycles: block V-cycles of 4-16 right-hand sides against single ones")
    ap.add_argument("--rhs", type=int, default=16, help="scale: right-hand sides")
    ap.add_argument("--systems", type=int, default=256, help="batch: number of systems")
    ap.add_argument("--threads", type=int, default=16, help="batch: host threads")
    ap.add_argument("--stall", type=int, default=None,
                    help="learned / scale: Arnoldi steps without progress after which a "
                         "learned solve falls back to the defaults (0: off; default: "
                         "the learned module's STALL_STEPS)")
    ap.add_argument("--no-int-single", action="store_true",
                    help="scale: skip the integer solves one right-hand side at a time")
    ap.add_argument("--extras", default="batch,learned,scale",
                    help="laplace3d: the other workloads run after it and reported as "
                         "extra keys of the same line (comma list; '' for none)")
    ap.add_argument("--extras-budget", type=float, default=500.0,
                    help="seconds since start by which the extras should have ended: "
                         "an extra starts only if its expected time fits")
    ap.add_argument("--extras-deadline", type=float, default=560.0,
                    help="seconds since start at which the line is printed with the "
                         "extras done so far and the process ends")
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
    if args.workload in ("batch", "learned", "scale", "vcycles"):
        if args.workload == "learned" and "--n" not in sys.argv:
            args.n = 96
        fn = {"batch": run_batch, "learned": run_learned, "scale": run_scale,
              "vcycles": run_vcycles}[args.workload]
        t = time.perf_counter()
        rec = fn(args)
        wall = time.perf_counter() - t
        rec = {"metric": rec.pop("metric"), "value": rec.pop("value"), "unit": rec.pop("unit"),
               "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": wall * 1e3,
               "higher_is_better": rec.pop("higher_is_better"), "scaling": "weak",
               "vs_baseline": None, "dtype": "int32/int64 (FP64 outer refinement)",
               "data": "synthetic (generated)", **rec}
        print(json.dumps(rec), flush=True)
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

    cc = copy_ceiling()
    log(f"copy ceiling: {cc['gbps']:.0f} GB/s")
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
                     "launches_timed": launches, "mean_ms": spmv_ms,
                     # the measured HBM traffic per launch over the mean
                     # duration, against this run's copy ceiling
                     "traffic_gbps": (traffic / (spmv_ms * 1e-3) / 1e9)
                                     if traffic and spmv_ms else None,
                     "traffic_frac_of_copy_ceiling":
                         (traffic / (spmv_ms * 1e-3) / 1e9 / cc["gbps"])
                         if traffic and spmv_ms and cc.get("gbps") else None},
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
        "copy_ceiling": cc,
        "cpu_baseline": cpu,
        "setup_s": {"host": setup_host, "gpu_interp_galerkin": setup_dev,
                    "identical": bool(same_setup), "precisions": len(precs)},
    }
    rec.update(extra)
    if cpu and cpu.get("value"):
        rec["speedup_vs_cpu"] = cpu["value"] / t_int
    if args.workload == "laplace3d" and args.n == 128:
        run_extras(args, rec)
    _emit(rec)


_EMIT_LOCK = threading.Lock()
_EMITTED = [False]


def _emit(rec):
    """Print the one JSON line (once, whichever of the main thread and the
    deadline watchdog gets here first)."""
    with _EMIT_LOCK:
        if _EMITTED[0]:
            return
        _EMITTED[0] = True
        print(json.dumps(rec), flush=True)


# Seconds an extra workload takes on one MI355X (setup included), with room:
# an extra starts only if it ends within the budget by this estimate.
EXTRA_EST_S = {"batch": 150.0, "learned": 100.0, "scale": 200.0}


def run_extras(args, rec):
    """The metric's third part (learned / batched vs plain) on its own
    workloads, after the main line's measurement: each as an extra key of the
    same JSON line, each started only if it should end within the time
    budget.  A watchdog prints the line with what is done by the hard
    deadline, should an extra run long, and ends the process."""
    import copy
    wanted = [w for w in args.extras.split(",") if w]

    def watchdog():
        with _EMIT_LOCK:
            done = _EMITTED[0]
        if not done:
            snap = {}
            for _ in range(10):  # the main thread may be adding a key
                try:
                    snap = dict(rec)
                    break
                except RuntimeError:
                    time.sleep(0.01)
            snap["extras_deadline"] = f"line printed at the {args.extras_deadline:.0f} s deadline"
            log(f"extras: deadline of {args.extras_deadline:.0f} s reached")
            _emit(snap)
            os._exit(0)

    left = args.extras_deadline - (time.time() - T_START)
    timer = threading.Timer(max(left, 1.0), watchdog)
    timer.daemon = True
    timer.start()
    for w in wanted:
        elapsed = time.time() - T_START
        est = EXTRA_EST_S.get(w, 120.0)
        if elapsed + est > args.extras_budget:
            rec[f"extra_{w}"] = {"skipped": f"time budget ({elapsed:.0f} s elapsed + "
                                            f"{est:.0f} s expected > {args.extras_budget:.0f} s)"}
            log(f"extra {w}: skipped, {elapsed:.0f} s elapsed")
            continue
        a = copy.copy(args)
        a.steps, a.warmup = 3, 1
        try:
            if w == "batch":
                a.m = 20
                r = run_batch(a)
                rec["batched_uneven"] = r
                rec["batched_systems_per_s"] = r["batched_systems_per_s"]
            elif w == "learned":
                a.n = 96
                r = run_learned(a)
                rec["learned"] = r
                rec["learned_vs_default"] = r["learned_vs_default"]
            elif w == "scale":
                a.n = 160
                a.no_int_single = True
                r = run_scale(a)
                rec["learned_batched_at_scale"] = r
        except Exception as e:  # reported, never fatal to the main line
            rec[f"extra_{w}"] = {"error": repr(e)}
            log(f"extra {w} failed: {e!r}")
        log(f"extra {w} done at {time.time() - T_START:.0f} s")


if __name__ == "__main__":
    main()
