"""HIP solve path against the oracle, through the C ABI (MI355X).

Bar: bit-exact for every integer vector, and for the FP64/FP32 solvers too,
since the device library's FP reductions follow a fixed documented order that
the oracle restates (1024 x 256 blocked partial sums, halving trees).  In the
device library's sequential-reduction mode the GPU reproduces the
reference's own results (tests/golden/reference.json) bit for bit.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as O
import sc23ad as S
from helpers import export, sha, system

pytestmark = pytest.mark.gpu

CASES = ["poisson2d_32", "poisson2d_48", "poisson2d_64", "laplace3d_16",
         "convdiff2d_60", "diffusion27_16", "inclusions3d_20"]


@pytest.fixture(scope="module")
def built(golden):
    """case -> (case record, system, {(prec, smoother): (Amg, OracleHier)})."""
    out = {}
    for c in golden["cases"]:
        if c["name"] not in CASES:
            continue
        print(f"[setup] {c['name']}", flush=True)
        sy = system(c)
        hs = {}
        for smoother in (0, 1, 2, 3):
            precs = (S.FP64, S.FP32) if smoother == 3 else (S.INT, S.FP64, S.FP32)
            multi = S.Hierarchy.build_multi(sy.a, list(precs), S.amg_param(smoother))
            for prec, h in zip(precs, multi):
                path = export(h)
                hs[(prec, smoother)] = (S.Amg(h), O.OracleHier(path))
                os.unlink(path)
        out[c["name"]] = (c, sy, hs)
    return out


def _rand(prec, n, rng):
    if prec == S.INT:
        return rng.integers(-(2**29), 2**29, n).astype(np.int32)
    return rng.standard_normal(n).astype(S.DT[prec])


def _same(a, b):
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("prec", [S.INT, S.FP64, S.FP32], ids=["int", "fp64", "fp32"])
def test_level_ops(built, name, prec):
    c, sy, hs = built[name]
    rng = np.random.default_rng(17)
    for smoother in (0, 1, 2, 3):
        if (prec, smoother) not in hs:
            continue  # Jacobi: FP only
        amg, oh = hs[(prec, smoother)]
        depth = amg.depth
        for lv in range(depth):
            n = oh.level_n(lv)
            nxt = oh.level_n(lv + 1)
            x = _rand(prec, n, rng)
            xc = _rand(prec, nxt, rng)
            print(f"[level ops] {name} prec {prec} smoother {smoother} level {lv}", flush=True)
            for op in (S.OP_SPMV_A, S.OP_SPMV_U, S.OP_TRSV_LD, S.OP_SMOOTH0):
                assert _same(amg.level_op(lv, op, x, n), oh.level_op(lv, op, x, n)), op
            assert _same(amg.level_op(lv, S.OP_SPMV_R, x, nxt), oh.level_op(lv, S.OP_SPMV_R, x, nxt))
            assert _same(amg.level_op(lv, S.OP_SPMV_P, xc, n), oh.level_op(lv, S.OP_SPMV_P, xc, n))
            init = _rand(prec, n, rng)
            assert _same(amg.level_op(lv, S.OP_SMOOTH1, x, n, init),
                         oh.level_op(lv, S.OP_SMOOTH1, x, n, init))
        nc = oh.level_n(depth)
        xc = _rand(prec, nc, rng)
        assert _same(amg.level_op(depth, S.OP_COARSE, xc, nc),
                     oh.level_op(depth, S.OP_COARSE, xc, nc))


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("prec", [S.INT, S.FP64, S.FP32], ids=["int", "fp64", "fp32"])
def test_vcycle(built, name, prec):
    c, sy, hs = built[name]
    rng = np.random.default_rng(23)
    for smoother in (0, 1, 2, 3):
        if (prec, smoother) not in hs:
            continue  # Jacobi: FP only
        amg, oh = hs[(prec, smoother)]
        for _ in range(2):  # the second call reuses every work buffer
            v = _rand(prec, sy.a.n, rng)
            assert _same(amg.precondition(v), oh.vcycle(v)), (name, smoother)


@pytest.mark.parametrize("name", ["poisson2d_64", "inclusions3d_20"])
def test_timed_vcycles_compute_the_vcycle(built, name):
    """The V-cycle timing entry point runs the real V-cycle: its output is the
    oracle's, and its per-cycle time and byte count are positive."""
    c, sy, hs = built[name]
    amg, oh = hs[(S.INT, 1)]
    v = _rand(S.INT, sy.a.n, np.random.default_rng(29))
    ms, out = amg.time_vcycles(v, 3)
    assert _same(out, oh.vcycle(v))
    assert ms > 0 and amg.vcycle_bytes() > 0


def test_default_frac_bits(built):
    for name in CASES:
        c, sy, hs = built[name]
        for smoother in (0, 1, 2):
            amg, _ = hs[(S.FP64, smoother)]
            assert amg.default_frac_bits(sy.b, sy.nrm_b) == c["solves"][f"2_{smoother}"]["bit"]


@pytest.mark.parametrize("fmt", ["csr", "sell", "sell16", "sellcol"])
def test_spmv_formats_match_oracle(fmt, monkeypatch):
    """Every SpMV kernel the tuner can pick (CSR with 1-64 lanes, SELL-64,
    length-sorted SELL-64, SELL-64 with 16-bit column offsets, column-ordered
    SELL-64 of the rectangular matrices), forced in
    turn, gives the oracle's A, U, R and P products bit for bit."""
    monkeypatch.setenv("SC23AD_SPMV_FORMAT", fmt)
    sy = S.make_system(S.GEN_INCLUSIONS3D, (24, 8, 2.0, 4.0, 1e4, 3), 5)  # 13,824 rows
    for prec in (S.INT, S.FP64):
        h = S.Hierarchy(sy.a, prec, S.amg_param(S.BLOCK_SOR))
        path = export(h)
        amg, oh = S.Amg(h), O.OracleHier(path)
        os.unlink(path)
        rng = np.random.default_rng(31)
        for lv in range(amg.depth):
            n, nxt = oh.level_n(lv), oh.level_n(lv + 1)
            x, xc = _rand(prec, n, rng), _rand(prec, nxt, rng)
            print(f"[spmv formats] {fmt} prec {prec} level {lv}", flush=True)
            for op in (S.OP_SPMV_A, S.OP_SPMV_U):
                assert _same(amg.level_op(lv, op, x, n), oh.level_op(lv, op, x, n)), (lv, op)
            assert _same(amg.level_op(lv, S.OP_SPMV_R, x, nxt), oh.level_op(lv, S.OP_SPMV_R, x, nxt))
            assert _same(amg.level_op(lv, S.OP_SPMV_P, xc, n), oh.level_op(lv, S.OP_SPMV_P, xc, n))


@pytest.mark.parametrize("knobs", [{"SC23AD_TRI_RP": "0"}, {"SC23AD_TRI_RP": "1"},
                                   {"SC23AD_TRI_RP": "2"}, {"SC23AD_RP_ROWS": "2"},
                                   {"SC23AD_TRI_RP": "0", "SC23AD_EP_COPIES": "1"},
                                   {"SC23AD_RP_E16": "1"}],
                         ids=["entry-parallel", "lane-per-row", "lanes-per-row", "two-rows",
                              "ep-copies", "chunks-of-16"])
def test_ld_solve_variants_match_oracle(knobs, monkeypatch):
    """Every LD-solve kernel layout, forced in turn (integer: entry-parallel
    with and without accumulator copies, row-parallel with one lane per row,
    several lanes per row, two rows per lane; FP64/FP32: row-parallel in
    stored order with one or two rows per lane, several lanes per row with
    chunks of 4 or 8 entries (FP32 also 16), or level-pipelined), gives the oracle's LD
    solves, smoothing steps and V-cycles bit for bit."""
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    for kind, params in ((S.GEN_LAPLACE3D, (24,)), (S.GEN_INCLUSIONS3D, (24, 8, 2.0, 4.0, 1e4, 3))):
        sy = S.make_system(kind, params, 7)
        for smoother in (S.SOR, S.BLOCK_SOR):
            for prec in (S.INT, S.FP64, S.FP32):  # FP: row-parallel in stored order, or not
                h = S.Hierarchy(sy.a, prec, S.amg_param(smoother))
                path = export(h)
                amg, oh = S.Amg(h), O.OracleHier(path)
                os.unlink(path)
                rng = np.random.default_rng(37)
                print(f"[ld variants] {knobs} kind {kind} smoother {smoother} prec {prec}",
                      flush=True)
                for lv in range(amg.depth):
                    n = oh.level_n(lv)
                    x = _rand(prec, n, rng)
                    for op in (S.OP_TRSV_LD, S.OP_SMOOTH0):
                        assert _same(amg.level_op(lv, op, x, n), oh.level_op(lv, op, x, n)), \
                            (lv, op)
                v = _rand(prec, sy.a.n, rng)
                assert _same(amg.precondition(v), oh.vcycle(v))


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("smoother", [0, 1, 2], ids=["sor", "blocksor", "jacobisor"])
def test_int_solve_matches_oracle(built, golden, name, smoother):
    c, sy, hs = built[name]
    amg, oh = hs[(S.INT, smoother)]
    bit = c["solves"][f"2_{smoother}"]["bit"]
    sol = S.Solver(S.INT, sy.a, amg, m=c["m"], eps=golden["eps"], bit=bit)
    sol.trace(8192)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    trace = sol.read_trace(8192)
    xo, ro, to = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=c["m"], eps=golden["eps"],
                          bit=bit, mode=O.RED_GPU, trace_cap=8192)
    assert np.array_equal(trace, to)
    assert (cv.iter, cv.is_converged) == (ro.iters, ro.converged)
    assert _same(x, xo)
    g = c["solves"][f"2_{smoother}"]
    assert abs(cv.iter - g["iters"]) <= 1  # vs the reference's own (sequential) order
    assert sol.relres(sy.b, sy.nrm_b, x) < 1e-9


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("prec", [S.FP64, S.FP32], ids=["fp64", "fp32"])
def test_fp_solve_matches_oracle(built, golden, name, prec):
    c, sy, hs = built[name]
    for smoother in (0, 1, 2, 3):
        amg, oh = hs[(prec, smoother)]
        sol = S.Solver(prec, sy.a, amg, m=c["m"], eps=golden["eps"])
        x, cv = sol.solve(sy.b, sy.nrm_b)
        xo, ro, _ = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=c["m"], eps=golden["eps"],
                             mode=O.RED_GPU)
        assert (cv.iter, cv.is_converged) == (ro.iters, ro.converged)
        assert _same(x, xo)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("solver", [S.INT, S.FP64, S.FP32], ids=["int", "fp64", "fp32"])
def test_sequential_mode_reproduces_reference(built, golden, name, solver):
    """With sequential FP reductions the device solve is the reference's."""
    c, sy, hs = built[name]
    for smoother in (0, 1, 2, 3):
        if (solver, smoother) not in hs:
            continue  # Jacobi: FP only
        g = c["solves"][f"{solver}_{smoother}"]
        amg, _ = hs[(solver, smoother)]
        kw = {"bit": g["bit"]} if solver == S.INT else {}
        sol = S.Solver(solver, sy.a, amg, m=c["m"], eps=golden["eps"], **kw)
        sol.set_sequential_reduction(True)
        if solver == S.INT:
            sol.trace(8192)
        x, cv = sol.solve(sy.b, sy.nrm_b)
        assert (cv.iter, cv.is_converged) == (g["iters"], g["converged"])
        if solver == S.INT:
            assert [format(int(v), "016x") for v in sol.read_trace(8192)] == g["trace"]
        assert sha(x) == g["x_sha"], (name, solver, smoother)


def test_device_pointer_solve_equals_host(built, golden):
    c, sy, hs = built["laplace3d_16"]
    amg, _ = hs[(S.INT, 1)]
    bit = c["solves"]["2_1"]["bit"]
    sol = S.Solver(S.INT, sy.a, amg, m=30, eps=golden["eps"], bit=bit)
    xh, cvh = sol.solve(sy.b, sy.nrm_b)
    db = S.DevArray(sy.b)
    dx = S.DevArray(np.zeros(sy.a.n))
    cvd = sol.solve_device(db, sy.nrm_b, dx)
    assert cvd.iter == cvh.iter and _same(dx.get(), xh)


@pytest.mark.parametrize("bit", [8, 12])
def test_small_bits_status_matches_oracle(built, bit):
    """Few fractional bits: at 8 the restart budget (1000 / m restarts,
    solver/irsolver.hpp:27-43) runs out without convergence and the status
    says so; at 12 the solve converges after more restarts."""
    c, sy, hs = built["poisson2d_64"]
    amg, oh = hs[(S.INT, 0)]
    sol = S.Solver(S.INT, sy.a, amg, m=30, eps=1e-10, bit=bit)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    xo, ro, _ = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=30, eps=1e-10, bit=bit,
                         mode=O.RED_GPU)
    assert cv.status == ro.status and cv.iter == ro.iters and cv.is_converged == ro.converged
    assert _same(x, xo)
    assert cv.is_converged == (bit == 12)


def test_overflowing_bits_flagged(built):
    """30 fractional bits overflow this system's Givens generation (a zero
    norm where lsqrt wraps): the solve stops with the overflow status, as the
    oracle does, instead of trapping on a division by zero."""
    c, sy, hs = built["poisson2d_64"]
    amg, oh = hs[(S.INT, 0)]
    sol = S.Solver(S.INT, sy.a, amg, m=30, eps=1e-10, bit=30)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    xo, ro, _ = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=30, eps=1e-10, bit=30,
                         mode=O.RED_GPU)
    assert ro.status == S.OVERFLOW
    assert (cv.status, cv.iter, cv.overflow_flag) == (ro.status, ro.iters, 1)
    assert _same(x, xo)


# ---- raw integer SpMV -----------------------------------------------------------

def _spmv_ref(rptr, cind, val, bit, x):
    y = np.zeros(len(rptr) - 1, np.int64)
    for i in range(len(rptr) - 1):
        acc = 0
        for p in range(rptr[i], rptr[i + 1]):
            acc += int(val[p]) * int(x[cind[p]])
        acc &= (1 << 64) - 1
        acc = acc - (1 << 64) if acc >> 63 else acc
        v = (acc >> bit) & 0xFFFFFFFF
        y[i] = v - (1 << 32) if v >> 31 else v
    return y.astype(np.int32)


@pytest.mark.parametrize("shape", ["empty_rows", "long_rows", "one_row", "wrap"])
def test_spmv_int_brute_force(shape):
    rng = np.random.default_rng(hash(shape) % 2**32)
    if shape == "one_row":
        lens = np.array([37])
    elif shape == "empty_rows":
        lens = rng.integers(0, 4, 3000) * (rng.random(3000) < 0.6)
    elif shape == "long_rows":
        lens = rng.integers(100, 700, 40)
    else:
        lens = rng.integers(1, 30, 2000)
    n = len(lens)
    m = 5000
    rptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cind = rng.integers(0, m, rptr[-1]).astype(np.int32)
    lim = 2**31 if shape == "wrap" else 2**20
    val = rng.integers(-lim, lim, rptr[-1]).astype(np.int32)
    x = rng.integers(-lim, lim, m).astype(np.int32)
    for bit in (0, 17, 30):
        assert np.array_equal(S.spmv_int(rptr, cind, val, bit, x),
                              _spmv_ref(rptr, cind, val, bit, x))


def test_batch_equals_individual_solves():
    """All solves issued at once (each solver on its own stream, several
    right-hand sides sharing one hierarchy) give the same x as one at a time."""
    draws = [(S.GEN_CONVDIFF2D, (40, 20.0, 0.3)), (S.GEN_POISSON2D, (50,)),
             (S.GEN_LAPLACE3D, (12,))]
    solvers, rhs, nrms, keep, hiers = [], [], [], [], []
    for kind, params in draws:
        a = S.CsrMat.generate(kind, *params)
        rptr, cind, val = a.arrays()
        rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
        bs = [a.matvec(S.xstar(a.n, 7 + k)) for k in range(3)]
        a.lmax1_scaling(bs[0].copy())
        hi = S.Hierarchy(a, S.INT, S.amg_param(S.BLOCK_SOR))
        amg = S.Amg(hi)
        keep.append((a, amg))
        hiers.append(hi)
        for b in bs:
            b = b / rmax
            solvers.append(S.Solver(S.INT, a, amg, m=20, eps=1e-10, bit=28))
            rhs.append(b)
            nrms.append(float(np.sqrt(np.add.accumulate(b * b)[-1])))
    seq = [s.solve(b, nb) for s, b, nb in zip(solvers, rhs, nrms)]
    xs, cvs = S.solve_batch(solvers, rhs, nrms, threads=4)
    for (x1, c1), x2, c2 in zip(seq, xs, cvs):
        assert c1.iter == c2.iter and c1.is_converged == c2.is_converged
        assert _same(x1, x2)
    # every batched answer is also the oracle's, in the device reduction order
    for q, (a, amg) in enumerate(keep):
        path = export(hiers[q])
        oh = O.OracleHier(path)
        os.unlink(path)
        for k in range(3):
            i = 3 * q + k
            xo, ro, _ = oh.solve(a.arrays(), rhs[i], nrms[i], m=20, eps=1e-10, bit=28,
                                 mode=O.RED_GPU)
            assert cvs[i].iter == ro.iters and cvs[i].is_converged == ro.converged, (q, k)
            assert _same(xs[i], xo), (q, k)


@pytest.mark.parametrize("kind,params", [(S.GEN_POISSON2D, (63,)), (S.GEN_LAPLACE3D, (17,))],
                         ids=["poisson63", "laplace17"])
def test_int_solve_odd_sizes_match_oracle(kind, params):
    """Systems whose size is not a multiple of four (3,969 and 4,913 rows),
    where the integer Gram-Schmidt kernels take their one-element loops
    instead of the 16-byte ones: x, the iteration count and every V-cycle
    checksum equal the oracle's."""
    sy = S.make_system(kind, params, 3)
    assert sy.a.n % 4 != 0
    hi, hd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    amg = S.Amg(hi)
    bit = S.Amg(hd).default_frac_bits(sy.b, sy.nrm_b)
    path = export(hi)
    oh = O.OracleHier(path)
    os.unlink(path)
    sol = S.Solver(S.INT, sy.a, amg, m=30, eps=1e-10, bit=bit)
    sol.trace(8192)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    xo, ro, to = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=30, eps=1e-10, bit=bit,
                          mode=O.RED_GPU, trace_cap=8192)
    print(f"[odd size] n {sy.a.n}: {cv.iter} iterations (oracle {ro.iters})", flush=True)
    assert (cv.iter, cv.is_converged) == (ro.iters, ro.converged)
    assert np.array_equal(sol.read_trace(8192), to)
    assert _same(x, xo)
