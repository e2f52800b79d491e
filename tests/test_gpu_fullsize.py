"""The 128^3 7-point Laplacian (2,097,152 rows), FGMRES(30), BlockSOR(40):
the benchmark workload at full size (MI355X), plus one integer V-cycle
with each single-chain smoother (SOR, JacobiSOR).

Checked bit for bit against the oracle: the level-0 integer SpMV, one integer
and one FP64 V-cycle, and the whole integer solve (every V-cycle input and
output checksum, and x).  Size-independent properties: the explicit relative
residual reaches 1e-8 and the iteration count equals the oracle's.  Against
the reference itself (tests/golden/reference_laplace3d_128.json, written by
the reference compiled in the container): the integer hierarchy file's hash,
and in the sequential FP reduction order every V-cycle checksum of the
37-step solve, the iteration count and x.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import tempfile

import numpy as np
import pytest

import oracle as O
import sc23ad as S

pytestmark = pytest.mark.gpu

N = 128


@pytest.fixture(scope="module")
def big():
    print(f"[setup] laplace3d {N}^3 BlockSOR", flush=True)
    sy = S.make_system(S.GEN_LAPLACE3D, (N,), 1)
    hi, hd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    out = {"sy": sy, "amg_i": S.Amg(hi), "amg_d": S.Amg(hd)}
    for key, h in (("or_i", hi), ("or_d", hd)):
        fd, path = tempfile.mkstemp(suffix=".bin")
        os.close(fd)
        h.export(path)
        out[key] = O.OracleHier(path)
        if key == "or_i":
            with open(path, "rb") as f:
                out["hier_sha"] = hashlib.sha256(f.read()).hexdigest()
        os.unlink(path)
    return out


def _sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_reference_trace_at_full_size(big):
    """The GPU solve in the reference's sequential FP order reproduces the
    reference's own 128^3 run bit for bit."""
    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "reference_laplace3d_128.json")) as f:
        g = json.load(f)
    sy = big["sy"]
    assert _sha(*sy.a.arrays()) == g["a_sha"] and _sha(sy.b) == g["b_sha"]
    assert big["hier_sha"] == g["hier_int"]
    bit = big["amg_d"].default_frac_bits(sy.b, sy.nrm_b)
    assert bit == g["solve"]["bit"]
    sol = S.Solver(S.INT, sy.a, big["amg_i"], m=g["m"], eps=g["eps"], bit=bit)
    sol.set_sequential_reduction(True)
    sol.trace(4096)
    print(f"[solve] int bit {bit}, sequential reductions", flush=True)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    trace = sol.read_trace(4096)
    assert [format(int(v), "016x") for v in trace] == g["solve"]["trace"]
    assert (cv.iter, cv.is_converged) == (g["solve"]["iters"], g["solve"]["converged"])
    assert _sha(x) == g["solve"]["x_sha"]


def test_spmv_and_vcycles(big):
    sy = big["sy"]
    n = sy.a.n
    assert n == N**3
    rng = np.random.default_rng(5)
    v = rng.integers(-(2**29), 2**29, n).astype(np.int32)
    print("[spmv] level 0", flush=True)
    assert np.array_equal(big["amg_i"].level_op(0, S.OP_SPMV_A, v, n),
                          big["or_i"].level_op(0, S.OP_SPMV_A, v, n))
    print("[vcycle] int", flush=True)
    assert np.array_equal(big["amg_i"].precondition(v), big["or_i"].vcycle(v))
    print("[vcycle] fp64", flush=True)
    d = rng.standard_normal(n)
    assert np.array_equal(big["amg_d"].precondition(d).view(np.uint64),
                          big["or_d"].vcycle(d).view(np.uint64))


def test_fp32_far_dependencies(big, monkeypatch, capfd):
    """FP32 LD solves over several lanes per row with far dependencies (read
    from global memory, at least 40 levels back), forced at 128^3 by the
    smallest LDS ring the far rule admits (the FP64 hierarchy takes them by
    itself, its ring being too small for level 1's span): every LD solve and
    one V-cycle are the oracle's bit for bit, and the far layout did occur."""
    monkeypatch.setenv("SC23AD_FP_RING_FAR", "1")
    monkeypatch.setenv("SC23AD_VERBOSE", "1")
    sy = big["sy"]
    h = S.Hierarchy(sy.a, S.FP32, S.amg_param(S.BLOCK_SOR))
    capfd.readouterr()
    amg = S.Amg(h)
    err = capfd.readouterr().err
    far = [ln for ln in err.splitlines() if ln.startswith("[sc23ad plan] kind 0")
           and re.search(r"ep waves \d+ far 1 ", ln)
           and int(re.search(r"rows \d+ lanes (\d+) identity", ln).group(1)) > 1]
    print(f"[fp32 far] {len(far)} far multi-lane LD plans", flush=True)
    assert far
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    h.export(path)
    oh = O.OracleHier(path)
    os.unlink(path)
    rng = np.random.default_rng(43)
    for lv in range(amg.depth):
        n = oh.level_n(lv)
        x = rng.standard_normal(n).astype(np.float32)
        assert np.array_equal(amg.level_op(lv, S.OP_TRSV_LD, x, n).view(np.uint32),
                              oh.level_op(lv, S.OP_TRSV_LD, x, n).view(np.uint32)), lv
    v = rng.standard_normal(sy.a.n).astype(np.float32)
    assert np.array_equal(amg.precondition(v).view(np.uint32), oh.vcycle(v).view(np.uint32))


def test_coarse_tiles_repeated(big):
    """The tiled coarse LU solve at 128^3 (1,769 rows, 28 blocks of 64)
    equals the oracle's coarse solve three times in a row (the hand-off
    epochs advance between calls)."""
    amg, oh = big["amg_i"], big["or_i"]
    depth = amg.depth
    nc = oh.level_n(depth)
    rng = np.random.default_rng(47)
    for rep in range(3):
        xc = rng.integers(-(2**29), 2**29, nc).astype(np.int32)
        print(f"[coarse tiles] rep {rep}, {nc} rows", flush=True)
        assert np.array_equal(amg.level_op(depth, S.OP_COARSE, xc, nc),
                              oh.level_op(depth, S.OP_COARSE, xc, nc))


def test_block_vcycle_at_full_size(big):
    """The block V-cycle at 128^3 (five right-hand sides: a full group of four
    and a padded one) gives every right-hand side its single V-cycle, which
    test_spmv_and_vcycles pins to the oracle; the block kernels cover every
    level (the mid levels read their far dependencies from global memory)."""
    n = big["sy"].a.n
    amg = big["amg_i"]
    rng = np.random.default_rng(11)
    vins = rng.integers(-(2**29), 2**29, (5, n)).astype(np.int32)
    native = S.block_native(amg)
    print(f"[block vcycle] 128^3 native {native}", flush=True)
    outs = S.precondition_block(amg, vins)
    for r in range(5):
        assert np.array_equal(outs[r], amg.precondition(vins[r])), r
    assert native


def test_int_solve_to_1e8(big):
    sy = big["sy"]
    bit = big["amg_d"].default_frac_bits(sy.b, sy.nrm_b)
    assert bit == big["or_d"].default_bits(sy.b, sy.nrm_b)
    sol = S.Solver(S.INT, sy.a, big["amg_i"], m=30, eps=1e-8, bit=bit)
    sol.trace(4096)
    print(f"[solve] int bit {bit}", flush=True)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    trace = sol.read_trace(4096)
    assert cv.is_converged
    rel = sol.relres(sy.b, sy.nrm_b, x)
    assert rel < 1e-8
    print(f"[oracle] int solve ({cv.iter} iterations on the GPU)", flush=True)
    xo, ro, to = big["or_i"].solve(sy.a.arrays(), sy.b, sy.nrm_b, m=30, eps=1e-8, bit=bit,
                                   mode=O.RED_GPU, trace_cap=4096)
    assert np.array_equal(trace, to)
    assert cv.iter == ro.iters
    assert np.array_equal(x.view(np.uint64), xo.view(np.uint64))


@pytest.mark.parametrize("smoother", [S.SOR, S.JACOBI_SOR], ids=["sor", "jacobisor"])
def test_int_vcycle_single_chain_smoothers(big, smoother):
    """SOR (382 wavefront levels at 128^3) and JacobiSOR (1200-row blocks):
    one integer V-cycle bit for bit against the oracle."""
    sy = big["sy"]
    print(f"[setup] int hierarchy, smoother {smoother}", flush=True)
    h = S.Hierarchy(sy.a, S.INT, S.amg_param(smoother))
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    h.export(path)
    oh = O.OracleHier(path)
    os.unlink(path)
    amg = S.Amg(h)
    v = np.random.default_rng(6).integers(-(2**29), 2**29, sy.a.n).astype(np.int32)
    print(f"[vcycle] int smoother {smoother}", flush=True)
    assert np.array_equal(amg.precondition(v), oh.vcycle(v))


@pytest.mark.parametrize("rhs", [0, 1, 2])
def test_inclusions160_matches_reference(rhs):
    """Right-hand sides 0-2 of the 160^3 inclusions workload (3D 7-point
    diffusion, 64 spheres of k = 1e6, seed 4; x* seed 400 + k): in the
    reference's sequential FP order the GPU integer solve reproduces the
    reference's own run (tests/golden/reference_inclusions160_2_<k>.json):
    400 iterations for k = 0, the exhausted restart budget for k = 1 (no
    convergence), and for k = 2 convergence after 838 iterations at an
    explicit residual of 1.16e-8, above the tolerance -- every V-cycle
    checksum and x bit for bit."""
    with open(os.path.join(os.path.dirname(__file__), "golden",
                           f"reference_inclusions160_2_{rhs}.json")) as f:
        g = json.load(f)
    a, rmax, amg_i, amg_d = _incl160()
    assert _sha(*a.arrays()) == g["a_sha"]
    # b from the unscaled operator, as the workload builds it
    a0 = S.CsrMat.generate(S.GEN_INCLUSIONS3D, 160, 64, 4.0, 12.0, 1e6, 4)
    b = a0.matvec(S.xstar(a0.n, 400 + rhs)) / rmax
    del a0
    assert _sha(b) == g["b_sha"]
    nrm = float(np.sqrt(np.add.accumulate(b * b)[-1]))
    bit = amg_d.default_frac_bits(b, nrm)
    assert bit == g["solve"]["bit"]
    sol = S.Solver(S.INT, a, amg_i, m=g["m"], eps=g["eps"], bit=bit)
    sol.set_sequential_reduction(True)
    sol.trace(4096)
    print(f"[solve] inclusions 160^3 rhs {rhs}, sequential reductions", flush=True)
    x, cv = sol.solve(b, nrm)
    trace = sol.read_trace(4096)
    print(f"  {cv.iter} iterations, status {cv.status}, {len(trace)} checksums", flush=True)
    assert [format(int(v), "016x") for v in trace] == g["solve"]["trace"]
    assert (cv.iter, cv.is_converged) == (g["solve"]["iters"], g["solve"]["converged"])
    assert _sha(x) == g["solve"]["x_sha"]


def test_block_vcycle_inclusions160():
    """The block V-cycle on the 160^3 inclusions hierarchy (five right-hand
    sides) gives every right-hand side its single V-cycle; every level runs
    on block kernels, including the ones whose rows span several lanes and
    read far dependencies from global memory."""
    a, _, amg_i, _ = _incl160()
    rng = np.random.default_rng(13)
    vins = rng.integers(-(2**29), 2**29, (5, a.n)).astype(np.int32)
    native = S.block_native(amg_i)
    print(f"[block vcycle] inclusions 160^3 native {native}", flush=True)
    outs = S.precondition_block(amg_i, vins)
    for r in range(5):
        assert np.array_equal(outs[r], amg_i.precondition(vins[r])), r
    assert native


def _incl160():
    """The 160^3 inclusions system (scaled), the row maxima of the unscaled
    operator and its integer and FP64 hierarchies, built once per module."""
    key = "incl160"
    if key not in _CACHE:
        print("[setup] inclusions 160^3 BlockSOR", flush=True)
        a = S.CsrMat.generate(S.GEN_INCLUSIONS3D, 160, 64, 4.0, 12.0, 1e6, 4)
        rptr, cind, val = a.arrays()
        rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
        a.lmax1_scaling(np.ones(a.n))
        hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
        _CACHE.clear()
        _CACHE[key] = (a, rmax, S.Amg(hi), S.Amg(hd))
    return _CACHE[key]


_CACHE: dict = {}
