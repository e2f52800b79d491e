"""Block solves: several right-hand sides of one hierarchy in shared launches.

The block V-cycle (INTAMG.hpp:269-293 on 4 * groups vectors) and the block
solver (k Restarted(FGMRESflex(m)) solves in lockstep, irsolver.hpp:27-43,
gmres.hpp:225-271) must give every right-hand side exactly what its own
single solve gives: every V-cycle output bit for bit against the oracle, and
for whole solves x, the iteration count, the status, the V-cycle checksum
trace and the residual history against both the single device solve and the
oracle.  Cases: the 4-RHS groups padded (k = 1, 3, 5), SOR hierarchies whose
LD solves have no block kernel (one right-hand side at a time inside the
block V-cycle), right-hand sides that converge at different steps and
restarts, one that overflows, and a batch of block solvers over systems of
different sizes.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as O
import sc23ad as S
from helpers import export, system

pytestmark = pytest.mark.gpu


def _oracle(h):
    path = export(h)
    try:
        return O.OracleHier(path)
    finally:
        os.unlink(path)


def _rhs(a, seeds):
    """Row-scaled right-hand sides b = A x* (x* from each seed), as the
    driver scales them, with their norms."""
    rptr, cind, val = a.arrays()
    rmax = np.maximum.reduceat(np.abs(val), rptr[:-1])
    bs = [a.matvec(S.xstar(a.n, sd)) / rmax for sd in seeds]
    nrms = [float(np.sqrt(np.add.accumulate(b * b)[-1])) for b in bs]
    return bs, nrms


def _scaled(kind, params):
    a = S.CsrMat.generate(kind, *params)
    b0 = np.ones(a.n)
    a.lmax1_scaling(b0)
    return a


@pytest.mark.parametrize("case,smoother", [
    ("laplace3d_16", S.BLOCK_SOR), ("convdiff2d_60", S.BLOCK_SOR),
    ("diffusion27_16", S.BLOCK_SOR), ("inclusions3d_20", S.BLOCK_SOR),
    ("poisson2d_64", S.SOR), ("laplace3d_16", S.JACOBI_SOR),
], ids=["laplace16_bsor", "convdiff60_bsor", "diff27_bsor", "incl20_bsor", "poisson64_sor",
        "laplace16_jsor"])
@pytest.mark.parametrize("k", [1, 3, 4, 5, 8])
def test_block_vcycle_matches_oracle(golden, case, smoother, k):
    c = next(c for c in golden["cases"] if c["name"] == case)
    sy = system(c)
    h = S.Hierarchy(sy.a, S.INT, S.amg_param(smoother))
    amg = S.Amg(h)
    oh = _oracle(h)
    rng = np.random.default_rng(k)
    vins = rng.integers(-(2**29), 2**29, (k, sy.a.n)).astype(np.int32)
    print(f"[block vcycle] {case} smoother {smoother} k {k} native {S.block_native(amg)}",
          flush=True)
    outs = S.precondition_block(amg, vins)
    for r in range(k):
        assert np.array_equal(outs[r], oh.vcycle(vins[r])), r
    # twice: no state left behind between calls
    outs2 = S.precondition_block(amg, vins[::-1].copy())
    assert np.array_equal(outs2, outs[::-1])


def _compare_block_with_singles(a, hi, hd, bs, nrms, m, eps, bits=None, seq_check=True):
    amg_i = S.Amg(hi)
    if bits is None:
        amg_d = S.Amg(hd)
        bits = [amg_d.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
    k = len(bs)
    blk = S.BlockSolver(a, amg_i, bits, m=m, eps=eps)
    blk.trace(8192)
    xb, cvb = blk.solve(bs, nrms)
    oh = _oracle(hi)
    for r in range(k):
        sol = S.Solver(S.INT, a, amg_i, m=m, eps=eps, bit=bits[r])
        sol.trace(8192)
        x1, c1 = sol.solve(bs[r], nrms[r])
        label = (r, bits[r], c1.status, c1.iter)
        assert (cvb[r].status, cvb[r].iter, cvb[r].is_converged, cvb[r].restarts) == \
            (c1.status, c1.iter, c1.is_converged, c1.restarts), label
        assert np.array_equal(xb[r].view(np.uint64), x1.view(np.uint64)), label
        assert np.array_equal(blk.read_trace(r, 8192), sol.read_trace(8192)), label
        assert np.array_equal(blk.history(r), sol.history()), label
        assert cvb[r].implicit_relres == c1.implicit_relres, label
        if seq_check:
            xo, ro, to = oh.solve(a.arrays(), bs[r], nrms[r], m=m, eps=eps, bit=bits[r],
                                  mode=O.RED_GPU, trace_cap=8192)
            assert (ro.status, ro.iters) == (c1.status, c1.iter), label
            assert np.array_equal(xo.view(np.uint64), xb[r].view(np.uint64)), label
            assert np.array_equal(to, blk.read_trace(r, 8192)), label
    return cvb


@pytest.mark.parametrize("kind,params,smoother,k,m", [
    (S.GEN_LAPLACE3D, (18,), S.BLOCK_SOR, 5, 30),
    (S.GEN_CONVDIFF2D, (80, 60.0, 0.9), S.BLOCK_SOR, 4, 20),
    (S.GEN_DIFFUSION27, (18, 6, 1.0, 1.0, 1e-3, 9), S.BLOCK_SOR, 3, 10),
    (S.GEN_INCLUSIONS3D, (22, 8, 2.0, 4.0, 1e6, 5), S.BLOCK_SOR, 6, 30),
    (S.GEN_POISSON2D, (70,), S.SOR, 2, 30),
    (S.GEN_CONVDIFF2D, (45, 30.0, 1.2), S.BLOCK_SOR, 3, 20),
], ids=["laplace18_k5", "convdiff80_k4_m20", "diff27_k3_m10", "incl22_k6", "poisson70_sor_k2",
        "convdiff45_odd_k3"])
def test_block_solve_matches_single_solves(kind, params, smoother, k, m):
    a = _scaled(kind, params)
    hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], S.amg_param(smoother))
    bs, nrms = _rhs(a, [31 + 7 * r for r in range(k)])
    print(f"[block solve] kind {kind} {params} k {k} m {m}", flush=True)
    cvb = _compare_block_with_singles(a, hi, hd, bs, nrms, m, 1e-9)
    print("  iterations", [c.iter for c in cvb], "statuses", [c.status for c in cvb], flush=True)


def test_block_solve_mixed_outcomes():
    """Right-hand sides with different F on the SOR omega = 1.3 hierarchy of the
    2D Poisson 64^2 system (where the reference's integer solver dies at
    gmres.hpp:244): some overflow, the others converge at different steps and
    restarts; each keeps its own single solve's outcome."""
    a = _scaled(S.GEN_POISSON2D, (64,))
    prm = S.amg_param(S.SOR, omega=1.3)
    hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], prm)
    bs, nrms = _rhs(a, [0, 3, 4, 5, 6, 7])
    # F = 29 and 30 overflow in the second restart; the others converge after
    # 39-66 steps, one of them after a second restart
    bits = [29, 30, 24, 20, 26, 27]
    cvb = _compare_block_with_singles(a, hi, hd, bs, nrms, 30, 1e-10, bits=bits)
    statuses = {c.status for c in cvb}
    print("  statuses", [c.status for c in cvb], "iterations", [c.iter for c in cvb], flush=True)
    assert len(statuses) >= 2  # not all the same outcome


def test_block_solve_batch_over_systems():
    """Block solvers of several systems of different sizes, all at once, give
    each system's block solve."""
    specs = [(S.GEN_CONVDIFF2D, (40, 20.0, 0.3)), (S.GEN_CONVDIFF2D, (90, 75.0, 4.0)),
             (S.GEN_POISSON2D, (50,)), (S.GEN_LAPLACE3D, (12,))]
    solvers, bss, nbs, keep = [], [], [], []
    for q, (kind, params) in enumerate(specs):
        a = _scaled(kind, params)
        hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
        amg_i, amg_d = S.Amg(hi), S.Amg(hd)
        bs, nrms = _rhs(a, [100 * q + r for r in range(4)])
        bits = [amg_d.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
        solvers.append(S.BlockSolver(a, amg_i, bits, m=20, eps=1e-10))
        bss.append(bs)
        nbs.append(nrms)
        keep.append((a, amg_i, amg_d, bits))
    xs, cvs = S.block_solve_batch(solvers, bss, nbs, threads=4)
    for q, (a, amg_i, _, bits) in enumerate(keep):
        one = S.BlockSolver(a, amg_i, bits, m=20, eps=1e-10)
        x1, c1 = one.solve(bss[q], nbs[q])
        assert np.array_equal(xs[q].view(np.uint64), x1.view(np.uint64)), q
        assert [c.iter for c in cvs[q]] == [c.iter for c in c1], q
        assert [c.status for c in cvs[q]] == [c.status for c in c1], q


def test_block_gram_schmidt_one_launch_and_many_agree(monkeypatch):
    """The one-launch Gram-Schmidt step (by default for systems of at most
    8,192 rows; forced here) and the multi-workgroup one (j + 3 launches)
    give the same block solves on an 8,100-row system: x, steps, statuses and
    residual histories, for right-hand sides that converge at different
    steps."""
    a = _scaled(S.GEN_CONVDIFF2D, (90, 75.0, 4.0))
    hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    amg_i, amg_d = S.Amg(hi), S.Amg(hd)
    bs, nrms = _rhs(a, [3, 5, 8, 13, 21])
    bits = [amg_d.default_frac_bits(b, nb) for b, nb in zip(bs, nrms)]
    runs = {}
    for bound in ("0", "100000"):
        monkeypatch.setenv("SC23AD_MGS_FUSED_MAX", bound)
        blk = S.BlockSolver(a, amg_i, bits, m=20, eps=1e-10)
        xb, cvb = blk.solve(bs, nrms)
        runs[bound] = (xb, [(c.status, c.iter, c.restarts) for c in cvb],
                       [blk.history(r) for r in range(len(bs))])
        print(f"[mgs] bound {bound}: {runs[bound][1]}", flush=True)
    (x0, c0, h0), (x1, c1, h1) = runs["0"], runs["100000"]
    assert c0 == c1
    for r in range(len(bs)):
        assert np.array_equal(x0[r].view(np.uint64), x1[r].view(np.uint64)), r
        assert np.array_equal(h0[r], h1[r]), r
