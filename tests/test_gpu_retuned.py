"""Hierarchies with relaxation weights other than 1 and retuned hierarchies
(per-level weights and bit caps, sc23ad_hier_retune) on the GPU.

* omega in {0.8, 1.3, 1.7} for SOR and BlockSOR: in the reference's
  sequential FP order the GPU integer and FP64 solves reproduce the
  reference's own results (tests/golden/reference_omega.json; the split of
  smoother/stationary.hpp:375-396,941-978), including the weights where the
  reference's integer solver overflows and dies at gmres.hpp:244 (the
  library stops with OVERFLOW there); in the default blocked order the
  integer solves equal the oracle's bit for bit.
* retuned hierarchies: the GPU solve equals the oracle's on the exported
  file (every V-cycle checksum, the iterations and x).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle as O
import sc23ad as S
from helpers import export, sha, system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden_omega():
    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_omega.json")) as f:
        return json.load(f)


def _oracle(h):
    path = export(h)
    try:
        return O.OracleHier(path)
    finally:
        os.unlink(path)


@pytest.mark.parametrize("smoother", [S.SOR, S.BLOCK_SOR], ids=["sor", "blocksor"])
def test_omega_solves_match_reference(golden_omega, smoother):
    for c in golden_omega["cases"]:
        sy = system(c)
        for om in golden_omega["omegas"]:
            print(f"[omega] {c['name']} omega {om} smoother {smoother}", flush=True)
            hd, hi = S.Hierarchy.build_multi(sy.a, [S.FP64, S.INT],
                                             S.amg_param(smoother, omega=om))
            amg_d, amg_i = S.Amg(hd), S.Amg(hi)
            bit = amg_d.default_frac_bits(sy.b, sy.nrm_b)
            for solver, amg in ((2, amg_i), (0, amg_d)):
                g = c["solves"][f"{om}_{solver}_{smoother}"]
                label = (c["name"], om, solver)
                sol = S.Solver(solver, sy.a, amg, m=c["m"], eps=golden_omega["eps"], bit=bit)
                sol.set_sequential_reduction(True)
                sol.trace(8192)
                x, cv = sol.solve(sy.b, sy.nrm_b)
                if "crashed" in g:
                    assert solver == 2 and cv.status == S.OVERFLOW and cv.overflow_flag, label
                    continue
                assert (cv.iter, cv.is_converged) == (g["iters"], g["converged"]), label
                assert sha(x) == g["x_sha"], label
                if solver == 2:
                    assert bit == g["bit"], label
                    tr = sol.read_trace(8192)
                    assert [format(int(v), "016x") for v in tr] == g["trace"], label
            # the default blocked order against the oracle
            oh = _oracle(hi)
            sol = S.Solver(S.INT, sy.a, amg_i, m=c["m"], eps=golden_omega["eps"], bit=bit)
            sol.trace(8192)
            x, cv = sol.solve(sy.b, sy.nrm_b)
            xo, ro, to = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=c["m"],
                                  eps=golden_omega["eps"], bit=bit, mode=O.RED_GPU,
                                  trace_cap=8192)
            assert cv.status == ro.status and cv.iter == ro.iters, (c["name"], om)
            assert np.array_equal(sol.read_trace(8192), to), (c["name"], om)
            assert np.array_equal(x.view(np.uint64), xo.view(np.uint64)), (c["name"], om)


@pytest.mark.parametrize("case", [
    (S.GEN_DIFFUSION27, (20, 5, 1.0, 1.0, 1e-3, 3), 3, S.BLOCK_SOR),
    (S.GEN_CONVDIFF2D, (70, 30.0, 2.2), 5, S.BLOCK_SOR),
    (S.GEN_LAPLACE3D, (18,), 6, S.SOR),
], ids=["diffusion27_20", "convdiff2d_70", "laplace3d_18_sor"])
def test_retuned_hierarchy_matches_oracle(case):
    kind, params, seed, smoother = case
    sy = S.make_system(kind, params, seed)
    base = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(smoother))
    depth = base[0].info()[0]
    rng = np.random.default_rng(seed)
    omega = rng.uniform(0.7, 1.6, depth)
    bits = rng.choice([0, 18, 22, 26, 30], depth).astype(np.int32)
    hi = base[0].retune(omega, bits, S.INT)
    hd = base[1].retune(omega, None, S.FP64)
    bit = S.Amg(hd).default_frac_bits(sy.b, sy.nrm_b)
    amg = S.Amg(hi)
    oh = _oracle(hi)
    print(f"[retuned] kind {kind} depth {depth} omega {np.round(omega, 2).tolist()} "
          f"bits {bits.tolist()} F {bit}", flush=True)
    sol = S.Solver(S.INT, sy.a, amg, m=30, eps=1e-9, bit=bit)
    sol.trace(8192)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    xo, ro, to = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=30, eps=1e-9, bit=bit,
                          mode=O.RED_GPU, trace_cap=8192)
    assert cv.status == ro.status and cv.iter == ro.iters
    assert np.array_equal(sol.read_trace(8192), to)
    assert np.array_equal(x.view(np.uint64), xo.view(np.uint64))
    # and the level operations of every retuned level
    rng2 = np.random.default_rng(1)
    for lv in range(depth):
        n = oh.level_n(lv)
        v = rng2.integers(-(2**29), 2**29, n).astype(np.int32)
        for op in (S.OP_SPMV_A, S.OP_SPMV_U, S.OP_TRSV_LD, S.OP_SMOOTH0):
            assert np.array_equal(amg.level_op(lv, op, v, n), oh.level_op(lv, op, v, n)), (lv, op)
