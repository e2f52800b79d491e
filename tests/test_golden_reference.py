"""Setup and oracle against the reference's own outputs (CPU).

tests/golden/reference.json holds what the reference, compiled in this
container (oracle/_ref), computes on seeded systems: hierarchy file hashes
for FP64/FP32/INT x SOR/BlockSOR(40) (preconditioner/INTAMG.hpp:25-106,
149-245 over helper/helper_multigrid.hpp:34-586) and the results of
Restarted(FGMRES) for solvers 0/1/2 (solver/irsolver.hpp:27-43,
solver/gmres.hpp:18-271), including the checksum of every integer V-cycle
input and output.  Here the product's host setup must write byte-identical
hierarchies, and the oracle in the reference's sequential reduction order
must reproduce every result bit for bit.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as O
import sc23ad as S
from helpers import export, file_sha, sha, system


def _cases(golden):
    return golden["cases"]


@pytest.fixture(scope="module")
def systems(golden):
    return {c["name"]: system(c) for c in _cases(golden)}


def test_generated_systems_match_fixture(golden, systems):
    for c in _cases(golden):
        sy = systems[c["name"]]
        assert sy.a.n == c["n"] and sy.a.nnz == c["nnz"], c["name"]
        assert sha(*sy.a.arrays()) == c["a_sha"], c["name"]
        assert sha(sy.b) == c["b_sha"], c["name"]
        assert sy.nrm_b == c["nrm_b"], c["name"]


SMOOTHERS = [0, 1, 2, 3]
SM_IDS = ["sor", "blocksor", "jacobisor", "jacobi"]


def _precs(smoother):
    """Jacobi exists for the FP hierarchies only."""
    return (S.FP64, S.FP32) if smoother == 3 else (S.INT, S.FP64, S.FP32)


@pytest.mark.parametrize("prec", [0, 1, 2], ids=["fp64", "fp32", "int"])
@pytest.mark.parametrize("smoother", SMOOTHERS, ids=SM_IDS)
def test_setup_byte_identical_to_reference(golden, systems, prec, smoother):
    if prec == S.INT and smoother == 3:
        with pytest.raises(S.Sc23adError) as e:
            S.Hierarchy(systems[golden["cases"][0]["name"]].a, prec, S.amg_param(smoother))
        assert e.value.code == S.BAD_PARAM
        return
    for c in _cases(golden):
        h = S.Hierarchy(systems[c["name"]].a, prec, S.amg_param(smoother))
        path = export(h)
        try:
            assert file_sha(path) == c["hier"][f"{prec}_{smoother}"], c["name"]
        finally:
            os.unlink(path)


@pytest.fixture(scope="module")
def oracle_hiers(golden, systems):
    out = {}
    for c in _cases(golden):
        a = systems[c["name"]].a
        for smoother in SMOOTHERS:
            precs = _precs(smoother)
            hs = S.Hierarchy.build_multi(a, list(precs), S.amg_param(smoother))
            for prec, h in zip(precs, hs):
                path = export(h)
                out[(c["name"], prec, smoother)] = O.OracleHier(path)
                os.unlink(path)
    return out


def test_default_bits_match_reference(golden, systems, oracle_hiers):
    for c in _cases(golden):
        sy = systems[c["name"]]
        for smoother in (0, 1, 2):
            bit = oracle_hiers[(c["name"], S.FP64, smoother)].default_bits(sy.b, sy.nrm_b)
            assert bit == c["solves"][f"2_{smoother}"]["bit"], c["name"]


@pytest.mark.parametrize("smoother", [0, 1, 2], ids=["sor", "blocksor", "jacobisor"])
def test_oracle_int_solve_matches_reference(golden, systems, oracle_hiers, smoother):
    for c in _cases(golden):
        sy = systems[c["name"]]
        g = c["solves"][f"2_{smoother}"]
        x, res, trace = oracle_hiers[(c["name"], S.INT, smoother)].solve(
            sy.a.arrays(), sy.b, sy.nrm_b, m=c["m"], eps=golden["eps"], bit=g["bit"],
            mode=O.RED_SEQ, trace_cap=8192)
        assert [format(int(v), "016x") for v in trace] == g["trace"], c["name"]
        assert (res.iters, res.converged) == (g["iters"], g["converged"]), c["name"]
        assert sha(x) == g["x_sha"], c["name"]


@pytest.mark.parametrize("solver", [0, 1], ids=["fp64", "fp32"])
@pytest.mark.parametrize("smoother", SMOOTHERS, ids=SM_IDS)
def test_oracle_fp_solve_matches_reference(golden, systems, oracle_hiers, solver, smoother):
    for c in _cases(golden):
        sy = systems[c["name"]]
        g = c["solves"][f"{solver}_{smoother}"]
        x, res, _ = oracle_hiers[(c["name"], solver, smoother)].solve(
            sy.a.arrays(), sy.b, sy.nrm_b, m=c["m"], eps=golden["eps"], mode=O.RED_SEQ)
        assert (res.iters, res.converged) == (g["iters"], g["converged"]), c["name"]
        assert sha(x) == g["x_sha"], c["name"]
        assert O.relres(sy.a.arrays(), sy.b, sy.nrm_b, x, O.RED_SEQ) == g["relres"]


# ---- live comparison where the reference is compiled (this container) -------

live = pytest.mark.skipif(not O.ref_available(),
                          reason="oracle/_ref not built (reference checkout absent)")


@live
@pytest.mark.parametrize("case", [
    (S.GEN_POISSON2D, (37,), 9), (S.GEN_LAPLACE3D, (13,), 8),
    (S.GEN_CONVDIFF2D, (45, 80.0, 4.0), 7),
    (S.GEN_DIFFUSION27, (14, 7, 1.0, 0.5, 1e-2, 12), 6),
], ids=["poisson37", "laplace13", "convdiff45", "diffusion27_14"])
def test_live_reference(case, tmp_path):
    kind, params, seed = case
    sy = S.make_system(kind, params, seed)
    arrs = sy.a.arrays()
    for smoother in SMOOTHERS:
        for prec in ((0, 1) if smoother == 3 else (0, 2)):
            ref_path = str(tmp_path / f"ref_{prec}_{smoother}.bin")
            O.ref_build_hier(arrs, prec, smoother, ref_path)
            h = S.Hierarchy(sy.a, prec, S.amg_param(smoother))
            mine = str(tmp_path / f"mine_{prec}_{smoother}.bin")
            h.export(mine)
            assert file_sha(mine) == file_sha(ref_path)
        if smoother == 3:
            # Jacobi: the FP64 solve through the oracle against the reference's
            R = O.ref_solve(arrs, sy.b, 0, smoother, m=30, eps=1e-10)
            oh = O.OracleHier(str(tmp_path / f"ref_0_{smoother}.bin"))
            x, res, _ = oh.solve(arrs, sy.b, sy.nrm_b, m=30, eps=1e-10, mode=O.RED_SEQ)
            assert res.iters == R["iters"]
            assert np.array_equal(x, R["x"])
            continue
        R = O.ref_solve(arrs, sy.b, 2, smoother, m=30, eps=1e-10, trace_cap=8192)
        oh = O.OracleHier(str(tmp_path / f"ref_2_{smoother}.bin"))
        x, res, trace = oh.solve(arrs, sy.b, sy.nrm_b, m=30, eps=1e-10, bit=R["bit"],
                                 mode=O.RED_SEQ, trace_cap=8192)
        assert np.array_equal(trace, R["trace"])
        assert res.iters == R["iters"]
        assert np.array_equal(x, R["x"])


# ---- relaxation weights other than 1 (tests/golden/reference_omega.json) ------
#
# With omega != 1 the SOR / BlockSOR split stores a_ii / omega in LD and a
# non-zero (1 - 1/omega) a_ii in U (smoother/stationary.hpp:375-396,
# 941-978).  Some weights make the reference's integer solver overflow: it
# wraps in blas1::Add and Nrm2, then lsqrt(<= 0) = 0 and the division at
# gmres.hpp:244 kills the process ("crashed" in the fixture).  The oracle
# must stop with OVERFLOW exactly there.

@pytest.fixture(scope="module")
def golden_omega():
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "reference_omega.json")) as f:
        return json.load(f)


def _omega_cases(golden_omega):
    for c in golden_omega["cases"]:
        yield c, system(c)


def test_omega_setup_byte_identical_to_reference(golden_omega):
    for c, sy in _omega_cases(golden_omega):
        for om in golden_omega["omegas"]:
            for sm in (0, 1):
                hs = S.Hierarchy.build_multi(sy.a, [S.FP64, S.INT], S.amg_param(sm, omega=om))
                for prec, h in zip((0, 2), hs):
                    path = export(h)
                    try:
                        assert file_sha(path) == c["hier"][f"{om}_{prec}_{sm}"], (c["name"], om)
                    finally:
                        os.unlink(path)


def test_omega_oracle_solves_match_reference(golden_omega):
    for c, sy in _omega_cases(golden_omega):
        for om in golden_omega["omegas"]:
            for sm in (0, 1):
                hs = S.Hierarchy.build_multi(sy.a, [S.FP64, S.INT], S.amg_param(sm, omega=om))
                ohs = []
                for h in hs:
                    path = export(h)
                    ohs.append(O.OracleHier(path))
                    os.unlink(path)
                # F by the driver's rule on the FP64 hierarchy (main.cpp:153-160)
                rule_bit = ohs[0].default_bits(sy.b, sy.nrm_b)
                for solver, oh in zip((0, 2), ohs):
                    g = c["solves"][f"{om}_{solver}_{sm}"]
                    if "bit" in g:
                        assert g["bit"] == rule_bit, (c["name"], om, sm)
                    bit = rule_bit
                    x, res, trace = oh.solve(sy.a.arrays(), sy.b, sy.nrm_b, m=c["m"],
                                             eps=golden_omega["eps"], bit=bit, mode=O.RED_SEQ,
                                             trace_cap=8192)
                    label = (c["name"], om, solver, sm)
                    if "crashed" in g:
                        assert solver == 2 and res.status == S.OVERFLOW, label
                        continue
                    assert (res.iters, res.converged) == (g["iters"], g["converged"]), label
                    assert sha(x) == g["x_sha"], label
                    if solver == 2:
                        assert [format(int(v), "016x") for v in trace] == g["trace"], label


def test_omega_fixture_covers_both_outcomes(golden_omega):
    """The fixture holds converged, diverged and overflowing weights, so each
    branch above is exercised."""
    outs = [s for c in golden_omega["cases"] for s in c["solves"].values()]
    assert any("crashed" in s for s in outs)
    assert any(s.get("converged") == 1 for s in outs)
    assert any(s.get("converged") == 0 for s in outs)


def test_full_size_hierarchy_matches_reference():
    """The host setup of the 128^3 Laplacian (BlockSOR, integer) writes the
    reference's own hierarchy file (tests/golden/reference_laplace3d_128.json)."""
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden",
                           "reference_laplace3d_128.json")) as f:
        g = json.load(f)
    sy = S.make_system(g["kind"], tuple(g["params"]), g["seed"])
    assert sha(*sy.a.arrays()) == g["a_sha"] and sha(sy.b) == g["b_sha"]
    h = S.Hierarchy(sy.a, S.INT, S.amg_param(g["smoother"]))
    path = export(h)
    try:
        assert file_sha(path) == g["hier_int"]
    finally:
        os.unlink(path)


def test_batched_convdiff_outcomes_match_reference():
    """A sample of the batched convection-diffusion workload
    (tests/golden/reference_convdiff_batch.json): where the reference's
    integer solver dies (the wrapped sum of squares and lsqrt(<= 0) = 0 as a
    divisor, gmres.hpp:243-244) the oracle stops with OVERFLOW; where the
    reference converges the oracle reproduces its iterations and x."""
    import json
    import importlib.util
    here = os.path.dirname(__file__)
    with open(os.path.join(here, "golden", "reference_convdiff_batch.json")) as f:
        g = json.load(f)
    spec = importlib.util.spec_from_file_location(
        "make_batch_golden", os.path.join(here, "golden", "make_batch_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    crashed = converged = 0
    for c in g["cases"]:
        a, rmax, _ = mk.system(c["system"])
        b = mk.rhs(a, rmax, c["system"], c["rhs"])
        a.lmax1_scaling(np.ones(a.n))
        assert sha(*a.arrays()) == c["a_sha"] and sha(b) == c["b_sha"]
        nrm = float(np.sqrt(np.add.accumulate(b * b)[-1]))
        hi, hd = S.Hierarchy.build_multi(a, [S.INT, S.FP64], S.amg_param(g["smoother"]))
        ohs = []
        for h in (hi, hd):
            path = export(h)
            ohs.append(O.OracleHier(path))
            os.unlink(path)
        bit = ohs[1].default_bits(b, nrm)
        x, res, _ = ohs[0].solve(a.arrays(), b, nrm, m=g["m"], eps=g["eps"], bit=bit,
                                 mode=O.RED_SEQ)
        label = (c["system"], c["rhs"])
        if "crashed" in c["int"]:
            assert res.status == S.OVERFLOW, label
            crashed += 1
        else:
            assert bit == c["int"]["bit"], label
            assert (res.iters, res.converged) == (c["int"]["iters"], c["int"]["converged"]), label
            assert sha(x) == c["int"]["x_sha"], label
            converged += 1
        assert c["fp64"]["converged"] == 1, label
    assert crashed > 0 and converged > 0
