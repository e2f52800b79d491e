"""Host-side AMG setup: invariants that fix its mathematics (CPU).

  * SOR / BlockSOR / JacobiSOR split (smoother/stationary.hpp:385-396,
    586-631, 941-978): A = strict lower part of LD + diag(1 / stored inverse)
    + U, where U holds the explicit (1 - 1/omega) a_ii slot (first in each
    SOR row, in column order otherwise); JacobiSOR's LD holds only columns
    left of the row's 1200-row block, so one block's rows are independent;
  * BlockSOR with one block is SOR;
  * classical RS interpolation (helper/helper_multigrid.hpp:457-532): the
    rows of C points are unit rows, one per coarse column, in order;
  * R = P^T and A_{l+1} = R A_l P within rounding (helper_sparse.hpp:213-285)
    in the FP64 hierarchy, which has no row scaling (INTAMG.hpp:25-106);
  * after the second pass (helper_multigrid.hpp:378-417) every F point i and
    strong F neighbour j share a strong C point;
  * 1D Poisson coarsens to alternating C/F points.
"""
from __future__ import annotations

import os

import numpy as np
import pytest
import scipy.sparse as sp

import sc23ad as S
from helpers import export, read_hier

CASES = {
    "poisson40": (S.GEN_POISSON2D, (40,), 0),
    "laplace12": (S.GEN_LAPLACE3D, (12,), 1),
    "convdiff40": (S.GEN_CONVDIFF2D, (40, 50.0, 0.9), 3),
    "diffusion27_12": (S.GEN_DIFFUSION27, (12, 4, 1.0, 1.0, 1e-3, 2), 2),
}


def _csr(M):
    return sp.csr_matrix((M.val.astype(np.float64), M.cind, M.rptr), shape=(M.n, M.m))


def _hier(case, prec=S.FP64, smoother=S.SOR, **kw):
    kind, params, seed = CASES[case]
    sy = S.make_system(kind, params, seed)
    h = S.Hierarchy(sy.a, prec, S.amg_param(smoother, **kw))
    path = export(h)
    try:
        return sy, read_hier(path)
    finally:
        os.unlink(path)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("smoother,omega", [(S.SOR, 1.0), (S.SOR, 1.3), (S.BLOCK_SOR, 1.0),
                                            (S.BLOCK_SOR, 0.8), (S.JACOBI_SOR, 1.0),
                                            (S.JACOBI_SOR, 1.2)])
def test_split_sums_to_a(case, smoother, omega):
    _, hf = _hier(case, S.FP64, smoother, omega=omega)
    for lvl, L in enumerate(hf.levels):
        A, LD, U = L["A"], L["LD"], L["U"]
        n = A.n
        rebuilt = np.zeros((n, n))
        for i in range(n):
            p0, p1 = LD.rptr[i], LD.rptr[i + 1]
            assert p1 > p0
            for p in range(p0, p1 - 1):
                assert LD.cind[p] < i
                if smoother == S.JACOBI_SOR:
                    assert LD.cind[p] < (i // 1200) * 1200
                rebuilt[i, LD.cind[p]] += LD.val[p]
            assert LD.cind[p1 - 1] == i          # inverse diagonal stored last
            diag_ld = 1.0 / LD.val[p1 - 1]       # a_ii / omega
            ucols = list(U.cind[U.rptr[i]:U.rptr[i + 1]])
            assert ucols.count(i) == 1           # explicit (1 - 1/omega) a_ii slot
            if smoother == S.SOR:
                assert ucols[0] == i             # first in the row
            else:
                assert ucols == sorted(ucols)    # column order, earlier blocks included
            rebuilt[i, i] += diag_ld
            for p in range(U.rptr[i], U.rptr[i + 1]):
                rebuilt[i, U.cind[p]] += U.val[p]
        dense = _csr(A).toarray()
        np.testing.assert_allclose(rebuilt, dense, rtol=1e-13, atol=1e-13 * abs(dense).max(),
                                   err_msg=f"level {lvl}")
        diag = np.diag(dense)
        udiag = np.array([U.val[U.rptr[i] + list(U.cind[U.rptr[i]:U.rptr[i + 1]]).index(i)]
                          for i in range(n)])
        np.testing.assert_allclose(udiag, (1 - 1 / omega) * diag, rtol=1e-13, atol=1e-15)


def test_blocksor_one_block_is_sor():
    _, a = _hier("poisson40", S.INT, S.SOR)
    _, b = _hier("poisson40", S.INT, S.BLOCK_SOR, b_num=1)
    assert a.depth == b.depth
    for la, lb in zip(a.levels, b.levels):
        for k in ("A", "LD", "U", "R", "P"):
            for f in ("rptr", "cind", "val"):
                assert np.array_equal(getattr(la[k], f), getattr(lb[k], f)), k
            assert (la[k].bit, la[k].dbit) == (lb[k].bit, lb[k].dbit)


def _strength(A, theta):
    S_rows = []
    for i in range(A.n):
        cols, vals = A.cind[A.rptr[i]:A.rptr[i + 1]], A.val[A.rptr[i]:A.rptr[i + 1]]
        off = [(c, abs(v)) for c, v in zip(cols, vals) if c != i]
        mx = max((v for _, v in off), default=0.0)
        S_rows.append({c for c, v in off if v >= theta * mx and mx > 0})
    return S_rows


@pytest.mark.parametrize("case", list(CASES))
def test_interpolation_and_galerkin(case):
    _, hf = _hier(case)
    for lvl, L in enumerate(hf.levels):
        A, R, P = _csr(L["A"]), _csr(L["R"]), _csr(L["P"])
        Pm = L["P"]
        unit = [i for i in range(Pm.n) if Pm.rptr[i + 1] - Pm.rptr[i] == 1
                and Pm.val[Pm.rptr[i]] == 1.0]
        # every coarse column has a unit row, and they can be taken in row
        # order (C points are numbered in order); an F point interpolating
        # from one C point only is a unit row too, so the C set is exact only
        # when the counts agree
        chosen, k = [], 0
        for i in unit:
            if k < Pm.m and Pm.cind[Pm.rptr[i]] == k:
                chosen.append(i)
                k += 1
        assert k == Pm.m, f"level {lvl}: C rows are unit rows, in order"
        exact_c = len(unit) == Pm.m
        assert (R != P.T).nnz == 0, "R = P^T"
        nxt = hf.levels[lvl + 1]["A"] if lvl + 1 < hf.depth else None
        if nxt is not None:
            Ac = _csr(nxt).toarray()
            want = (R @ A @ P).toarray()
            np.testing.assert_allclose(Ac, want, rtol=1e-12, atol=1e-12 * abs(want).max())
            assert np.all(nxt.val != 0), "exact zeros are dropped"
            assert np.all((Ac != 0) <= (want != 0) | (abs(want) < 1e-12 * abs(want).max()))
        if not exact_c:
            continue
        # second-pass property on this level's strength graph
        Sg = _strength(L["A"], hf.theta)
        is_c = np.zeros(L["A"].n, bool)
        is_c[unit] = True
        for i in range(L["A"].n):
            if is_c[i]:
                continue
            for j in Sg[i]:
                if not is_c[j]:
                    assert any(is_c[k] for k in Sg[i] & Sg[j]), (lvl, i, j)


def test_1d_poisson_alternates():
    n = 101
    rptr, cind, val = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                cind.append(j)
                val.append(v)
        rptr.append(len(cind))
    a = S.CsrMat.from_arrays(n, np.array(rptr, np.int32), np.array(cind, np.int32),
                             np.array(val))
    h = S.Hierarchy(a, S.FP64, S.amg_param(S.SOR))
    path = export(h)
    hf = read_hier(path)
    os.unlink(path)
    P = hf.levels[0]["P"]
    is_c = [P.rptr[i + 1] - P.rptr[i] == 1 and P.val[P.rptr[i]] == 1.0 for i in range(n)]
    assert all(is_c[i] != is_c[i + 1] for i in range(n - 1))
    # F points interpolate with weight 1/2 from both neighbours
    for i in range(1, n - 1):
        if not is_c[i]:
            np.testing.assert_allclose(P.val[P.rptr[i]:P.rptr[i + 1]], [0.5, 0.5])


def test_jacobisor_single_block_is_damped_jacobi():
    """n <= 1200: LD = D / omega (diagonal only), U = A - D / omega."""
    sy = S.make_system(S.GEN_POISSON2D, (32,), 0)     # 1024 rows, one smoothed level
    path = export(S.Hierarchy(sy.a, S.FP64, S.amg_param(S.JACOBI_SOR, omega=0.7)))
    try:
        h = read_hier(path)
    finally:
        os.unlink(path)
    L = h.levels[0]
    assert L["A"].n <= 1200
    A, LD, U = L["A"], L["LD"], L["U"]
    assert np.array_equal(LD.rptr, np.arange(A.n + 1))
    assert np.array_equal(LD.cind, np.arange(A.n))
    dense = _csr(A).toarray()
    np.testing.assert_allclose(1.0 / LD.val, np.diag(dense) / 0.7, rtol=1e-14)
    np.testing.assert_allclose(_csr(U).toarray(), dense - np.diag(np.diag(dense)) / 0.7,
                               rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("omega", [1.0, 0.7])
def test_jacobi_factors(omega):
    """Jacobi (FP): LD holds omega / a_ii as its one entry per row (omega
    converted to T, divided in double by the FP64 diagonal, stored as T, as
    Jacobi<CSR<T>> computes it), U is A."""
    sy = S.make_system(S.GEN_LAPLACE3D, (20,), 1)      # two smoothed levels
    hs = S.Hierarchy.build_multi(sy.a, [S.FP64, S.FP32], S.amg_param(S.JACOBI, omega=omega))
    paths = [export(h) for h in hs]
    try:
        h64, h32 = (read_hier(p) for p in paths)
    finally:
        for p in paths:
            os.unlink(p)
    assert h64.depth >= 2
    for h, w in ((h64, omega), (h32, float(np.float32(omega)))):
        for L, L64 in zip(h.levels, h64.levels):
            A, LD, U = L["A"], L["LD"], L["U"]
            assert np.array_equal(LD.rptr, np.arange(A.n + 1))
            assert np.array_equal(LD.cind, np.arange(A.n))
            diag = np.diag(_csr(L64["A"]).toarray())
            np.testing.assert_array_equal(LD.val, (w / diag).astype(LD.val.dtype))
            for f in ("rptr", "cind", "val"):
                assert np.array_equal(getattr(U, f), getattr(A, f))
