"""The allocation-free level operation (sc23ad_amg_level_op_noalloc) against
sc23ad_amg_level_op: the same results for every operation of every level,
called twice in a row with the second call on a smaller level (the static
scratch keeps the first call's contents), and levels whose vectors exceed
the static scratch rejected with BAD_PARAM."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import sc23ad as S

pytestmark = pytest.mark.gpu

S.lib.sc23ad_amg_level_op_noalloc.restype = C.c_int
S.lib.sc23ad_amg_level_op_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                              C.c_void_p]


def _noalloc(amg, level, op, vin, out_n, out_init=None):
    din = S.DevArray(vin)
    dout = S.DevArray(np.zeros(out_n, vin.dtype) if out_init is None else out_init)
    S.check(S.lib.sc23ad_amg_level_op_noalloc(amg.h, level, op, din.p, dout.p))
    return dout.get()


@pytest.mark.parametrize("prec", [S.INT, S.FP64, S.FP32], ids=["int", "fp64", "fp32"])
def test_noalloc_matches_level_op(prec):
    sy = S.make_system(S.GEN_LAPLACE3D, (40,), 1)
    h = S.Hierarchy(sy.a, prec, S.amg_param(S.BLOCK_SOR))
    amg = S.Amg(h)
    depth, level_n, _ = h.info()
    rng = np.random.default_rng(3)
    dt = S.DT[prec]

    def rand(n):
        if prec == S.INT:
            return rng.integers(-(2**29), 2**29, n).astype(np.int32)
        return rng.standard_normal(n).astype(dt)

    # every level from the finest down: each call after the first runs on a
    # smaller level than the one before it
    for lv in range(depth):
        n, nxt = int(level_n[lv]), int(level_n[lv + 1])
        x, xc, init = rand(n), rand(nxt), rand(n)
        print(f"[noalloc] prec {prec} level {lv} n {n}", flush=True)
        for op, vin, on in ((S.OP_SPMV_A, x, n), (S.OP_SPMV_U, x, n), (S.OP_TRSV_LD, x, n),
                            (S.OP_SMOOTH0, x, n), (S.OP_SPMV_R, x, nxt), (S.OP_SPMV_P, xc, n)):
            ref = amg.level_op(lv, op, vin, on)
            got = _noalloc(amg, lv, op, vin, on)
            assert np.array_equal(ref.view(np.uint8), got.view(np.uint8)), (lv, op)
        ref = amg.level_op(lv, S.OP_SMOOTH1, x, n, init)
        got = _noalloc(amg, lv, S.OP_SMOOTH1, x, n, init)
        assert np.array_equal(ref.view(np.uint8), got.view(np.uint8)), (lv, "smooth1")
    nc = int(level_n[depth])
    xc = rand(nc)
    ref = amg.level_op(depth, S.OP_COARSE, xc, nc)
    got = _noalloc(amg, depth, S.OP_COARSE, xc, nc)
    assert np.array_equal(ref.view(np.uint8), got.view(np.uint8))
    # a second round, largest level again after the coarsest: same answers
    x = rand(int(level_n[0]))
    ref = amg.level_op(0, S.OP_SMOOTH0, x, int(level_n[0]))
    got = _noalloc(amg, 0, S.OP_SMOOTH0, x, int(level_n[0]))
    assert np.array_equal(ref.view(np.uint8), got.view(np.uint8))


def test_noalloc_rejects_levels_beyond_the_scratch():
    # 1,440,000 FP64 values (11.5 MB) exceed the 8 MiB per staging vector
    a = S.CsrMat.generate(S.GEN_POISSON2D, 1200)
    h = S.Hierarchy(a, S.FP64, S.amg_param(S.BLOCK_SOR))
    amg = S.Amg(h)
    n = a.n
    din = S.DevArray(np.ones(n))
    dout = S.DevArray(np.zeros(n))
    rc = S.lib.sc23ad_amg_level_op_noalloc(amg.h, 0, S.OP_SPMV_A, din.p, dout.p)
    assert rc == S.BAD_PARAM
    assert b"static scratch" in S.lib.sc23ad_last_error()
    # a coarser level of the same hierarchy fits
    n1 = int(h.info()[1][1])
    x = np.random.default_rng(1).standard_normal(n1)
    ref = amg.level_op(1, S.OP_SPMV_A, x, n1)
    got = _noalloc(amg, 1, S.OP_SPMV_A, x, n1)
    assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))
