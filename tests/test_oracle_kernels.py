"""The oracle's kernels against closed forms and brute force (CPU).

Each check states the reference's arithmetic independently of the oracle's
C code: Python integers with explicit two's-complement wrapping for the
integer path (Library/blas/blas1.hpp:245-307, matrix/spmat.hpp:77-86,
trimat/ldtri.hpp:169-183, ltri.hpp:120-133, dutri.hpp:140-156), exact
integer square roots for lsqrt (blas1.hpp:106-119), invariants of the Givens
rotation, and left-to-right / documented blocked orders for FP reductions.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
import sc23ad as S
from helpers import export, read_hier

M64 = (1 << 64) - 1


def s64(v: int) -> int:
    v &= M64
    return v - (1 << 64) if v >> 63 else v


def s32(v: int) -> int:
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >> 31 else v


# ---- lsqrt ------------------------------------------------------------------

@pytest.mark.parametrize("x", [0, 1, 2, 3, 4, 15, 16, 17, 2**31 - 1, 2**32, 2**32 + 1,
                               2**62, 2**62 - 1, 2**62 + 1, 2**63 - 1, 10**18,
                               -1, -(2**63)])
def test_lsqrt_edges(x):
    want = math.isqrt(x) if x > 0 else 0
    assert O.lib().or_lsqrt(x) == want


def test_lsqrt_random():
    rng = np.random.default_rng(7)
    for x in rng.integers(1, 2**63 - 1, 2000, dtype=np.int64):
        assert O.lib().or_lsqrt(int(x)) == math.isqrt(int(x))


# ---- Givens -----------------------------------------------------------------

def _ggen(a, b, bit):
    c, s, ov = C.c_int64(), C.c_int64(), C.c_int(0)
    r = O.lib().or_ggen_int(a, b, C.byref(c), C.byref(s), bit, C.byref(ov))
    return r, c.value, s.value, ov.value


@pytest.mark.parametrize("bit", [20, 25, 29, 30])
def test_ggen_closed_form_and_invariants(bit):
    rng = np.random.default_rng(bit)
    one = 1 << bit
    for _ in range(500):
        a = int(rng.integers(-(2**31), 2**31))
        b = int(rng.integers(1, 2**31))
        r, c, s, ov = _ggen(a, b, bit)
        assert ov == 0
        rr = math.isqrt(a * a + b * b)
        inv = ((1 << 62) // rr) >> (62 - 2 * bit)
        assert r == rr
        assert c == (a * inv) >> bit            # Python >> floors, like arithmetic shift
        assert s == (-(b * inv)) >> bit          # floor of the negated product
        # rotation invariants: inv = 2^2F / r - d with 0 <= d < 1, so
        # c = a 2^F / r + e_c with |e_c| <= |a| / 2^F + 1 (likewise s), and
        # c^2 + s^2 = 2^2F (a^2 + b^2) / r^2 up to 2 2^F (E_c + E_s) + E_c^2 + E_s^2,
        # where r = floor(sqrt(a^2 + b^2)) puts (a^2 + b^2) / r^2 in [1, (1 + 1/r)^2]
        ec, es = abs(a) / one + 1, abs(b) / one + 1
        assert abs(c - a * one / rr) <= ec
        assert abs(s + b * one / rr) <= es
        slack = one * one * (2.0 / rr + 1.0 / rr**2)
        bound = 2.02 * one * (ec + es) + ec * ec + es * es + slack
        assert abs(c * c + s * s - one * one) <= bound


def test_ggen_zero_reports_overflow():
    r, c, s, ov = _ggen(0, 0, 29)
    assert (r, c, s, ov) == (0, 0, 0, 1)


def test_grot_closed_form():
    rng = np.random.default_rng(3)
    bit = 29
    for _ in range(500):
        c, s = (int(v) for v in rng.integers(-(2**29), 2**29, 2))
        a0, b0 = (int(v) for v in rng.integers(-(2**40), 2**40, 2))
        a, b = C.c_int64(a0), C.c_int64(b0)
        O.lib().or_grot_int(c, s, C.byref(a), C.byref(b), bit)
        assert a.value == s64(c * a0 - s * b0) >> bit
        assert b.value == s64(s * a0 + c * b0) >> bit


# ---- integer BLAS-1 ---------------------------------------------------------

def _rand_i32(rng, n, lim=2**31):
    return rng.integers(-lim, lim, n).astype(np.int32)


def test_int_dot_nrm2_axpy_scal():
    rng = np.random.default_rng(11)
    L = O.lib()
    for n in (0, 1, 7, 1000):
        x, y = _rand_i32(rng, n), _rand_i32(rng, n)
        bit = 29
        px, py = O._p(x), O._p(y)
        want = s64(sum(int(a) * int(b) for a, b in zip(x, y))) >> bit
        assert L.or_dot_int(px, py, n, bit) == want
        want = math.isqrt(s64(sum(int(a) * int(a) for a in x))) if n and \
            s64(sum(int(a) * int(a) for a in x)) > 0 else 0
        assert L.or_nrm2_int(px, n) == want
        alpha = int(rng.integers(-(2**40), 2**40))
        yy = y.copy()
        L.or_axpy_int(alpha, px, O._p(yy), n, bit)
        assert [int(v) for v in yy] == [s32(int(b) + s32(s64(alpha * int(a)) >> bit))
                                       for a, b in zip(x, y)]
        xx = x.copy()
        L.or_scal_int(alpha, O._p(xx), n, bit)
        assert [int(v) for v in xx] == [s32(s64(alpha * int(a)) >> bit) for a in x]


# ---- FP reductions ----------------------------------------------------------

def _blocked_sum(terms: np.ndarray, dtype) -> float:
    """The device library's documented order: 1024 x 256 virtual threads,
    thread g sums elements g, g + 262144, ... left to right, then halving
    trees inside each 256-thread block and over the 1024 block sums."""
    G = 1024 * 256
    part = np.zeros(G, dtype)
    for k in range(0, len(terms), G):
        chunk = terms[k:k + G].astype(dtype)
        part[:len(chunk)] = part[:len(chunk)] + chunk
    blk = part.reshape(1024, 256)
    stride = 128
    while stride >= 1:
        blk[:, :stride] = blk[:, :stride] + blk[:, stride:2 * stride]
        stride //= 2
    s = blk[:, 0].copy()
    stride = 512
    while stride >= 1:
        s[:stride] = s[:stride] + s[stride:2 * stride]
        stride //= 2
    return s[0]


@pytest.mark.parametrize("n", [1, 5, 1000, 300000, 600001])
def test_fp64_reductions_both_orders(n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    L = O.lib()
    seq = np.add.accumulate(x * y)[-1]
    assert L.or_dot_f64(O._p(x), O._p(y), n, O.RED_SEQ) == seq
    assert L.or_nrm2_f64(O._p(x), n, O.RED_SEQ) == math.sqrt(np.add.accumulate(x * x)[-1])
    assert L.or_dot_f64(O._p(x), O._p(y), n, O.RED_GPU) == _blocked_sum(x * y, np.float64)
    assert L.or_nrm2_f64(O._p(x), n, O.RED_GPU) == math.sqrt(_blocked_sum(x * x, np.float64))


def test_checksum_definition():
    rng = np.random.default_rng(5)
    v = _rand_i32(rng, 257)

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    want = sum(mix(((i * 0x9E3779B97F4A7C15) & M64) ^ (int(x) & 0xFFFFFFFF))
               for i, x in enumerate(v)) & M64
    assert O.checksum_i32(v) == want
    assert S.checksum_i32(v) == want


# ---- level operations on a real hierarchy, brute force ------------------------

@pytest.fixture(scope="module", params=[0, 1], ids=["sor", "blocksor"])
def int_hier(request):
    sy = S.make_system(S.GEN_POISSON2D, (40,), 0)
    h = S.Hierarchy(sy.a, S.INT, S.amg_param(request.param))
    path = export(h)
    return read_hier(path), O.OracleHier(path)


def _spmv(M, x):
    y = []
    for i in range(M.n):
        acc = 0
        for p in range(M.rptr[i], M.rptr[i + 1]):
            acc += int(M.val[p]) * int(x[M.cind[p]])
        y.append(s32(s64(acc) >> M.bit))
    return np.array(y, np.int32)


def _ld_solve(M, v):
    out = [0] * M.n
    for i in range(M.n):
        p0, p1 = M.rptr[i], M.rptr[i + 1]
        t = int(v[i]) << M.bit
        for p in range(p0, p1 - 1):
            t -= int(M.val[p]) * out[M.cind[p]]
        t = s64(t)
        out[i] = s32(s64((t >> M.bit) * int(M.val[p1 - 1])) >> M.dbit)
    return np.array(out, np.int32)


def _ltri(M, v):
    out = [0] * M.n
    for i in range(M.n):
        t = int(v[i]) << M.bit
        for p in range(M.rptr[i], M.rptr[i + 1]):
            if M.cind[p] < i:
                t -= int(M.val[p]) * out[M.cind[p]]
        out[i] = s32(s64(t) >> M.bit)
    return np.array(out, np.int32)


def _dutri(M, v):
    out = [0] * M.n
    for i in range(M.n - 1, -1, -1):
        p0, p1 = M.rptr[i], M.rptr[i + 1]
        t = int(v[i]) << M.bit
        for p in range(p0 + 1, p1):
            t -= int(M.val[p]) * out[M.cind[p]]
        t = s64(t)
        out[i] = s32(s64((t >> M.bit) * int(M.val[p0])) >> M.dbit)
    return np.array(out, np.int32)


def test_level_ops_brute_force(int_hier):
    hf, oh = int_hier
    rng = np.random.default_rng(1)
    for lvl in range(hf.depth):
        L = hf.levels[lvl]
        n = L["A"].n
        x = _rand_i32(rng, n, 2**29)
        for op, key in ((S.OP_SPMV_A, "A"), (S.OP_SPMV_U, "U")):
            assert np.array_equal(oh.level_op(lvl, op, x, n), _spmv(L[key], x))
        assert np.array_equal(oh.level_op(lvl, S.OP_SPMV_R, x, L["R"].n), _spmv(L["R"], x))
        xc = _rand_i32(rng, L["P"].m, 2**29)
        assert np.array_equal(oh.level_op(lvl, S.OP_SPMV_P, xc, n), _spmv(L["P"], xc))
        assert np.array_equal(oh.level_op(lvl, S.OP_TRSV_LD, x, n), _ld_solve(L["LD"], x))
        # Smooth(in, out, false) = LD^-1 in; Smooth(in, out, true): out = LD^-1 (in - U out)
        assert np.array_equal(oh.level_op(lvl, S.OP_SMOOTH0, x, n), _ld_solve(L["LD"], x))
        out0 = _rand_i32(rng, n, 2**29)
        t = np.array([s32(int(a) - int(b)) for a, b in zip(x, _spmv(L["U"], out0))], np.int32)
        assert np.array_equal(oh.level_op(lvl, S.OP_SMOOTH1, x, n, out0), _ld_solve(L["LD"], t))
    xc = _rand_i32(rng, hf.lc.n, 2**29)
    want = _dutri(hf.uc, _ltri(hf.lc, xc))
    assert np.array_equal(oh.level_op(hf.depth, S.OP_COARSE, xc, hf.lc.n), want)


def test_vcycle_composed_from_level_ops(int_hier):
    """intAMGflex::Precondition (preconditioner/INTAMG.hpp:269-293) composed by
    hand from the level operations checked above."""
    hf, oh = int_hier
    rng = np.random.default_rng(2)
    n0 = hf.levels[0]["A"].n
    vin = _rand_i32(rng, n0, 2**29)
    d = hf.depth

    def sub(a, b):
        return (a.astype(np.int64) - b).astype(np.int32)

    def add(a, b):
        return (a.astype(np.int64) + b).astype(np.int32)

    def n_of(level):
        return hf.levels[level]["A"].n if level < d else hf.lc.n

    out = oh.level_op(0, S.OP_SMOOTH0, vin, n0)
    r = sub(vin, oh.level_op(0, S.OP_SPMV_A, out, n0))
    ee = [None] * d
    xx = [None] * d
    ee[0] = oh.level_op(0, S.OP_SPMV_R, r, n_of(1))
    for i in range(d - 1):
        n = n_of(i + 1)
        xx[i] = oh.level_op(i + 1, S.OP_SMOOTH0, ee[i], n)
        rr = sub(ee[i], oh.level_op(i + 1, S.OP_SPMV_A, xx[i], n))
        ee[i + 1] = oh.level_op(i + 1, S.OP_SPMV_R, rr, n_of(i + 2))
    xx[d - 1] = oh.level_op(d, S.OP_COARSE, ee[d - 1], n_of(d))
    for i in range(d - 2, -1, -1):
        n = n_of(i + 1)
        xx[i] = add(oh.level_op(i + 1, S.OP_SPMV_P, xx[i + 1], n), xx[i])
        xx[i] = oh.level_op(i + 1, S.OP_SMOOTH1, ee[i], n, xx[i])
    out = add(oh.level_op(0, S.OP_SPMV_P, xx[0], n0), out)
    out = oh.level_op(0, S.OP_SMOOTH1, vin, n0, out)
    assert np.array_equal(oh.vcycle(vin), out)
