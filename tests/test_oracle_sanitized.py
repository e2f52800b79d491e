"""Undefined behaviour and memory errors, checked on the CPU.

The oracle (oracle/oracle.c) is built with AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/Makefile, target sanitize_driver) and driven
through V-cycles, every level operation and an integer and an FP64 solve of
every smoother.  The reference's integer kernels rely on two's-complement
wrap-around of signed types, which C++ leaves undefined (blas1.hpp:49,
:294); the restatement does that arithmetic in unsigned types, so a clean run
shows that none of its results depends on undefined behaviour, and that every
index stays in bounds.  Where the reference is present, its own build under
UBSan (target _ref/ubsan_ref) shows the overflow that the restatement turns
into the overflow status.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
import sc23ad as S
from helpers import export

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
REF_LIB = "/root/reference/Library"


def _make(target):
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc / make not found")
    r = subprocess.run(["make", "-s", "-C", ORACLE, target], capture_output=True, text=True)
    if r.returncode != 0:
        if "asan" in r.stderr or "ubsan" in r.stderr:
            pytest.skip("sanitizer runtimes not installed: " + r.stderr[-200:])
        raise AssertionError(r.stderr)
    return os.path.join(ORACLE, target)


@pytest.fixture(scope="module")
def driver():
    return _make("sanitize_driver")


def _write_scaled(path, sy):
    """int32 n, nnz; rptr, cind; f64 val, b (row-scaled)."""
    rptr, cind, val = sy.a.arrays()
    with open(path, "wb") as f:
        f.write(np.array([sy.a.n, len(cind)], dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(rptr, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(cind, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(val, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(sy.b, dtype=np.float64).tobytes())


def _run_driver(driver, tmp_path, sy, smoother, bit=None):
    hi, hd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(smoother))
    hd_path = export(hd)
    if bit is None:
        bit = O.OracleHier(hd_path).default_bits(sy.b, sy.nrm_b)
    sysf = str(tmp_path / "system.bin")
    _write_scaled(sysf, sy)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([driver, export(hi), hd_path, sysf, str(bit)], capture_output=True,
                       text=True, env=env, timeout=300)
    for word in ("runtime error", "AddressSanitizer", "LeakSanitizer"):
        assert word not in r.stderr, r.stderr[-3000:]
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr[-3000:])
    m = re.search(r"int solve status (\d+) iters (\d+); fp64 solve status (\d+) iters (\d+)",
                  r.stdout)
    assert m, r.stdout
    return tuple(int(g) for g in m.groups())


CASES = [
    ("laplace3d_10", S.GEN_LAPLACE3D, (10,)),
    ("inclusions3d_12", S.GEN_INCLUSIONS3D, (12, 4, 3.0, 6.0, 1e6, 4)),
    ("convdiff2d_40", S.GEN_CONVDIFF2D, (40, 60.0, 0.7)),
]


@pytest.mark.parametrize("smoother", [S.SOR, S.BLOCK_SOR, S.JACOBI_SOR])
@pytest.mark.parametrize("name,kind,params", CASES, ids=[c[0] for c in CASES])
def test_oracle_clean_under_sanitizers(driver, tmp_path, name, kind, params, smoother):
    st_i, it_i, st_d, it_d = _run_driver(driver, tmp_path, S.make_system(kind, params, 3),
                                         smoother)
    assert (st_i, st_d) == (0, 0)
    assert it_i > 0 and it_d > 0


def test_overflow_is_a_status_not_undefined_behaviour(driver, tmp_path):
    """30 fractional bits overflow poisson2d_64's Givens generation: the
    restatement stops with the overflow status without a sanitizer report."""
    st_i, it_i, st_d, _ = _run_driver(driver, tmp_path, S.make_system(S.GEN_POISSON2D, (64,), 0),
                                      S.SOR, bit=30)
    assert st_i == S.OVERFLOW and it_i == 0
    assert st_d == 0


@pytest.mark.skipif(not os.path.isdir(REF_LIB), reason="reference checkout absent")
def test_reference_relies_on_signed_wrap(tmp_path):
    """The same solve in the reference (seq_ref built under UBSan): signed
    int32 and int64 additions overflow (blas1.hpp:49 Add, :294 Ggen's
    a*a + b*b) before lsqrt(<= 0) = 0 divides by zero."""
    exe = _make("_ref/ubsan_ref")
    rptr, cind, val = S.CsrMat.generate(S.GEN_POISSON2D, 64).arrays()
    sy = S.make_system(S.GEN_POISSON2D, (64,), 0)
    sysf = str(tmp_path / "system_unscaled.bin")
    with open(sysf, "wb") as f:  # "SC23SYS1", n, nnz, rptr, cind, val, b (unscaled)
        f.write(b"SC23SYS1" + np.array([len(rptr) - 1, len(cind)], np.int32).tobytes())
        for arr, t in ((rptr, np.int32), (cind, np.int32), (val, np.float64),
                       (sy.b_unscaled, np.float64)):
            f.write(np.ascontiguousarray(arr, t).tobytes())
    r = subprocess.run([exe, sysf, "0", "2", "30", "30", "1e-10"], capture_output=True,
                       text=True, timeout=300)
    errs = [l for l in r.stderr.splitlines() if "runtime error: signed integer overflow" in l]
    assert any("blas1.hpp:294" in l for l in errs), r.stderr[-2000:]
