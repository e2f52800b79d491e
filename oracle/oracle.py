"""ctypes binding of the CPU oracle (oracle.c) and of the reference harness
(_ref/libref.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
SEQ_REF = os.path.join(HERE, "_ref", "seq_ref")
MULTI_REF = os.path.join(HERE, "_ref", "multi_ref")

RED_SEQ, RED_GPU = 0, 1


class OrResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int),
                ("restarts", C.c_int), ("status", C.c_int),
                ("trace_len", C.c_int)]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} not built (make -C oracle)")
        L = C.CDLL(ORACLE_SO)
        vp, ip = C.c_void_p, C.c_int
        L.or_hier_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.or_hier_free.argtypes = [vp]
        L.or_hier_level_n.argtypes = [vp, ip]
        L.or_vcycle.argtypes = [vp, vp, vp]
        L.or_level_op.argtypes = [vp, ip, ip, vp, vp]
        L.or_default_bits.argtypes = [vp, vp, C.c_double]
        L.or_solve.argtypes = [vp, ip, vp, vp, vp, vp, C.c_double, ip, C.c_double,
                               ip, ip, ip, vp, vp, ip, C.POINTER(OrResult)]
        L.or_relres.restype = C.c_double
        L.or_relres.argtypes = [ip, vp, vp, vp, vp, C.c_double, vp, ip]
        L.or_checksum_i32.restype = C.c_uint64
        L.or_checksum_i32.argtypes = [vp, ip]
        L.or_lsqrt.restype = C.c_int64
        L.or_lsqrt.argtypes = [C.c_int64]
        L.or_nrm2_f64.restype = C.c_double
        L.or_nrm2_f64.argtypes = [vp, ip, ip]
        L.or_dot_f64.restype = C.c_double
        L.or_dot_f64.argtypes = [vp, vp, ip, ip]
        L.or_ggen_int.restype = C.c_int64
        L.or_ggen_int.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), ip, C.POINTER(C.c_int)]
        L.or_grot_int.argtypes = [C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), ip]
        L.or_dot_int.restype = C.c_int64
        L.or_dot_int.argtypes = [vp, vp, ip, ip]
        L.or_nrm2_int.restype = C.c_int64
        L.or_nrm2_int.argtypes = [vp, ip]
        L.or_axpy_int.argtypes = [C.c_int64, vp, vp, ip, ip]
        L.or_scal_int.argtypes = [C.c_int64, vp, ip, ip]
        _lib = L
    return _lib


DTYPES = {0: np.float64, 1: np.float32, 2: np.int32}


class OracleHier:
    """A hierarchy file loaded by the oracle."""

    def __init__(self, path: str):
        h = C.c_void_p()
        rc = lib().or_hier_load(path.encode(), C.byref(h))
        if rc != 0:
            raise IOError(f"oracle could not load {path} ({rc})")
        self.h = h
        raw = open(path, "rb").read(44)
        self.precision = int(np.frombuffer(raw[8:12], np.int32)[0])
        self.depth = int(np.frombuffer(raw[12:16], np.int32)[0])
        self.dtype = DTYPES[self.precision]

    def level_n(self, level: int) -> int:
        return lib().or_hier_level_n(self.h, level)

    def vcycle(self, vin: np.ndarray) -> np.ndarray:
        vin = np.ascontiguousarray(vin, dtype=self.dtype)
        out = np.zeros(self.level_n(0), self.dtype)
        lib().or_vcycle(self.h, _p(vin), _p(out))
        return out

    def level_op(self, level: int, op: int, vin: np.ndarray, out_n: int,
                 out_init: np.ndarray | None = None) -> np.ndarray:
        vin = np.ascontiguousarray(vin, dtype=self.dtype)
        out = (np.zeros(out_n, self.dtype) if out_init is None
               else np.array(out_init, dtype=self.dtype, copy=True))
        rc = lib().or_level_op(self.h, level, op, _p(vin), _p(out))
        if rc != 0:
            raise ValueError(f"or_level_op failed ({rc})")
        return out

    def default_bits(self, b: np.ndarray, nrm_b: float) -> int:
        return lib().or_default_bits(self.h, _p(np.ascontiguousarray(b)), nrm_b)

    def solve(self, a_arrays, b, nrm_b, m=30, eps=1e-10, bit=29,
              max_outer=None, mode=RED_GPU, trace_cap=0, x0=None):
        rptr, cind, val = a_arrays
        n = len(b)
        x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
        tr = np.zeros(max(trace_cap, 1), np.uint64)
        res = OrResult()
        lib().or_solve(self.h, n, _p(rptr), _p(cind), _p(val), _p(b), nrm_b,
                       m, eps, bit, 1000 // m if max_outer is None else max_outer,
                       mode, _p(x), _p(tr) if trace_cap else None, trace_cap,
                       C.byref(res))
        return x, res, tr[:res.trace_len]

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib().or_hier_free(self.h)
            self.h = C.c_void_p()


def relres(a_arrays, b, nrm_b, x, mode=RED_GPU) -> float:
    rptr, cind, val = a_arrays
    return lib().or_relres(len(b), _p(rptr), _p(cind), _p(val), _p(b), nrm_b,
                           _p(np.ascontiguousarray(x)), mode)


def checksum_i32(v: np.ndarray) -> int:
    v = np.ascontiguousarray(v, dtype=np.int32)
    return int(lib().or_checksum_i32(_p(v), len(v)))


# ---- reference harness (only where the reference was compiled) ----------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        R.ref_build_hier.restype = ip
        R.ref_build_hier.argtypes = [ip, vp, vp, vp, ip, ip, dp, ip, dp, C.c_char_p]
        R.ref_int_amg_create.restype = vp
        R.ref_int_amg_create.argtypes = [ip, vp, vp, vp, ip, dp, ip, dp]
        R.ref_int_amg_apply.argtypes = [vp, vp, vp]
        R.ref_int_amg_destroy.argtypes = [vp]
        R.ref_solve.argtypes = [ip, vp, vp, vp, vp, ip, ip, ip, dp, ip, dp, ip, dp,
                                vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp),
                                C.POINTER(ip), vp, ip, C.POINTER(ip)]
        _ref = R
    return _ref


def ref_build_hier(a_arrays, precision, smoother, path, omega=1.0, b_num=40,
                   theta=0.8) -> int:
    rptr, cind, val = a_arrays
    return ref().ref_build_hier(len(rptr) - 1, _p(rptr), _p(cind), _p(val),
                                precision, smoother, omega, b_num, theta,
                                path.encode())


def ref_solve(a_arrays, b, solver, smoother, m=30, eps=1e-10, bit=-1,
              omega=1.0, b_num=40, theta=0.8, trace_cap=0):
    rptr, cind, val = a_arrays
    n = len(b)
    x = np.zeros(n)
    iters, conv, bit_used, tlen = C.c_int(), C.c_int(), C.c_int(-1), C.c_int()
    rel = C.c_double()
    tr = np.zeros(max(trace_cap, 1), np.uint64)
    # The reference prints every iteration's residual; keep it off stdout.
    libc = C.CDLL(None)
    libc.fflush(None)
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        ref().ref_solve(n, _p(rptr), _p(cind), _p(val), _p(np.ascontiguousarray(b)),
                        solver, smoother, m, eps, bit, omega, b_num, theta, _p(x),
                        C.byref(iters), C.byref(conv), C.byref(rel),
                        C.byref(bit_used), _p(tr), trace_cap, C.byref(tlen))
    finally:
        libc.fflush(None)
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    return dict(x=x, iters=iters.value, converged=conv.value,
                relres=rel.value, bit=bit_used.value, trace=tr[:tlen.value])
