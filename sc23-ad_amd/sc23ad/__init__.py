"""ctypes binding of libsc23ad.so (include/sc23ad.h).

The library is the product: host-side AMG setup in C++ and the solve phase as
HIP kernels for gfx950.  This module only marshals numpy arrays across the C
ABI; it computes nothing itself.  Importing it loads the in-tree library and
fails loudly if it has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsc23ad.so")

OK, BAD_PARAM, SETUP_FAIL, OVERFLOW, STAGNATION, HIP_ERROR, NOT_CONVERGED, \
    IO_ERROR, NO_DEVICE = range(9)
FP64, FP32, INT = 0, 1, 2
SOR, BLOCK_SOR, JACOBI_SOR, JACOBI = 0, 1, 2, 3
GEN_POISSON2D, GEN_LAPLACE3D, GEN_DIFFUSION27, GEN_CONVDIFF2D, \
    GEN_INCLUSIONS3D = range(5)
OP_SPMV_A, OP_SPMV_U, OP_SPMV_R, OP_SPMV_P, OP_TRSV_LD, OP_SMOOTH0, \
    OP_SMOOTH1, OP_COARSE = range(8)


class Sc23adError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sc23ad error {code}: {msg}")
        self.code = code


class AmgParam(C.Structure):
    _fields_ = [("max_levels", C.c_int), ("theta", C.c_double),
                ("cycle", C.c_int), ("smoother", C.c_int),
                ("omega", C.c_double), ("iter", C.c_int),
                ("b_num", C.c_int), ("coarse_n", C.c_int)]


class Converge(C.Structure):
    _fields_ = [("is_converged", C.c_int), ("iter", C.c_int),
                ("restarts", C.c_int), ("status", C.c_int),
                ("overflow_flag", C.c_int), ("fallback_used", C.c_int),
                ("frac_bits", C.c_int), ("implicit_relres", C.c_double),
                ("solve_ms", C.c_double)]


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c "
            "'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    sig = {
        "sc23ad_last_error": (C.c_char_p, []),
        "sc23ad_version": (C.c_char_p, []),
        "sc23ad_amg_param_default": (None, [C.POINTER(AmgParam), C.c_int]),
        "sc23ad_csrmat_create": (C.c_int, [C.c_int, C.c_int, vp, vp, vp, C.POINTER(vp)]),
        "sc23ad_csrmat_read_mm": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "sc23ad_csrmat_gen": (C.c_int, [C.c_int, dp, C.c_int, C.POINTER(vp)]),
        "sc23ad_csrmat_info": (C.c_int, [vp, ip, ip, ip]),
        "sc23ad_csrmat_export": (C.c_int, [vp, vp, vp, vp]),
        "sc23ad_csrmat_lmax1_scaling": (C.c_int, [vp, vp]),
        "sc23ad_csrmat_matvec": (C.c_int, [vp, vp, vp]),
        "sc23ad_csrmat_destroy": (None, [vp]),
        "sc23ad_xstar": (C.c_int, [C.c_int, C.c_uint64, vp]),
        "sc23ad_convdiff_draws": (C.c_int, [C.c_int, C.c_uint64, vp, vp, vp]),
        "sc23ad_hier_build": (C.c_int, [vp, C.POINTER(AmgParam), C.c_int, C.POINTER(vp)]),
        "sc23ad_hier_info": (C.c_int, [vp, ip, vp, vp]),
        "sc23ad_hier_build_multi": (C.c_int, [vp, C.POINTER(AmgParam), vp, C.c_int, vp]),
        "sc23ad_hier_build_multi_dev": (C.c_int, [vp, C.POINTER(AmgParam), vp, C.c_int, vp]),
        "sc23ad_amg_vcycle_bytes": (C.c_int, [vp, dp]),
        "sc23ad_amg_time_vcycles": (C.c_int, [vp, C.c_int, vp, vp, dp]),
        "sc23ad_hier_export": (C.c_int, [vp, C.c_char_p]),
        "sc23ad_hier_destroy": (None, [vp]),
        "sc23ad_checksum_i32": (C.c_uint64, [vp, C.c_int]),
        "sc23ad_hier_retune": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.POINTER(vp)]),
        "sc23ad_hier_saturated": (C.c_int, [vp, ip]),
        "sc23ad_hier_retune_levels": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.POINTER(vp)]),
        "sc23ad_hier_features": (C.c_int, [vp, vp, C.c_int, ip]),
    }
    optional = {
        "sc23ad_device_count": (C.c_int, [ip]),
        "sc23ad_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
        "sc23ad_dev_free": (C.c_int, [vp]),
        "sc23ad_dev_upload": (C.c_int, [vp, vp, C.c_size_t]),
        "sc23ad_dev_download": (C.c_int, [vp, vp, C.c_size_t]),
        "sc23ad_dev_sync": (C.c_int, []),
        "sc23ad_dev_copy": (C.c_int, [vp, vp, C.c_size_t]),
        "sc23ad_spmv_int": (C.c_int, [C.c_int, vp, vp, vp, C.c_int, vp, vp]),
        "sc23ad_amg_create": (C.c_int, [vp, C.POINTER(vp)]),
        "sc23ad_amg_destroy": (None, [vp]),
        "sc23ad_amg_precision": (C.c_int, [vp, ip]),
        "sc23ad_amg_precondition": (C.c_int, [vp, vp, vp, C.c_int]),
        "sc23ad_amg_level_op": (C.c_int, [vp, C.c_int, C.c_int, vp, vp]),
        "sc23ad_solver_create": (C.c_int, [C.c_int, vp, vp, C.c_int, C.c_double,
                                           C.c_int, C.c_int, C.POINTER(vp)]),
        "sc23ad_solver_destroy": (None, [vp]),
        "sc23ad_solver_solve": (C.c_int, [vp, vp, C.c_double, vp, C.c_int,
                                          C.POINTER(Converge)]),
        "sc23ad_solver_trace": (C.c_int, [vp, vp, C.c_int, ip]),
        "sc23ad_solver_set_reduction": (C.c_int, [vp, C.c_int]),
        "sc23ad_solve_batch": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp),
                                         C.POINTER(C.c_double), C.POINTER(vp), C.c_int,
                                         C.c_int, C.POINTER(Converge)]),
        "sc23ad_solver_history": (C.c_int, [vp, vp, C.c_int, ip]),
        "sc23ad_solver_time_spmv": (C.c_int, [vp, C.c_int]),
        "sc23ad_solver_spmv_stats": (C.c_int, [vp, ip, dp]),
        "sc23ad_solver_trsv_stats": (C.c_int, [vp, ip, dp]),
        "sc23ad_solver_relres": (C.c_int, [vp, vp, C.c_double, vp, C.c_int, dp]),
        "sc23ad_default_frac_bits": (C.c_int, [vp, vp, C.c_double, ip]),
        "sc23ad_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(vp)]),
        "sc23ad_mlp_destroy": (None, [vp]),
        "sc23ad_mlp_nparams": (C.c_int, [ip]),
        "sc23ad_mlp_params": (C.c_int, [vp, vp, vp]),
        "sc23ad_mlp_forward": (C.c_int, [vp, vp, C.c_int, vp]),
        "sc23ad_mlp_grad": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(C.c_float), vp]),
        "sc23ad_mlp_train": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_float,
                                       C.POINTER(C.c_float)]),
        "sc23ad_learned_levels": (C.c_int, [vp, vp, vp, vp, C.c_int, ip]),
        "sc23ad_amg_precondition_block": (C.c_int, [vp, C.c_int, vp, vp, C.c_int]),
        "sc23ad_amg_time_vcycles_block": (C.c_int, [vp, C.c_int, C.c_int, vp, vp,
                                                    C.POINTER(C.c_double)]),
        "sc23ad_amg_block_native": (C.c_int, [vp, ip]),
        "sc23ad_block_solver_create": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_double, vp,
                                                 C.c_int, C.POINTER(vp)]),
        "sc23ad_block_solver_destroy": (None, [vp]),
        "sc23ad_block_solver_solve": (C.c_int, [vp, vp, vp, vp, C.c_int,
                                                C.POINTER(Converge)]),
        "sc23ad_block_solve_batch": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp),
                                               C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int,
                                               C.POINTER(vp)]),
        "sc23ad_block_solver_trace": (C.c_int, [vp, C.c_int, vp, C.c_int, ip]),
        "sc23ad_block_solver_history": (C.c_int, [vp, C.c_int, vp, C.c_int, ip]),
        "sc23ad_solver_set_stall": (C.c_int, [vp, C.c_int]),
        "sc23ad_block_solver_set_stall": (C.c_int, [vp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    for name, (res, args) in optional.items():
        f = getattr(lib, name, None)
        if f is not None:
            f.restype = res
            f.argtypes = args
    return lib


lib = _load()


def check(code: int) -> None:
    if code != OK:
        raise Sc23adError(code, lib.sc23ad_last_error().decode())


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def amg_param(smoother: int = SOR, **kw) -> AmgParam:
    p = AmgParam()
    lib.sc23ad_amg_param_default(C.byref(p), smoother)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class CsrMat:
    """Host CSR matrix (the reference's CSRMat)."""

    def __init__(self, handle: int):
        self.h = C.c_void_p(handle)

    @classmethod
    def from_arrays(cls, n, rptr, cind, val, m=None):
        rptr = np.ascontiguousarray(rptr, dtype=np.int32)
        cind = np.ascontiguousarray(cind, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        h = C.c_void_p()
        check(lib.sc23ad_csrmat_create(n, n if m is None else m, ptr(rptr),
                                       ptr(cind), ptr(val), C.byref(h)))
        return cls(h.value)

    @classmethod
    def generate(cls, kind: int, *params: float):
        p = np.asarray(params, dtype=np.float64)
        h = C.c_void_p()
        check(lib.sc23ad_csrmat_gen(kind, p.ctypes.data_as(C.POINTER(C.c_double)),
                                    len(p), C.byref(h)))
        return cls(h.value)

    @classmethod
    def read_mm(cls, path: str):
        h = C.c_void_p()
        check(lib.sc23ad_csrmat_read_mm(path.encode(), C.byref(h)))
        return cls(h.value)

    def info(self):
        n, m, nnz = C.c_int(), C.c_int(), C.c_int()
        check(lib.sc23ad_csrmat_info(self.h, C.byref(n), C.byref(m), C.byref(nnz)))
        return n.value, m.value, nnz.value

    @property
    def n(self) -> int:
        return self.info()[0]

    @property
    def nnz(self) -> int:
        return self.info()[2]

    def arrays(self):
        n, _, nnz = self.info()
        rptr = np.empty(n + 1, np.int32)
        cind = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        check(lib.sc23ad_csrmat_export(self.h, ptr(rptr), ptr(cind), ptr(val)))
        return rptr, cind, val

    def lmax1_scaling(self, b: np.ndarray) -> None:
        assert b.dtype == np.float64 and b.flags.c_contiguous
        check(lib.sc23ad_csrmat_lmax1_scaling(self.h, ptr(b)))

    def matvec(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n, np.float64)
        check(lib.sc23ad_csrmat_matvec(self.h, ptr(x), ptr(y)))
        return y

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.sc23ad_csrmat_destroy(self.h)
            self.h = C.c_void_p()


def xstar(n: int, seed: int) -> np.ndarray:
    x = np.empty(n, np.float64)
    check(lib.sc23ad_xstar(n, seed, ptr(x)))
    return x


def convdiff_draws(count: int, seed: int):
    n = np.empty(count, np.int32)
    pe = np.empty(count, np.float64)
    ang = np.empty(count, np.float64)
    check(lib.sc23ad_convdiff_draws(count, seed, ptr(n), ptr(pe), ptr(ang)))
    return n, pe, ang


class Hierarchy:
    """Host AMG hierarchy (setup only; no GPU needed)."""

    def __init__(self, a: CsrMat | None, precision: int = INT,
                 param: AmgParam | None = None, _handle=None):
        self.precision = precision
        if _handle is not None:
            self.h = _handle
            return
        h = C.c_void_p()
        prm = param if param is not None else amg_param()
        check(lib.sc23ad_hier_build(a.h, C.byref(prm), precision, C.byref(h)))
        self.h = h

    @classmethod
    def build_multi(cls, a: CsrMat, precisions, param: AmgParam | None = None,
                    device_setup: bool = False):
        """Hierarchies for several precisions from one coarsening chain;
        device_setup computes the interpolation weights and the Galerkin
        products on the GPU (identical results)."""
        prm = param if param is not None else amg_param()
        pr = np.asarray(precisions, dtype=np.int32)
        outs = (C.c_void_p * len(pr))()
        fn = lib.sc23ad_hier_build_multi_dev if device_setup else lib.sc23ad_hier_build_multi
        check(fn(a.h, C.byref(prm), ptr(pr), len(pr), outs))
        return [cls(None, int(p), _handle=C.c_void_p(outs[k])) for k, p in enumerate(pr)]

    def info(self):
        depth = C.c_int()
        check(lib.sc23ad_hier_info(self.h, C.byref(depth), None, None))
        n = np.empty(depth.value + 1, np.int32)
        nnz = np.empty(depth.value + 1, np.int32)
        check(lib.sc23ad_hier_info(self.h, C.byref(depth), ptr(n), ptr(nnz)))
        return depth.value, n, nnz

    def retune(self, level_omega=None, level_bits=None, precision: int | None = None):
        """Hierarchy with per-level smoother weights / bit caps, from this
        one's coarsening."""
        nl = self.info()[0]
        om = None if level_omega is None else np.ascontiguousarray(level_omega, np.float64)
        bi = None if level_bits is None else np.ascontiguousarray(level_bits, np.int32)
        n = len(om) if om is not None else (len(bi) if bi is not None else nl)
        h = C.c_void_p()
        prec = self.precision if precision is None else precision
        check(lib.sc23ad_hier_retune(self.h, ptr(om), ptr(bi), n, prec, C.byref(h)))
        return Hierarchy(None, prec, _handle=h)

    def retune_levels(self, level_omega=None, level_vbits=None,
                      precision: int | None = None):
        """Hierarchy with per-level smoother weights and per-level vector
        fractional bits (level_vbits: depth + 1 entries, the first 0)."""
        nl = self.info()[0]
        om = None if level_omega is None else np.ascontiguousarray(level_omega, np.float64)
        vb = None if level_vbits is None else np.ascontiguousarray(level_vbits, np.int32)
        if om is not None and len(om) != nl:
            raise ValueError(f"level_omega: expected {nl} values")
        if vb is not None and len(vb) != nl + 1:
            raise ValueError(f"level_vbits: expected {nl + 1} values")
        h = C.c_void_p()
        prec = self.precision if precision is None else precision
        check(lib.sc23ad_hier_retune_levels(self.h, ptr(om), ptr(vb), nl, prec, C.byref(h)))
        return Hierarchy(None, prec, _handle=h)

    def saturated(self) -> int:
        v = C.c_int()
        check(lib.sc23ad_hier_saturated(self.h, C.byref(v)))
        return v.value

    def features(self) -> np.ndarray:
        """8 features per level (levels x 8, float32)."""
        nl = C.c_int()
        check(lib.sc23ad_hier_features(self.h, None, 0, C.byref(nl)))
        f = np.zeros((max(nl.value, 1), 8), np.float32)
        check(lib.sc23ad_hier_features(self.h, ptr(f), nl.value, C.byref(nl)))
        return f[:nl.value]

    def export(self, path: str) -> None:
        check(lib.sc23ad_hier_export(self.h, path.encode()))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.sc23ad_hier_destroy(self.h)
            self.h = C.c_void_p()


def checksum_i32(v: np.ndarray) -> int:
    v = np.ascontiguousarray(v, dtype=np.int32)
    return int(lib.sc23ad_checksum_i32(ptr(v), len(v)))


@dataclass
class System:
    """A generated, row-scaled linear system as the reference driver sees it
    (Work/main.cpp:55-63): A and b = A x* scaled by LMax1Scaling."""
    a: CsrMat
    b: np.ndarray
    xstar: np.ndarray
    nrm_b: float
    b_unscaled: np.ndarray


def make_system(kind: int, params, seed: int) -> System:
    a = CsrMat.generate(kind, *params)
    xs = xstar(a.n, seed)
    b = a.matvec(xs)
    b0 = b.copy()
    a.lmax1_scaling(b)
    nrm_b = _seq_nrm2(b)
    return System(a, b, xs, nrm_b, b0)


def _seq_nrm2(b: np.ndarray) -> float:
    """sqrt of the left-to-right sum of squares, as blas1::Nrm2 in a
    single-threaded build computes nrm_b (Work/main.cpp:63)."""
    if len(b) == 0:
        return 0.0
    # np.add.accumulate adds strictly left to right (unlike np.sum).
    return float(np.sqrt(np.add.accumulate(b * b)[-1]))


# ---- device -----------------------------------------------------------------

def device_count() -> int:
    n = C.c_int()
    check(lib.sc23ad_device_count(C.byref(n)))
    return n.value


class DevArray:
    """A device buffer holding a copy of a numpy array."""

    def __init__(self, a: np.ndarray | None = None, nbytes: int | None = None,
                 dtype=None, shape=None):
        if a is not None:
            a = np.ascontiguousarray(a)
            nbytes, dtype, shape = a.nbytes, a.dtype, a.shape
        self.nbytes, self.dtype, self.shape = nbytes, np.dtype(dtype), shape
        p = C.c_void_p()
        check(lib.sc23ad_dev_alloc(max(nbytes, 4), C.byref(p)))
        self.p = p
        if a is not None and nbytes:
            check(lib.sc23ad_dev_upload(self.p, ptr(a), nbytes))

    @classmethod
    def empty(cls, shape, dtype):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return cls(nbytes=n, dtype=dtype, shape=shape)

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib.sc23ad_dev_download(ptr(out), self.p, self.nbytes))
        return out

    def set(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib.sc23ad_dev_upload(self.p, ptr(a), self.nbytes))

    def view(self, offset: int, count: int) -> "DevView":
        """Elements [offset, offset + count) of a one-dimensional array, as a
        device pointer the solvers accept (no copy; the view does not own)."""
        if offset < 0 or count < 0 or (offset + count) * self.dtype.itemsize > self.nbytes:
            raise ValueError("view outside the array")
        return DevView(self, offset, count)

    def __del__(self):
        if getattr(self, "owner", True) and getattr(self, "p", None) is not None \
                and self.p.value:
            lib.sc23ad_dev_free(self.p)
            self.p = C.c_void_p()


class DevView:
    """A slice of a DevArray (keeps its parent alive)."""

    def __init__(self, parent: DevArray, offset: int, count: int):
        self.parent = parent
        self.dtype = parent.dtype
        self.shape = (count,)
        self.nbytes = count * parent.dtype.itemsize
        self.p = C.c_void_p(parent.p.value + offset * parent.dtype.itemsize)

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib.sc23ad_dev_download(ptr(out), self.p, self.nbytes))
        return out


DT = {FP64: np.float64, FP32: np.float32, INT: np.int32}


class Amg:
    """A hierarchy resident on the GPU (the reference's preconditioner)."""

    def __init__(self, hier: Hierarchy):
        h = C.c_void_p()
        check(lib.sc23ad_amg_create(hier.h, C.byref(h)))
        self.h = h
        self.precision = hier.precision
        self.dtype = DT[hier.precision]
        self.depth, self.level_n, _ = hier.info()

    def precondition(self, vin: np.ndarray) -> np.ndarray:
        vin = np.ascontiguousarray(vin, dtype=self.dtype)
        out = np.zeros(self.level_n[0], self.dtype)
        check(lib.sc23ad_amg_precondition(self.h, ptr(vin), ptr(out), 0))
        return out

    def vcycle_bytes(self) -> float:
        b = C.c_double()
        check(lib.sc23ad_amg_vcycle_bytes(self.h, C.byref(b)))
        return b.value

    def time_vcycles(self, vin: np.ndarray, reps: int = 20):
        """(mean milliseconds of one V-cycle on device vectors, its output)."""
        din = DevArray(np.ascontiguousarray(vin, dtype=self.dtype))
        dout = DevArray(np.zeros(self.level_n[0], self.dtype))
        ms = C.c_double()
        check(lib.sc23ad_amg_time_vcycles(self.h, reps, din.p, dout.p, C.byref(ms)))
        return ms.value, dout.get()

    def level_op(self, level: int, op: int, vin: np.ndarray, out_n: int,
                 out_init: np.ndarray | None = None) -> np.ndarray:
        din = DevArray(np.ascontiguousarray(vin, dtype=self.dtype))
        init = (np.zeros(out_n, self.dtype) if out_init is None
                else np.ascontiguousarray(out_init, dtype=self.dtype))
        dout = DevArray(init)
        check(lib.sc23ad_amg_level_op(self.h, level, op, din.p, dout.p))
        return dout.get()

    def default_frac_bits(self, b: np.ndarray, nrm_b: float) -> int:
        bits = C.c_int()
        check(lib.sc23ad_default_frac_bits(self.h, ptr(np.ascontiguousarray(b)),
                                           nrm_b, C.byref(bits)))
        return bits.value

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.sc23ad_amg_destroy(self.h)
            self.h = C.c_void_p()


class Solver:
    """Restarted(FGMRES(m)) on the GPU for solver 0/1/2."""

    def __init__(self, precision: int, a: CsrMat, amg: Amg, m: int = 30,
                 eps: float = 1e-10, bit: int = 29, max_restarts: int = 0):
        h = C.c_void_p()
        check(lib.sc23ad_solver_create(precision, a.h, amg.h, m, eps, bit,
                                       max_restarts, C.byref(h)))
        self.h = h
        self.amg = amg
        self.n = a.n

    def set_sequential_reduction(self, on: bool) -> None:
        check(lib.sc23ad_solver_set_reduction(self.h, 1 if on else 0))

    def set_stall(self, steps: int) -> None:
        """Stop (status STAGNATION) once the implicit residual has not fallen
        for `steps` consecutive Arnoldi steps (0: never, the reference's
        behaviour)."""
        check(lib.sc23ad_solver_set_stall(self.h, steps))

    def trace(self, cap: int) -> None:
        check(lib.sc23ad_solver_trace(self.h, None, cap, None))

    def read_trace(self, cap: int) -> np.ndarray:
        buf = np.zeros(max(cap, 1), np.uint64)
        n = C.c_int()
        check(lib.sc23ad_solver_trace(self.h, ptr(buf), cap, C.byref(n)))
        return buf[:n.value]

    def time_spmv(self, on: bool) -> None:
        check(lib.sc23ad_solver_time_spmv(self.h, 1 if on else 0))

    def history(self) -> np.ndarray:
        n = C.c_int()
        check(lib.sc23ad_solver_history(self.h, None, 0, C.byref(n)))
        buf = np.zeros(max(n.value, 1))
        check(lib.sc23ad_solver_history(self.h, ptr(buf), n.value, C.byref(n)))
        return buf[:n.value]

    def spmv_stats(self):
        n, ms = C.c_int(), C.c_double()
        check(lib.sc23ad_solver_spmv_stats(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def trsv_stats(self):
        """(launches, mean ms) of the fine-level LD solves timed since
        time_spmv(True)."""
        n, ms = C.c_int(), C.c_double()
        check(lib.sc23ad_solver_trsv_stats(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def solve(self, b: np.ndarray, nrm_b: float, x0: np.ndarray | None = None):
        b = _fine_vector(b, self.n, "b")
        x = np.zeros(self.n) if x0 is None else _fine_vector(x0, self.n, "x0").copy()
        cv = Converge()
        check(lib.sc23ad_solver_solve(self.h, ptr(b), nrm_b, ptr(x), 0, C.byref(cv)))
        return x, cv

    def solve_device(self, b: DevArray, nrm_b: float, x: DevArray) -> Converge:
        cv = Converge()
        check(lib.sc23ad_solver_solve(self.h, b.p, nrm_b, x.p, 1, C.byref(cv)))
        return cv

    def relres(self, b: np.ndarray, nrm_b: float, x: np.ndarray) -> float:
        out = C.c_double()
        check(lib.sc23ad_solver_relres(self.h, ptr(_fine_vector(b, self.n, "b")), nrm_b,
                                       ptr(_fine_vector(x, self.n, "x")), 0, C.byref(out)))
        return out.value

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.sc23ad_solver_destroy(self.h)
            self.h = C.c_void_p()


class Mlp:
    """The learned-parameter network (parameters on the GPU)."""

    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        check(lib.sc23ad_mlp_create(seed, C.byref(h)))
        self.h = h
        n = C.c_int()
        check(lib.sc23ad_mlp_nparams(C.byref(n)))
        self.nparams = n.value

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, np.float32)
        check(lib.sc23ad_mlp_params(self.h, ptr(p), None))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(lib.sc23ad_mlp_params(self.h, None, ptr(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        y = np.zeros((x.shape[0], 2), np.float32)
        check(lib.sc23ad_mlp_forward(self.h, ptr(x), x.shape[0], ptr(y)))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_float()
        check(lib.sc23ad_mlp_grad(self.h, ptr(x), ptr(t), x.shape[0], C.byref(loss), ptr(g)))
        return loss.value, g

    def train(self, x, t, steps: int, lr: float = 1e-2) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        loss = C.c_float()
        check(lib.sc23ad_mlp_train(self.h, ptr(x), ptr(t), x.shape[0], steps, lr,
                                   C.byref(loss)))
        return loss.value

    def levels(self, hier: "Hierarchy"):
        """(omega per level, vector bits per level incl. the coarsest, the
        first 0) predicted for hier's levels."""
        nl = hier.info()[0]
        om = np.zeros(max(nl, 1))
        vb = np.zeros(nl + 1, np.int32)
        n = C.c_int()
        check(lib.sc23ad_learned_levels(self.h, hier.h, ptr(om), ptr(vb), nl, C.byref(n)))
        return om[:n.value], vb[:n.value + 1]

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.sc23ad_mlp_destroy(self.h)
            self.h = C.c_void_p()


def solve_batch(solvers, bs, nrm_bs, xs=None, threads: int = 0):
    """Solves solver k for right-hand side bs[k] (host arrays), all at once:
    each solver on its own stream, `threads` host threads (0: up to 16).
    Returns (xs, converge records)."""
    k = len(solvers)
    if len(bs) != k or len(nrm_bs) != k or (xs is not None and len(xs) != k):
        raise ValueError("solve_batch: one b, nrm_b (and x) per solver")
    bs = [_fine_vector(b, s.n, "b") for s, b in zip(solvers, bs)]
    # the library reads and writes solver.n doubles of every x: work on
    # float64 contiguous copies and hand the results back in the caller's arrays
    xw = [np.zeros(s.n) if xs is None else _fine_vector(x, s.n, "x").copy()
          for s, x in zip(solvers, xs if xs is not None else [None] * k)]
    hs = (C.c_void_p * k)(*[s.h.value for s in solvers])
    bp = (C.c_void_p * k)(*[b.ctypes.data for b in bs])
    xp = (C.c_void_p * k)(*[x.ctypes.data for x in xw])
    nb = (C.c_double * k)(*nrm_bs)
    out = (Converge * k)()
    check(lib.sc23ad_solve_batch(k, hs, bp, nb, xp, 0, threads, out))
    if xs is None:
        return xw, list(out)
    for dst, src in zip(xs, xw):
        dst[...] = src
    return xs, list(out)


def _fine_vector(v, n: int, what: str) -> np.ndarray:
    """v as a contiguous float64 array of length n (the library reads exactly
    n doubles), or ValueError."""
    a = np.ascontiguousarray(v, dtype=np.float64)
    if a.ndim != 1 or a.shape[0] != n:
        raise ValueError(f"{what}: expected {n} values, got shape {a.shape}")
    return a


def to_block(cols: np.ndarray) -> np.ndarray:
    """(k, n) int32 vectors -> the block layout of the library (element (i, r)
    at ((r // 4) * n + i) * 4 + r % 4; missing right-hand sides are zero)."""
    cols = np.asarray(cols, dtype=np.int32)
    k, n = cols.shape
    g = (k + 3) // 4
    pad = np.zeros((4 * g, n), np.int32)
    pad[:k] = cols
    return np.ascontiguousarray(pad.reshape(g, 4, n).transpose(0, 2, 1)).reshape(-1)


def from_block(blk: np.ndarray, k: int, n: int) -> np.ndarray:
    g = (k + 3) // 4
    return np.asarray(blk).reshape(g, n, 4).transpose(0, 2, 1).reshape(4 * g, n)[:k].copy()


def precondition_block(amg: "Amg", vins: np.ndarray) -> np.ndarray:
    """One integer V-cycle on each row of vins ((k, n) int32), in one block
    call; returns (k, n)."""
    vins = np.asarray(vins, dtype=np.int32)
    k, n = vins.shape
    blk = to_block(vins)
    out = np.zeros_like(blk)
    check(lib.sc23ad_amg_precondition_block(amg.h, k, ptr(blk), ptr(out), 0))
    return from_block(out, k, n)


def time_vcycles_block(amg: "Amg", vins: np.ndarray, reps: int = 20) -> float:
    """Mean milliseconds of one block V-cycle on the rows of vins ((k, n)
    int32), on device vectors."""
    vins = np.asarray(vins, dtype=np.int32)
    k, n = vins.shape
    blk = to_block(vins)
    din = DevArray(blk)
    dout = DevArray(np.zeros_like(blk))
    ms = C.c_double()
    check(lib.sc23ad_amg_time_vcycles_block(amg.h, k, reps, din.p, dout.p, C.byref(ms)))
    return ms.value


def block_native(amg: "Amg") -> bool:
    v = C.c_int()
    check(lib.sc23ad_amg_block_native(amg.h, C.byref(v)))
    return bool(v.value)


class BlockSolver:
    """k integer Restarted(FGMRESflex(m)) solves on one hierarchy in lockstep
    (one launch per V-cycle / SpMV / Gram-Schmidt pass for all of them)."""

    def __init__(self, a: CsrMat, amg: Amg, bits, m: int = 30, eps: float = 1e-10,
                 max_restarts: int = 0):
        bits = np.ascontiguousarray(bits, dtype=np.int32)
        h = C.c_void_p()
        check(lib.sc23ad_block_solver_create(a.h, amg.h, len(bits), m, eps, ptr(bits),
                                             max_restarts, C.byref(h)))
        self.h = h
        self.amg = amg
        self.n = a.n
        self.k = len(bits)

    def solve(self, bs, nrm_bs, x0=None):
        """bs: k right-hand sides (sequence or (k, n) array); returns ((k, n)
        x, list of Converge)."""
        b = np.ascontiguousarray(np.stack([_fine_vector(v, self.n, "b") for v in bs]))
        if b.shape[0] != self.k or len(nrm_bs) != self.k:
            raise ValueError(f"solve: expected {self.k} right-hand sides")
        x = (np.zeros((self.k, self.n)) if x0 is None
             else np.ascontiguousarray(np.asarray(x0, dtype=np.float64).reshape(self.k, self.n)).copy())
        nb = np.ascontiguousarray(nrm_bs, dtype=np.float64)
        out = (Converge * self.k)()
        check(lib.sc23ad_block_solver_solve(self.h, ptr(b), ptr(nb), ptr(x), 0, out))
        return x, list(out)

    def solve_device(self, b: DevArray, nrm_bs, x: DevArray):
        """b, x: device arrays of k * n FP64 values (right-hand side r at r * n)."""
        nb = np.ascontiguousarray(nrm_bs, dtype=np.float64)
        if len(nb) != self.k or b.nbytes < 8 * self.k * self.n or x.nbytes < 8 * self.k * self.n:
            raise ValueError("solve_device: sizes")
        out = (Converge * self.k)()
        check(lib.sc23ad_block_solver_solve(self.h, b.p, ptr(nb), x.p, 1, out))
        return list(out)

    def set_stall(self, steps: int) -> None:
        """Stop a right-hand side (status STAGNATION) whose implicit residual
        has not fallen for `steps` consecutive Arnoldi steps (0: never)."""
        check(lib.sc23ad_block_solver_set_stall(self.h, steps))

    def trace(self, cap: int) -> None:
        check(lib.sc23ad_block_solver_trace(self.h, 0, None, cap, None))

    def read_trace(self, rhs: int, cap: int) -> np.ndarray:
        buf = np.zeros(max(cap, 1), np.uint64)
        n = C.c_int()
        check(lib.sc23ad_block_solver_trace(self.h, rhs, ptr(buf), cap, C.byref(n)))
        return buf[:n.value]

    def history(self, rhs: int) -> np.ndarray:
        n = C.c_int()
        check(lib.sc23ad_block_solver_history(self.h, rhs, None, 0, C.byref(n)))
        buf = np.zeros(max(n.value, 1))
        check(lib.sc23ad_block_solver_history(self.h, rhs, ptr(buf), n.value, C.byref(n)))
        return buf[:n.value]

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            lib.sc23ad_block_solver_destroy(self.h)
            self.h = C.c_void_p()


def block_solve_batch(solvers, bss, nrm_bss, threads: int = 0):
    """Block solver k solves right-hand sides bss[k] (host arrays); all
    solvers at once, each on its own stream.  Returns (xs, converge lists)."""
    cnt = len(solvers)
    if len(bss) != cnt or len(nrm_bss) != cnt:
        raise ValueError("block_solve_batch: one set of right-hand sides per solver")
    bl, xl, nl, ol = [], [], [], []
    for s_, bs, nbs in zip(solvers, bss, nrm_bss):
        b = np.ascontiguousarray(np.stack([_fine_vector(v, s_.n, "b") for v in bs]))
        if b.shape[0] != s_.k or len(nbs) != s_.k:
            raise ValueError("block_solve_batch: right-hand side count differs from the solver's")
        bl.append(b)
        xl.append(np.zeros((s_.k, s_.n)))
        nl.append(np.ascontiguousarray(nbs, dtype=np.float64))
        ol.append((Converge * s_.k)())
    hs = (C.c_void_p * cnt)(*[s_.h.value for s_ in solvers])
    bp = (C.c_void_p * cnt)(*[b.ctypes.data for b in bl])
    np_ = (C.c_void_p * cnt)(*[v.ctypes.data for v in nl])
    xp = (C.c_void_p * cnt)(*[x.ctypes.data for x in xl])
    op = (C.c_void_p * cnt)(*[C.addressof(o) for o in ol])
    check(lib.sc23ad_block_solve_batch(cnt, hs, bp, np_, xp, 0, threads, op))
    return xl, [list(o) for o in ol]


def spmv_int(rptr, cind, val, bit, x) -> np.ndarray:
    """Raw device integer SpMV on host arrays (copied in and out)."""
    n = len(rptr) - 1
    dr, dc, dv = DevArray(np.asarray(rptr, np.int32)), DevArray(np.asarray(cind, np.int32)), \
        DevArray(np.asarray(val, np.int32))
    dx = DevArray(np.asarray(x, np.int32))
    dy = DevArray.empty((n,), np.int32)
    check(lib.sc23ad_spmv_int(n, dr.p, dc.p, dv.p, bit, dx.p, dy.p))
    return dy.get()
