"""Host side of the product (CPU): the C ABI surface, the generators, the
MatrixMarket reader and error reporting.  No device calls."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import sc23ad as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    with open(os.path.join(ROOT, "include", "sc23ad.h")) as f:
        text = f.read()
    names = sorted(set(re.findall(r"\b(sc23ad_[a-z0-9_]+)\s*\(", text)))
    assert len(names) > 30
    lib = C.CDLL(S.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert S.lib.sc23ad_version().decode().startswith("sc23ad")


def test_hip_kernels_are_in_the_library():
    """The library carries gfx950 device code (a HIP fat binary section)."""
    with open(S.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"hipv4-amdgcn-amd-amdhsa--gfx950" in blob


# ---- generators ---------------------------------------------------------------

def _poisson2d(n):
    t = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(n, n))
    i = sp.identity(n)
    m = (sp.kron(i, t) + sp.kron(t, i)).tocsr()
    m.eliminate_zeros()
    return m


def _laplace3d(n):
    t = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(n, n))
    i = sp.identity(n)
    m = (sp.kron(sp.kron(i, i), t) + sp.kron(sp.kron(i, t), i)
         + sp.kron(sp.kron(t, i), i)).tocsr()
    m.eliminate_zeros()
    return m


@pytest.mark.parametrize("kind,n,ref", [(S.GEN_POISSON2D, 17, _poisson2d),
                                        (S.GEN_LAPLACE3D, 7, _laplace3d)])
def test_stencil_generators(kind, n, ref):
    a = S.CsrMat.generate(kind, n)
    rptr, cind, val = a.arrays()
    want = ref(n)
    want.sort_indices()
    assert np.array_equal(rptr, want.indptr) and np.array_equal(cind, want.indices)
    assert np.array_equal(val, want.data)


def test_laplace3d_128_size():
    a = S.CsrMat.generate(S.GEN_LAPLACE3D, 128)
    assert (a.n, a.nnz) == (2097152, 7 * 2097152 - 6 * 128 * 128)


def _mt19937_64(seed):
    """std::mt19937_64 (C++ [rand.predef]) in Python integers."""
    w, n, m, r = 64, 312, 156, 31
    a, u, d = 0xB5026F5AA96619E9, 29, 0x5555555555555555
    s, b, t, c, l, f = 17, 0x71D67FFFEDA60000, 37, 0xFFF7EEE000000000, 43, 6364136223846793005
    mask = (1 << w) - 1
    mt = [seed & mask]
    for i in range(1, n):
        mt.append((f * (mt[-1] ^ (mt[-1] >> (w - 2))) + i) & mask)
    idx = n
    lower = (1 << r) - 1
    upper = mask & ~lower
    while True:
        if idx >= n:
            for i in range(n):
                x = (mt[i] & upper) | (mt[(i + 1) % n] & lower)
                xa = x >> 1
                if x & 1:
                    xa ^= a
                mt[i] = mt[(i + m) % n] ^ xa
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= (y >> u) & d
        y ^= (y << s) & b
        y ^= (y << t) & c
        y ^= y >> l
        yield y & mask


def test_mt19937_64_restatement_known_answer():
    g = _mt19937_64(5489)
    for _ in range(9999):
        next(g)
    assert next(g) == 9981545732273789042  # required by the C++ standard


def test_xstar_is_uniform_map_of_mt19937_64():
    for seed in (0, 1, 4):
        g = _mt19937_64(seed)
        want = np.array([2.0 * ((next(g) >> 11) * 2.0**-53) - 1.0 for _ in range(1000)])
        assert np.array_equal(S.xstar(1000, seed), want)


def test_system_is_row_scaled():
    sy = S.make_system(S.GEN_POISSON2D, (20,), 0)
    rptr, cind, val = sy.a.arrays()
    raw = S.CsrMat.generate(S.GEN_POISSON2D, 20)
    np.testing.assert_array_equal(raw.matvec(sy.xstar), sy.b_unscaled)
    # LMax1Scaling (matrix/csrmat.hpp:70-80): row max |a_ij| becomes 1, b_i /= max
    for i in range(sy.a.n):
        row = val[rptr[i]:rptr[i + 1]]
        assert np.max(np.abs(row)) == 1.0
    np.testing.assert_array_equal(sy.b, sy.b_unscaled / 4.0)


@pytest.mark.parametrize("kind,params", [
    (S.GEN_CONVDIFF2D, (30, 50.0, 1.0)),
    (S.GEN_DIFFUSION27, (10, 4, 1.0, 1.0, 1e-3, 2)),
    (S.GEN_INCLUSIONS3D, (12, 4, 2.0, 3.0, 1e6, 4)),
])
def test_random_generators_are_deterministic_and_well_formed(kind, params):
    a1, a2 = S.CsrMat.generate(kind, *params), S.CsrMat.generate(kind, *params)
    for x, y in zip(a1.arrays(), a2.arrays()):
        assert np.array_equal(x, y)
    rptr, cind, val = a1.arrays()
    m = sp.csr_matrix((val, cind, rptr), shape=(a1.n, a1.n))
    d = m.diagonal()
    assert np.all(d > 0)
    # weakly diagonally dominant M-matrix rows (Dirichlet boundary: strict somewhere)
    off = np.asarray(abs(m).sum(axis=1)).ravel() - d
    assert np.all(off <= d * (1 + 1e-12))
    for i in range(a1.n):
        cols = cind[rptr[i]:rptr[i + 1]]
        assert np.all(np.diff(cols) > 0)


# ---- MatrixMarket -------------------------------------------------------------

def test_read_mm_symmetric_lower(tmp_path):
    full = _poisson2d(5).tocoo()
    lower = [(r, c, v) for r, c, v in zip(full.row, full.col, full.data) if r >= c]
    rng = np.random.default_rng(0)
    rng.shuffle(lower)
    path = tmp_path / "sym.mtx"
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% comment\n")
        f.write(f"25 25 {len(lower) + 1}\n")
        for r, c, v in lower:
            f.write(f"{r + 1} {c + 1} {v:.17g}\n")
        f.write("3 1 0\n")  # explicit zero: dropped
    a = S.CsrMat.read_mm(str(path))
    rptr, cind, val = a.arrays()
    want = _poisson2d(5)
    want.sort_indices()
    assert np.array_equal(rptr, want.indptr) and np.array_equal(cind, want.indices)
    assert np.array_equal(val, want.data)


def test_read_mm_general_integer(tmp_path):
    rng = np.random.default_rng(1)
    dense = np.where(rng.random((9, 9)) < 0.3, rng.integers(-5, 6, (9, 9)), 0)
    np.fill_diagonal(dense, 10)
    ent = [(r, c, dense[r, c]) for r in range(9) for c in range(9) if dense[r, c] != 0]
    rng.shuffle(ent)
    path = tmp_path / "gen.mtx"
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer general\n")
        f.write(f"9 9 {len(ent)}\n")
        for r, c, v in ent:
            f.write(f"{r + 1} {c + 1} {v}\n")
    a = S.CsrMat.read_mm(str(path))
    want = sp.csr_matrix(dense.astype(float))
    want.sort_indices()
    rptr, cind, val = a.arrays()
    assert np.array_equal(rptr, want.indptr) and np.array_equal(cind, want.indices)
    assert np.array_equal(val, want.data)


# ---- errors -------------------------------------------------------------------

def test_errors_are_status_codes(tmp_path):
    with pytest.raises(S.Sc23adError) as e:
        S.CsrMat.read_mm(str(tmp_path / "missing.mtx"))
    assert e.value.code == S.IO_ERROR
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(S.Sc23adError) as e:
        S.CsrMat.read_mm(str(bad))
    assert e.value.code in (S.IO_ERROR, S.BAD_PARAM)
    with pytest.raises(S.Sc23adError) as e:
        S.CsrMat.from_arrays(2, np.array([0, 3, 2], np.int32), np.array([0, 1], np.int32),
                             np.array([1.0, 2.0]))
    assert e.value.code == S.BAD_PARAM
    a = S.CsrMat.generate(S.GEN_POISSON2D, 30)
    for theta in (0.0, -0.5, 1.5):  # helper_multigrid.hpp:39-42 exits for theta outside (0, 1]
        with pytest.raises(S.Sc23adError) as e:
            S.Hierarchy(a, S.INT, S.amg_param(S.SOR, theta=theta))
        assert e.value.code == S.BAD_PARAM
    with pytest.raises(S.Sc23adError) as e:
        S.Hierarchy(a, 7, S.amg_param(S.SOR))
    assert e.value.code == S.BAD_PARAM


def test_device_setup_needs_a_gpu():
    if S.device_count() > 0:
        pytest.skip("a GPU is visible: the device setup is checked by the gpu tests")
    a = S.CsrMat.generate(S.GEN_POISSON2D, 30)
    with pytest.raises(S.Sc23adError) as e:
        S.Hierarchy.build_multi(a, [S.INT], S.amg_param(S.SOR), device_setup=True)
    assert e.value.code == S.NO_DEVICE


def test_bench_traffic_from_committed_counter_summaries():
    """bench.py's `traffic` comes from the committed rocprofv3 summaries:
    FETCH_SIZE and WRITE_SIZE per dispatch, each times its copy-kernel
    calibration factor, in bytes; it must equal the traffic the summary
    states (the first kernel of each file) and stay below the algorithmic
    bytes where the summary says so."""
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for name in ("spmv_int_level0_laplace3d_128_pmc.txt",
                 "sptrsv_int_level0_laplace3d_128_pmc.txt"):
        path = os.path.join(root, "profiles", name)
        txt = open(path).read()
        stated = float(re.search(r"^traffic_bytes_per_dispatch:\s*(\S+)", txt, re.M).group(1))
        got = bench.read_traffic(path)
        assert got is not None and abs(got - stated) <= 1e-6 * stated + 512, (name, got, stated)
    assert bench.read_traffic(os.path.join(root, "profiles", "missing.txt")) is None
