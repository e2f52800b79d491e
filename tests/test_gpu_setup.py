"""AMG setup with the interpolation weights and the Galerkin products on the
GPU (sc23ad_hier_build_multi_dev): every hierarchy file must be byte-identical
to the host setup's, which the CPU suite pins to the reference's own setup
(tests/golden/reference.json).  Cases: the golden systems plus larger ones
whose coarse levels have long rows (the hash tables of the row-merge product
see many collisions) and the amk = 0 substitution of the 1e6-contrast
inclusions."""
from __future__ import annotations

import hashlib
import os
import time

import pytest

import sc23ad as S
from helpers import export, system

pytestmark = pytest.mark.gpu


def _sha(h):
    path = export(h)
    try:
        with open(path, "rb") as f:
            return hashlib.sha256(f.read()).hexdigest()
    finally:
        os.unlink(path)


def _compare(a, smoother, label, golden_hashes=None):
    """Host setup against the device setup with the Galerkin products forced
    onto the GPU at every level (by default only large levels use it); where
    the reference's own hashes are known (golden_hashes: precision ->
    SHA-256 of the reference's hierarchy file) the device setup must match
    them too."""
    prm = S.amg_param(smoother)
    t = time.perf_counter()
    host = S.Hierarchy.build_multi(a, [S.INT, S.FP64, S.FP32], prm)
    th = time.perf_counter() - t
    os.environ["SC23AD_SETUP_GPU_PRODUCTS"] = "1"
    try:
        t = time.perf_counter()
        dev = S.Hierarchy.build_multi(a, [S.INT, S.FP64, S.FP32], prm, device_setup=True)
        td = time.perf_counter() - t
    finally:
        del os.environ["SC23AD_SETUP_GPU_PRODUCTS"]
    print(f"[setup] {label} smoother {smoother}: host {th:.2f} s, device {td:.2f} s", flush=True)
    for prec, h, d in zip((S.INT, S.FP64, S.FP32), host, dev):
        assert h.info()[0] == d.info()[0], label
        dsha = _sha(d)
        assert _sha(h) == dsha, label
        if golden_hashes is not None:
            assert dsha == golden_hashes[prec], (label, prec)


def test_golden_systems(golden):
    """Device setup against the reference's hierarchy files (written by the
    reference compiled in the container: tests/golden/reference.json)."""
    for c in golden["cases"]:
        sy = system(c)
        for smoother in (0, 1):
            _compare(sy.a, smoother, c["name"],
                     {prec: c["hier"][f"{prec}_{smoother}"] for prec in (0, 1, 2)})


@pytest.mark.parametrize("kind,params,seed", [
    (S.GEN_LAPLACE3D, (40,), 1),
    (S.GEN_DIFFUSION27, (24, 4, 1.0, 1.0, 1e-3, 2), 2),
    (S.GEN_INCLUSIONS3D, (40, 16, 3.0, 6.0, 1e6, 4), 4),
    (S.GEN_CONVDIFF2D, (150, 60.0, 0.7), 3),
], ids=["laplace3d_40", "diffusion27_24", "inclusions3d_40", "convdiff2d_150"])
def test_larger_systems(kind, params, seed):
    sy = S.make_system(kind, params, seed)
    _compare(sy.a, 1, f"kind {kind} {params}")
