"""Learned-parameter plumbing on the host (CPU): per-level retuning of a
hierarchy from its kept coarsening, bit caps and their saturation, level
features against a numpy restatement, and the network oracle's gradient
against finite differences."""
from __future__ import annotations

import os

import numpy as np
import pytest

import mlp_oracle as MO
import sc23ad as S
from helpers import export, file_sha, read_hier


@pytest.fixture(scope="module")
def base():
    sy = S.make_system(S.GEN_DIFFUSION27, (20, 5, 1.0, 1.0, 1e-3, 3), 3)
    hi, hd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    return sy, hi, hd


def _sha(h):
    p = export(h)
    try:
        return file_sha(p)
    finally:
        os.unlink(p)


def test_retune_defaults_reproduce_base(base):
    sy, hi, hd = base
    d = hi.info()[0]
    assert _sha(hi.retune([1.0] * d, [0] * d, S.INT)) == _sha(hi)
    assert _sha(hd.retune([1.0] * d, None, S.FP64)) == _sha(hd)


def _same_matrices(h1, h2):
    a, b = read_hier(export(h1)), read_hier(export(h2))
    assert a.depth == b.depth
    for la, lb in zip(a.levels + [{"lc": a.lc, "uc": a.uc}], b.levels + [{"lc": b.lc, "uc": b.uc}]):
        for k in la:
            for f in ("rptr", "cind", "val"):
                assert np.array_equal(getattr(la[k], f), getattr(lb[k], f)), k
            assert (la[k].bit, la[k].dbit) == (lb[k].bit, lb[k].dbit)


def test_retune_uniform_omega_equals_build(base):
    """Same matrices as a setup run with that omega (the header keeps the
    base's scalar omega)."""
    sy, hi, hd = base
    d = hi.info()[0]
    built = S.Hierarchy(sy.a, S.INT, S.amg_param(S.BLOCK_SOR, omega=1.2))
    _same_matrices(hi.retune([1.2] * d, None, S.INT), built)


def test_per_level_omega_only_changes_its_level(base):
    sy, hi, hd = base
    d = hi.info()[0]
    om = [1.0] * d
    om[1] = 1.3
    a = read_hier(export(hd))
    b = read_hier(export(hd.retune(om, None, S.FP64)))
    for l in range(d):
        same = all(np.array_equal(getattr(a.levels[l][k], "val"), getattr(b.levels[l][k], "val"))
                   for k in ("LD", "U"))
        assert same == (l != 1)


def test_bit_caps_and_saturation(base):
    sy, hi, hd = base
    d = hi.info()[0]
    capped = read_hier(export(hi.retune(None, [20] * d, S.INT)))
    ref = read_hier(export(hi))
    assert hi.retune(None, [20] * d, S.INT).saturated() == 0
    for l in range(d):
        for k in ("U", "LD") + (("A",) if l > 0 else ()):
            assert capped.levels[l][k].bit == min(20, ref.levels[l][k].bit)
        assert capped.levels[l]["LD"].dbit == ref.levels[l]["LD"].dbit
        for k in ("R", "P"):
            assert capped.levels[l][k].bit == ref.levels[l][k].bit
    # a cap above the rule's maximum is not applied and is reported
    sat = hi.retune(None, [31] * d, S.INT)
    assert sat.saturated() == d
    assert _sha(sat) == _sha(hi)


def test_features_against_numpy(base):
    import scipy.sparse as sp
    sy, hi, hd = base
    f = hd.features()
    h = read_hier(export(hd))
    d = h.depth
    assert f.shape == (d, 8)
    for l in range(d):
        A = h.levels[l]["A"]
        n = A.n
        M = sp.csr_matrix((A.val, A.cind, A.rptr), shape=(n, n))
        diag = np.abs(M.diagonal())
        rowsum = np.asarray(abs(M).sum(axis=1)).ravel()
        O = (M - sp.diags(M.diagonal())).tocsr()
        O.eliminate_zeros()
        offmax = np.asarray(abs(O).max(axis=1).todense()).ravel()
        coo = M.tocoo()
        offd = coo.row != coo.col
        MT = M.T.tocsr()
        vt = np.asarray(MT[coo.row[offd], coo.col[offd]]).ravel()
        asym = np.abs(coo.data[offd] - vt).sum()
        want = [np.log2(n) / 24, A.val.size / n / 32, np.mean(diag / rowsum),
                np.min(diag / rowsum), np.mean(offmax / diag),
                np.count_nonzero(coo.data[offd] > 0) / np.count_nonzero(offd),
                asym / rowsum.sum(), l / d]
        np.testing.assert_allclose(f[l], want, rtol=2e-6, atol=1e-7)


def test_mlp_oracle_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, MO.NPARAMS)
    x = rng.random((7, 8))
    t = rng.random((7, 2))
    loss, g = MO.loss_grad(p, x, t)
    for k in rng.choice(MO.NPARAMS, 40, replace=False):
        e = np.zeros_like(p)
        e[k] = 1e-6
        fd = (MO.loss_grad(p + e, x, t)[0] - MO.loss_grad(p - e, x, t)[0]) / 2e-6
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd)


def _train_module():
    import importlib.util
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location(
        "train", os.path.join(here, "sc23-ad_amd", "learn", "train.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_training_range_guard(tmp_path):
    """A level is in range when every feature lies within the training
    [min, max] widened by a tenth of its span on either side; without a
    range file every level is."""
    from sc23ad import learned as LR
    path = str(tmp_path / "range.npy")
    lo, hi = np.array([0.0, 10.0]), np.array([1.0, 20.0])
    np.save(path, np.stack([lo, hi]))
    feats = np.array([[0.5, 15.0],     # inside
                      [1.09, 10.0],    # inside the widened band of feature 0
                      [-0.09, 21.0],   # both inside their bands
                      [1.11, 15.0],    # beyond the band of feature 0
                      [0.5, 8.9]])     # beyond the band of feature 1
    assert LR.in_range(feats, path).tolist() == [True, True, True, False, False]
    assert LR.in_range(feats, str(tmp_path / "missing.npy")).all()


def test_margin_targets_keep_defaults_unless_the_gain_is_real():
    """A swept value becomes a level's target only if it beats the defaults
    by at least max(2 iterations, the margin); unconverged variants (counted
    as 10^6 iterations) never do; targets are scaled like the network's
    outputs (omega 0.6..1.9 and vector bits -21..1 onto [0, 1])."""
    tr = _train_module()
    # instance 3: defaults 100 iterations, two levels
    rows = [(3, -1, 0, 0, 100),
            (3, 0, 1, 1.3, 97),     # gain 3 < 5 % of 100: not a target
            (3, 0, 1, 0.6, 10**6),  # unconverged
            (3, 0, 2, -1, 100),
            (3, 1, 1, 1.6, 90),     # gain 10: target
            (3, 1, 1, 1.9, 92),
            (3, 1, 2, -4, 94),      # gain 6: target
            (3, 1, 2, 1, 95)]
    # instance 5: defaults 20 iterations, one level; the margin is 2 iterations
    rows += [(5, -1, 0, 0, 20), (5, 0, 1, 0.8, 18), (5, 0, 2, 1, 19)]
    sweep = np.array(rows, np.float64)
    inst = np.array([3, 3, 5])
    t = tr.margin_targets(sweep, inst, 0.05)
    omega = 0.6 + 1.3 * t[:, 0]
    vbits = -21 + 22 * t[:, 1]
    np.testing.assert_allclose(omega, [1.0, 1.6, 0.8], atol=1e-6)
    np.testing.assert_allclose(vbits, [0, -4, 0], atol=1e-5)
    assert t.dtype == np.float32 and t.shape == (3, 2)
