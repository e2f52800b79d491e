"""The learned-parameter network on the GPU against its CPU oracle
(oracle/mlp_oracle.py): forward pass, loss and gradient, training, and the
learned solve's fallback to the reference's defaults."""
from __future__ import annotations

import numpy as np
import pytest

import mlp_oracle as MO
import sc23ad as S
from sc23ad import learned as LR

pytestmark = pytest.mark.gpu


def test_forward_and_gradient_match_oracle():
    m = S.Mlp(seed=5)
    p = m.params()
    assert p.size == MO.NPARAMS
    rng = np.random.default_rng(1)
    x = rng.random((300, 8)).astype(np.float32)
    t = rng.random((300, 2)).astype(np.float32)
    y = m.forward(x)
    yo, _ = MO.forward(p, x)
    np.testing.assert_allclose(y, yo, rtol=1e-5, atol=1e-6)
    loss, g = m.grad(x, t)
    lo, go = MO.loss_grad(p, x, t)
    assert abs(loss - lo) <= 1e-5 * lo
    np.testing.assert_allclose(g, go, rtol=1e-3, atol=1e-6 * np.abs(go).max())


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(2)
    x = rng.random((512, 8)).astype(np.float32)
    t = np.stack([0.2 + 0.6 * x[:, 2], 0.9 - 0.5 * x[:, 0] * x[:, 1]], 1).astype(np.float32)
    m = S.Mlp(seed=3)
    l0, _ = m.grad(x, t)
    l1 = m.train(x, t, steps=400, lr=1e-2)
    assert l1 < 0.1 * l0
    # the loss reported by training is the oracle's loss at the new parameters
    assert abs(l1 - MO.loss_grad(m.params(), x, t)[0]) <= 1e-4 * max(l1, 1e-6)


def _forced(omega_y, bits_y):
    """Parameters whose outputs are constant: y = (omega_y, bits_y)."""
    p = np.zeros(MO.NPARAMS, np.float32)
    logit = lambda v: np.log(v / (1 - v))
    p[-2], p[-1] = logit(omega_y), logit(bits_y)
    m = S.Mlp(0)
    m.set_params(p)
    return m


def test_learned_levels_follow_the_network():
    sy = S.make_system(S.GEN_DIFFUSION27, (20, 5, 1.0, 1.0, 1e-3, 3), 3)
    hi, hd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    m = _forced((1.1 - 0.6) / 1.3, (24 - 8) / 22)
    om, bits = m.levels(hd)
    assert len(om) == hi.info()[0]
    np.testing.assert_allclose(om, 1.1, rtol=1e-6)
    assert np.all(bits == 24)


def test_failing_prediction_falls_back_to_defaults():
    sy = S.make_system(S.GEN_DIFFUSION27, (20, 5, 1.0, 1.0, 1e-3, 3), 3)
    base = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    # caps of 8 fractional bits make the smoother useless: the learned solve
    # fails and the defaults take over
    m = _forced((1.0 - 0.6) / 1.3, 1e-6)
    res = LR.learned_solve(sy.a, sy.b, sy.nrm_b, m, S.amg_param(S.BLOCK_SOR), base=base)
    bit = S.Amg(base[1]).default_frac_bits(sy.b, sy.nrm_b)
    x0, cv0 = S.Solver(S.INT, sy.a, S.Amg(base[0]), m=30, eps=1e-8, bit=bit).solve(sy.b, sy.nrm_b)
    assert np.all(res.learned.bits == 8)
    if res.fallback_used:
        assert np.array_equal(res.x.view(np.uint64), x0.view(np.uint64))
        assert res.converge.iter == cv0.iter
    else:
        assert res.converge.is_converged
