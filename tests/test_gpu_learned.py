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
    m = _forced((1.1 - 0.6) / 1.3, (-3 + 21) / 22)
    om, vb = m.levels(hd)
    depth = hi.info()[0]
    assert len(om) == depth and len(vb) == depth + 1
    np.testing.assert_allclose(om, 1.1, rtol=1e-6)
    assert vb[0] == 0 and np.all(vb[1:] == -3)


def test_failing_prediction_falls_back_to_defaults():
    sy = S.make_system(S.GEN_DIFFUSION27, (20, 5, 1.0, 1.0, 1e-3, 3), 3)
    base = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    # 8 fractional bits on every coarse level (vector bits -21) make the
    # coarse corrections useless: the learned solve stalls and the defaults
    # take over
    m = _forced((1.0 - 0.6) / 1.3, 1e-6)
    res = LR.learned_solve(sy.a, sy.b, sy.nrm_b, m, S.amg_param(S.BLOCK_SOR), base=base,
                           stall_steps=10, guard=False)
    bit = S.Amg(base[1]).default_frac_bits(sy.b, sy.nrm_b)
    x0, cv0 = S.Solver(S.INT, sy.a, S.Amg(base[0]), m=30, eps=1e-8, bit=bit).solve(sy.b, sy.nrm_b)
    assert np.all(res.learned.vbits[1:] == -21)
    print(f"[fallback] used {res.fallback_used}, iterations {res.converge.iter}", flush=True)
    if res.fallback_used:
        assert np.array_equal(res.x.view(np.uint64), x0.view(np.uint64))
        assert res.converge.iter == cv0.iter
    else:
        assert res.converge.is_converged and res.relres <= 1e-8


def test_stall_stop_single_and_block():
    """A solve that stops improving ends with STAGNATION once its implicit
    residual has not fallen for the set number of steps (single and block
    solvers alike, the block's right-hand sides each as their single solve);
    without the stall limit the same solve runs on until the reference's own
    tests end it."""
    sy = S.make_system(S.GEN_DIFFUSION27, (20, 5, 1.0, 1.0, 1e-3, 3), 3)
    bi, bd = S.Hierarchy.build_multi(sy.a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
    bit = S.Amg(bd).default_frac_bits(sy.b, sy.nrm_b)
    amg = S.Amg(bi)
    # 1e-30 is out of reach: the residual levels off at the FP64 outer
    # refinement's floor
    sol = S.Solver(S.INT, sy.a, amg, m=30, eps=1e-30, bit=bit)
    sol.set_stall(5)
    x, cv = sol.solve(sy.b, sy.nrm_b)
    h = sol.history()
    print(f"[stall] status {cv.status} after {len(h)} steps", flush=True)
    assert cv.status == S.STAGNATION and not cv.is_converged
    blk = S.BlockSolver(sy.a, amg, [bit, bit, 24], m=30, eps=1e-30)
    blk.set_stall(5)
    xb, cvb = blk.solve([sy.b, sy.b, sy.b], [sy.nrm_b] * 3)
    for r in range(2):
        assert cvb[r].status == S.STAGNATION
        assert np.array_equal(blk.history(r), h)
        assert np.array_equal(xb[r].view(np.uint64), x.view(np.uint64))
    one = S.Solver(S.INT, sy.a, amg, m=30, eps=1e-30, bit=24)
    one.set_stall(5)
    x3, cv3 = one.solve(sy.b, sy.nrm_b)
    assert cvb[2].status == cv3.status and np.array_equal(blk.history(2), one.history())
    assert np.array_equal(xb[2].view(np.uint64), x3.view(np.uint64))
    sol.set_stall(0)
    x2, cv2 = sol.solve(sy.b, sy.nrm_b)
    print(f"[stall] without the limit: status {cv2.status} after {len(sol.history())} steps",
          flush=True)
    assert len(sol.history()) > len(h)


def test_learned_on_held_out_instances():
    """The committed network on instances of both training families that the
    sweep never saw (another seed): every learned answer (after the fallback
    where it ran) converges and meets rtol in the explicit residual; its
    steps against the defaults', a fallback's wasted steps counted in, are
    printed as the gain (a gain <= 1 records its absence)."""
    import importlib.util
    import os
    import time
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location(
        "make_dataset", os.path.join(here, "sc23-ad_amd", "learn", "make_dataset.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    mlp = LR.load_mlp()
    it_d = it_l = 0
    t_d = t_l = 0.0
    for k in range(8):
        a, b, nrm, desc = mk.instance(k, 97, 32, 48)
        base = S.Hierarchy.build_multi(a, [S.INT, S.FP64], S.amg_param(S.BLOCK_SOR))
        bit = S.Amg(base[1]).default_frac_bits(b, nrm)
        sol = S.Solver(S.INT, a, S.Amg(base[0]), m=30, eps=1e-8, bit=bit)
        sol.solve(b, nrm)
        t = time.perf_counter()
        x0, cv0 = sol.solve(b, nrm)
        td = time.perf_counter() - t
        t = time.perf_counter()
        res = LR.learned_solve(a, b, nrm, mlp, S.amg_param(S.BLOCK_SOR), base=base)
        tl = time.perf_counter() - t
        print(f"[held-out] {desc}: default {cv0.iter} it, learned {res.converge.iter} it "
              f"(fallback {res.fallback_used}, omega {np.round(res.learned.omega, 2).tolist()}, "
              f"vbits {res.learned.vbits.tolist()})", flush=True)
        assert res.converge.is_converged and res.relres <= 1e-8
        it_d += cv0.iter
        it_l += res.attempt_iter + (res.converge.iter if res.fallback_used else 0)
        t_d += cv0.solve_ms
        t_l += res.attempt_ms + (res.converge.solve_ms if res.fallback_used else 0.0)
    print(f"[held-out] steps default {it_d} learned {it_l} (fallbacks included): gain "
          f"{it_d / it_l:.3f}; device solve time {t_d:.1f} vs {t_l:.1f} ms", flush=True)
