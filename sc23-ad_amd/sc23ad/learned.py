"""Learned per-level parameters (smoother weight omega_l and the fractional
bits of every coarse level's vectors) with the reference's defaults as the
fallback.

The network (`Mlp`, HIP kernels) reads 8 features of every level's operator
and predicts omega_l in [0.6, 1.9] and the fractional bits of the next
coarser level's vectors relative to the solver's F, in [-21, 1]
(sc23ad_learned_levels).  The integer kernels do not change: the bits are
folded into the restriction / prolongation shifts
(sc23ad_hier_retune_levels).  Fallbacks to the reference's rule (one omega,
one F for every vector):
  * a level whose features lie outside the training set's range (each
    feature's [min, max] widened by a tenth of its span, RANGE) keeps the
    defaults: the network does not extrapolate;
  * a level whose bit step leaves the shifts' range is cut to it (a
    saturated prediction, counted);
  * a solve that overflows, stops improving for STALL_STEPS Arnoldi steps,
    stagnates or runs out of restarts is repeated with the default
    hierarchy; the explicit FP64 residual decides whether the result stands.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import (FP64, INT, OK, Amg, Hierarchy, Mlp, Solver, amg_param)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(os.path.dirname(HERE), "learn", "weights_levels.npy")
# per-feature [min, max] over the training levels (2 x 8), written by train.py
RANGE = os.path.join(os.path.dirname(HERE), "learn", "weights_levels_range.npy")
# Arnoldi steps without a fall of the implicit residual after which a
# learned solve gives up and the defaults take over.
STALL_STEPS = 30


def load_mlp(path: str = WEIGHTS, seed: int = 0) -> Mlp:
    m = Mlp(seed)
    m.set_params(np.load(path, allow_pickle=False))
    return m


@dataclass
class LearnedSetup:
    omega: np.ndarray        # per level
    vbits: np.ndarray        # per level incl. the coarsest (first 0)
    saturated: int           # bit steps cut to the shifts' range
    hier_int: Hierarchy
    hier_fp64: Hierarchy
    abstained: int = 0       # levels outside the training range (defaults kept)


def in_range(feats: np.ndarray, path: str = RANGE) -> np.ndarray:
    """Per level: whether every feature lies inside the training range
    widened by a tenth of its span (all True when no range file exists)."""
    if not os.path.exists(path):
        return np.ones(len(feats), bool)
    lo, hi = np.load(path, allow_pickle=False)
    pad = 0.1 * (hi - lo)
    return np.all((feats >= lo - pad) & (feats <= hi + pad), axis=1)


def learned_setup(base_int: Hierarchy, base_fp64: Hierarchy, mlp: Mlp,
                  guard: bool = True) -> LearnedSetup:
    """Hierarchies with the network's per-level parameters, derived from the
    default ones' coarsening; with guard, levels outside the training range
    keep the defaults (omega 1 and the next level's vector bits unchanged)."""
    om, vb = mlp.levels(base_fp64)
    ok = in_range(base_fp64.features()) if guard else np.ones(len(om), bool)
    om = np.where(ok, om, 1.0)
    vb = np.array(vb)
    vb[1:][~ok] = 0
    hi = base_int.retune_levels(om, vb, INT)
    hd = base_fp64.retune_levels(om, None, FP64)
    return LearnedSetup(om, vb, hi.saturated(), hi, hd, int(np.sum(~ok)))


@dataclass
class LearnedResult:
    x: np.ndarray
    converge: object
    fallback_used: bool
    learned: LearnedSetup
    bit: int
    relres: float
    # the learned attempt's own steps (all of them when the fallback ran:
    # the total cost is then attempt_iter + converge.iter)
    attempt_iter: int = 0
    attempt_ms: float = 0.0


def learned_solve(a, b, nrm_b, mlp: Mlp, param=None, m: int = 30, eps: float = 1e-8,
                  base=None, stall_steps: int = STALL_STEPS, guard: bool = True) -> LearnedResult:
    """Integer AMG-FGMRES(m) with learned per-level parameters; falls back
    to the default hierarchy when the learned solve does not converge
    cleanly (overflow, a stall of stall_steps steps, stagnation, budget) or
    its explicit residual misses eps.  `base` = (int, fp64) default
    hierarchies if already built."""
    param = param if param is not None else amg_param()
    if base is None:
        base = Hierarchy.build_multi(a, [INT, FP64], param)
    bi, bd = base
    ls = learned_setup(bi, bd, mlp, guard)
    bit = Amg(ls.hier_fp64).default_frac_bits(b, nrm_b)
    sol = Solver(INT, a, Amg(ls.hier_int), m=m, eps=eps, bit=bit)
    sol.set_stall(stall_steps)
    x, cv = sol.solve(b, nrm_b)
    rel = sol.relres(b, nrm_b, x)
    if cv.status == OK and cv.is_converged and rel <= eps:
        return LearnedResult(x, cv, False, ls, bit, rel, cv.iter, cv.solve_ms)
    bit0 = Amg(bd).default_frac_bits(b, nrm_b)
    sol0 = Solver(INT, a, Amg(bi), m=m, eps=eps, bit=bit0)
    x0, cv0 = sol0.solve(b, nrm_b)
    cv0.fallback_used = 1
    return LearnedResult(x0, cv0, True, ls, bit0, sol0.relres(b, nrm_b, x0),
                         len(sol.history()), cv.solve_ms)
