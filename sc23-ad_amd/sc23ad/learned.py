"""Learned per-level smoother parameters (omega_l, bit cap b_l) with the
reference's defaults as the fallback.

The network (`Mlp`, HIP kernels) reads 8 features of every level's operator
and predicts omega_l in [0.6, 1.9] and a cap b_l in [8, 30] on the fractional
bits of A_l, U_l and LD_l's off-diagonal entries.  The integer kernels do not
change.  Fallbacks to the reference's rule (one omega, the CSRflex bits):
  * a level whose cap exceeds the flex rule's maximum (a saturated
    prediction) keeps the rule's bits;
  * a solve that overflows, stagnates or runs out of restarts is repeated
    with the default hierarchy.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import (FP64, INT, OK, Amg, Hierarchy, Mlp, Solver, amg_param)

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(os.path.dirname(HERE), "learn", "weights_diffusion27.npy")


def load_mlp(path: str = WEIGHTS, seed: int = 0) -> Mlp:
    m = Mlp(seed)
    m.set_params(np.load(path, allow_pickle=False))
    return m


@dataclass
class LearnedSetup:
    omega: np.ndarray        # per level
    bits: np.ndarray         # per-level caps as predicted
    saturated: int           # levels where the cap exceeded the rule's maximum
    hier_int: Hierarchy
    hier_fp64: Hierarchy


def learned_setup(base_int: Hierarchy, base_fp64: Hierarchy, mlp: Mlp) -> LearnedSetup:
    """Hierarchies with the network's per-level parameters, derived from the
    default ones' coarsening."""
    om, bi = mlp.levels(base_fp64)
    hi = base_int.retune(om, bi, INT)
    hd = base_fp64.retune(om, None, FP64)
    return LearnedSetup(om, bi, hi.saturated(), hi, hd)


@dataclass
class LearnedResult:
    x: np.ndarray
    converge: object
    fallback_used: bool
    learned: LearnedSetup
    bit: int


def learned_solve(a, b, nrm_b, mlp: Mlp, param=None, m: int = 30, eps: float = 1e-8,
                  base=None) -> LearnedResult:
    """Integer AMG-FGMRES(m) with learned per-level parameters; falls back
    to the default hierarchy when the learned solve does not converge
    cleanly.  `base` = (int, fp64) default hierarchies if already built."""
    param = param if param is not None else amg_param()
    if base is None:
        base = Hierarchy.build_multi(a, [INT, FP64], param)
    bi, bd = base
    ls = learned_setup(bi, bd, mlp)
    bit = Amg(ls.hier_fp64).default_frac_bits(b, nrm_b)
    sol = Solver(INT, a, Amg(ls.hier_int), m=m, eps=eps, bit=bit)
    x, cv = sol.solve(b, nrm_b)
    if cv.status == OK and cv.is_converged:
        return LearnedResult(x, cv, ls.saturated > 0, ls, bit)
    bit0 = Amg(bd).default_frac_bits(b, nrm_b)
    sol0 = Solver(INT, a, Amg(bi), m=m, eps=eps, bit=bit0)
    x0, cv0 = sol0.solve(b, nrm_b)
    return LearnedResult(x0, cv0, True, ls, bit0)
