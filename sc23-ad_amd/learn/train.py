"""Trains the learned-parameter network on the GPU (the library's own HIP
forward/backward kernels, full-batch Adam) from the solver's sweep data:

    python sc23-ad_amd/learn/train.py [--data ...npz] [--steps 3000] [--out ...npy]

Instances are split 80/20 into training and validation sets (by instance);
the script reports both losses and how often the predicted omega and vector
bits, rounded to the sweep grids, are the level's best, and writes the
parameters (float32) of the step with the lowest validation loss, and beside
them (<out>_range.npy) every feature's [min, max] over the training levels.

Targets come from the sweep table: a level's omega (vector bits) target is
the sweep value with the fewest iterations only if it beats the defaults by
at least max(2 iterations, 5 %); otherwise the default.  One-iteration wins
are mostly the fixed-point arithmetic's run-to-run luck, not a property of the
level that the features could predict.
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import sc23ad as S  # noqa: E402


OMEGA_LO, OMEGA_SPAN = 0.6, 1.3
VB_LO, VB_HI = -21, 1


def margin_targets(sweep, inst, margin):
    """Per-level (omega, vector bits) targets, scaled to [0, 1] like the
    network's outputs, rows in the order of the feature rows (instances in
    order, levels 0..depth-1).  sweep rows: (instance, level, kind, value,
    iterations) with kind 0 the defaults, 1 an omega value, 2 vector bits."""
    out = []
    for k in np.unique(inst):
        rows = sweep[sweep[:, 0] == k]
        base = rows[rows[:, 2] == 0][0, 4]
        need = base - max(2.0, margin * base)
        depth = int(np.sum(inst == k))
        for lv in range(depth):
            tgt = [1.0, 0.0]
            for kind in (1, 2):
                r = rows[(rows[:, 1] == lv) & (rows[:, 2] == kind)]
                if len(r) and r[:, 4].min() <= need:
                    tgt[kind - 1] = float(r[r[:, 4].argmin(), 3])
            out.append(((tgt[0] - OMEGA_LO) / OMEGA_SPAN, (tgt[1] - VB_LO) / (VB_HI - VB_LO)))
    return np.asarray(out, np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", default=os.path.join(HERE, "data", "levels_sweeps.npz"))
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "weights_levels.npy"))
    ap.add_argument("--margin", type=float, default=0.05,
                    help="relative gain a swept value needs to become a target")
    args = ap.parse_args()
    d = np.load(args.data, allow_pickle=False)
    X, inst = d["X"], d["instance"]
    T = margin_targets(d["sweep"], inst, args.margin)
    ids = np.unique(inst)
    rng = np.random.default_rng(args.seed)
    val_ids = set(rng.choice(ids, max(1, len(ids) // 5), replace=False).tolist())
    va = np.array([i in val_ids for i in inst])
    m = S.Mlp(args.seed)
    l0 = m.grad(X[~va], T[~va])[0]
    done = 0
    best = (m.grad(X[va], T[va])[0] if va.any() else l0, 0, m.params())
    while done < args.steps:
        k = min(250, args.steps - done)
        lt = m.train(X[~va], T[~va], steps=k, lr=args.lr)
        done += k
        lv = m.grad(X[va], T[va])[0] if va.any() else lt
        if lv < best[0]:
            best = (lv, done, m.params())
        print(f"step {done}: train loss {lt:.5f} validation loss {lv:.5f} (initial {l0:.5f})",
              flush=True)
    print(f"kept the parameters of step {best[1]} (validation loss {best[0]:.5f})", flush=True)
    m.set_params(best[2])
    y = m.forward(X)
    for col, name in ((0, "omega"), (1, "vector bits")):
        pred, true = y[:, col], T[:, col]
        grid = np.array(sorted(set(np.round(true, 6))))
        snap = grid[np.abs(pred[:, None] - grid[None, :]).argmin(1)]
        default = grid[np.abs(grid - (0.4 / 1.3 if col == 0 else 21 / 22)).argmin()]
        for part, sel in (("train", ~va), ("validation", va)):
            if sel.any():
                acc = np.mean(np.isclose(snap[sel], true[sel]))
                base = np.mean(np.isclose(default, true[sel]))
                print(f"{part}: {name} on the best grid point for {100 * acc:.1f}% of levels "
                      f"(the default for {100 * base:.1f}%)", flush=True)
    np.save(args.out, m.params())
    # the features' range over the training levels: outside it the learned
    # setup keeps the defaults
    rng_path = os.path.splitext(args.out)[0] + "_range.npy"
    np.save(rng_path, np.stack([X.min(0), X.max(0)]).astype(np.float32))
    print("wrote", rng_path, flush=True)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
