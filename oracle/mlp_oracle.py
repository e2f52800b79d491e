"""CPU oracle of the learned-parameter network (TEST INFRASTRUCTURE: used
only by tests/ to check the HIP kernels).  float64 numpy restatement of the
shape and parameter layout the library uses: 8 -> 32 (tanh) -> 32 (tanh) ->
2 (sigmoid); parameters W1[32][8], b1[32], W2[32][32], b2[32], W3[2][32],
b3[2]; loss = mean over samples of sum_k (y_k - t_k)^2."""
from __future__ import annotations

import numpy as np

IN, H, OUT = 8, 32, 2
SHAPES = [(H, IN), (H,), (H, H), (H,), (OUT, H), (OUT,)]
NPARAMS = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, k = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(p[k:k + n].reshape(s))
        k += n
    return out


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, np.float64).reshape(-1, IN)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = 1.0 / (1.0 + np.exp(-(h2 @ W3.T + b3)))
    return y, (x, h1, h2)


def loss_grad(p, x, t):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (x, h1, h2) = forward(p, x)
    t = np.asarray(t, np.float64).reshape(-1, OUT)
    B = x.shape[0]
    e = y - t
    loss = float((e * e).sum() / B)
    d3 = 2.0 * e / B * y * (1.0 - y)
    d2 = (d3 @ W3) * (1.0 - h2 * h2)
    d1 = (d2 @ W2) * (1.0 - h1 * h1)
    g = [d1.T @ x, d1.sum(0), d2.T @ h1, d2.sum(0), d3.T @ h2, d3.sum(0)]
    return loss, np.concatenate([a.ravel() for a in g])
