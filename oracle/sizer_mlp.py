"""CPU checker for the engine's learned sizer network -- TEST INFRASTRUCTURE ONLY:
used by tests/ to check the HIP forward/backward kernels, never by the product.

The network has no counterpart in the reference (it sizes buffers that the
reference always counts first).  This is a plain numpy statement of it:
16 -> 64 -> 64 -> 64 -> 1, ReLU on the hidden layers, parameters packed as
W1[64][16], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4, and the
asymmetric loss mean(w * (y_hat - y)^2) with w = alpha where y_hat < y, else 1.
Pinned by central finite differences in float64 (tests/test_sizer.py).
"""
from __future__ import annotations

import numpy as np

F, H = 16, 64
SHAPES = [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)), ("W4", (H,)), ("b4", (1,))]
P = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, at = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[at:at + n].reshape(shape)
        at += n
    return out


def forward(p, x):
    """Returns (y, cache) for inputs x[n, 16]."""
    w = unpack(p)
    z1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(z2, 0)
    z3 = h2 @ w["W3"].T + w["b3"]
    h3 = np.maximum(z3, 0)
    y = h3 @ w["W4"] + w["b4"][0]
    return y, (x, h1, h2, h3)


def loss(p, x, y, alpha):
    yh, _ = forward(p, x)
    r = yh - y
    wgt = np.where(r >= 0, 1.0, alpha)
    return float(np.mean(wgt * r * r))


def grad(p, x, y, alpha):
    """(loss, dloss/dp) by backpropagation."""
    w = unpack(p)
    yh, (x, h1, h2, h3) = forward(p, x)
    n = len(y)
    r = yh - y
    wgt = np.where(r >= 0, 1.0, alpha)
    dy = 2 * wgt * r / n
    g = {}
    g["W4"] = dy @ h3
    g["b4"] = np.array([dy.sum()])
    d3 = np.outer(dy, w["W4"]) * (h3 > 0)
    g["W3"] = d3.T @ h2
    g["b3"] = d3.sum(0)
    d2 = (d3 @ w["W3"]) * (h2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ w["W2"]) * (h1 > 0)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SHAPES])
    return float(np.mean(wgt * r * r)), flat
