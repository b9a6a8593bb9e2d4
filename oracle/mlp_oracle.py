"""CPU oracle of the step-size MLP -- TEST INFRASTRUCTURE ONLY.

numpy forward and backward of the 6 -> 64 -> 64 -> 64 -> 1 tanh network with
the parameter layout of pyquake3d_amd/csrc/mlp_kernels.hip (W1 64x6, b1,
W2 64x64, b2, W3 64x64, b3, W4 64, b4), loss sum 0.5 (y - t)^2 / n. Pinned by
central finite differences in tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np

IN, H = 6, 64
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
         ("b3", (H,)), ("W4", (H,)), ("b4", (1,))]
P = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, s in SIZES:
        k = int(np.prod(s))
        out[name] = p[o:o + k].reshape(s)
        o += k
    return out


def forward(p, X):
    w = unpack(np.asarray(p, dtype=np.float64))
    a1 = np.tanh(X @ w["W1"].T + w["b1"])
    a2 = np.tanh(a1 @ w["W2"].T + w["b2"])
    a3 = np.tanh(a2 @ w["W3"].T + w["b3"])
    return a3 @ w["W4"] + w["b4"][0], (a1, a2, a3)


def loss_and_grad(p, X, T):
    """Mean of 0.5 (y - t)^2 and its gradient (layout P)."""
    X = np.asarray(X, dtype=np.float64)
    w = unpack(np.asarray(p, dtype=np.float64))
    y, (a1, a2, a3) = forward(p, X)
    n = len(X)
    dy = (y - T) / n
    g = {"W4": a3.T @ dy, "b4": np.array([dy.sum()])}
    d3 = np.outer(dy, w["W4"]) * (1 - a3 ** 2)
    g["W3"], g["b3"] = d3.T @ a2, d3.sum(0)
    d2 = (d3 @ w["W3"]) * (1 - a2 ** 2)
    g["W2"], g["b2"] = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ w["W2"]) * (1 - a1 ** 2)
    g["W1"], g["b1"] = d1.T @ X, d1.sum(0)
    grad = np.concatenate([g[name].ravel() for name, _ in SIZES])
    return float(0.5 * np.mean((y - T) ** 2)), grad
