"""CPU restatement of the planner cost model (test infrastructure only).

The library's learned planner scores candidate plans with a perceptron
(16 features -> 3 x 64 ReLU -> 1, FP32) whose forward pass, gradient and
Adam step run as HIP kernels (aleph_amd/csrc/k_mlp.hip).  This module is
the same computation in float64 numpy, used only by the tests to check the
device kernels; its own gradient is pinned by central finite differences
(tests/test_cost_model.py).  There is no counterpart in the reference,
whose fusion widths are fixed rules (crates/aleph-ir/src/passes/fuse_kq.rs).

Flat parameter layout (shared with the device): W1[16][64], b1[64],
W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4.
"""

from __future__ import annotations

import numpy as np

IN, H = 16, 64
SIZES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("w4", (H,)), ("b4", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def forward(p: np.ndarray, X: np.ndarray) -> np.ndarray:
    w = unpack(p)
    h1 = np.maximum(X @ w["W1"] + w["b1"], 0.0)
    h2 = np.maximum(h1 @ w["W2"] + w["b2"], 0.0)
    h3 = np.maximum(h2 @ w["W3"] + w["b3"], 0.0)
    return h3 @ w["w4"] + w["b4"][0]


def loss_and_grad(p: np.ndarray, X: np.ndarray, y: np.ndarray):
    """Mean squared error and its gradient with respect to the flat parameters."""
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    B = X.shape[0]
    z1 = X @ w["W1"] + w["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ w["W3"] + w["b3"]
    h3 = np.maximum(z3, 0.0)
    out = h3 @ w["w4"] + w["b4"][0]
    r = out - y
    loss = float(np.mean(r * r))
    d4 = 2.0 * r / B
    g = {"w4": h3.T @ d4, "b4": np.array([d4.sum()])}
    d3 = np.outer(d4, w["w4"]) * (z3 > 0)
    g["W3"], g["b3"] = h2.T @ d3, d3.sum(0)
    d2 = (d3 @ w["W3"].T) * (z2 > 0)
    g["W2"], g["b2"] = h1.T @ d2, d2.sum(0)
    d1 = (d2 @ w["W2"].T) * (z1 > 0)
    g["W1"], g["b1"] = X.T @ d1, d1.sum(0)
    return loss, pack(g)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
