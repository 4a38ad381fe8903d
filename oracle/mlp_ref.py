"""float64 numpy restatement of the schedule cost model (test infrastructure).

TEST INFRASTRUCTURE ONLY: the checker for the library's HIP forward /
backward kernels (efficient-dpp_amd/csrc/learn.hip).  Network and parameter
layout as documented in include/edpp_learn.h: 12 -> 64 -> 64 -> 15 with ReLU,
softmax restricted to the row's mask, mean cross-entropy against soft targets.
The gradient here is pinned by central finite differences in
tests/test_mlp_ref.py, so it checks the GPU and is not a copy of it.
"""
import numpy as np

F, H, K = 12, 64, 15
SHAPES = [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (K, H)), ("b3", (K,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = p[o:o + size].reshape(shape)
        o += size
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def forward(p, x, mask):
    """Probabilities (zero outside the mask) and the hidden activations."""
    w = unpack(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ w["W1"].T + w["b1"], 0.0)
    h2 = np.maximum(h1 @ w["W2"].T + w["b2"], 0.0)
    z = h2 @ w["W3"].T + w["b3"]
    valid = np.asarray(mask) > 0
    zm = np.where(valid, z, -np.inf)
    zm = zm - zm.max(axis=1, keepdims=True)
    e = np.where(valid, np.exp(zm), 0.0)
    return e / e.sum(axis=1, keepdims=True), (x, h1, h2, z)


def loss(p, x, target, mask):
    prob, _ = forward(p, x, mask)
    t = np.asarray(target, np.float64)
    valid = (np.asarray(mask) > 0) & (t > 0)
    return float(-np.sum(np.where(valid, t * np.log(np.where(valid, prob, 1.0)), 0.0)) / x.shape[0])


def grad(p, x, target, mask):
    """Analytic gradient of `loss` (backpropagation in float64)."""
    w = unpack(p)
    prob, (x, h1, h2, _) = forward(p, x, mask)
    rows = x.shape[0]
    dz = np.where(np.asarray(mask) > 0, prob - np.asarray(target, np.float64), 0.0) / rows
    g = {"W3": dz.T @ h2, "b3": dz.sum(0)}
    dh2 = (dz @ w["W3"]) * (h2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ w["W2"]) * (h1 > 0)
    g["W1"] = dh1.T @ x
    g["b1"] = dh1.sum(0)
    return pack(g)
