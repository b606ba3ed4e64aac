"""Float64 CPU reference of the cost-model MLP (12 -> 64 -> 64 -> 1, ReLU, MSE).

TEST INFRASTRUCTURE: the checker for the GPU MLP kernels (mlp.hip).  Its
analytic gradient is itself pinned by central finite differences in
tests/test_mlp_oracle.py.  Parameter layout matches the GPU one:
W1[64][12], b1[64], W2[64][64], b2[64], W3[64], b3.
"""
import numpy as np

IN, H = 12, 64
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]
COUNT = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def forward(p, X):
    w = unpack(p)
    X = np.asarray(X, np.float64)
    a1 = X @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ w["W3"] + w["b3"][0]
    return y, (X, a1, h1, a2, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - np.asarray(T, np.float64)) ** 2))


def gradients(p, X, T):
    """Loss and analytic gradient (same layout as the parameters)."""
    w = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    n = X.shape[0]
    d3 = 2 * (y - np.asarray(T, np.float64)) / n
    d2 = np.outer(d3, w["W3"]) * (a2 > 0)
    d1 = (d2 @ w["W2"]) * (a1 > 0)
    g = {
        "W1": d1.T @ X, "b1": d1.sum(0),
        "W2": d2.T @ h1, "b2": d2.sum(0),
        "W3": d3 @ h2, "b3": np.array([d3.sum()]),
    }
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SIZES])
    return float(np.mean((y - T) ** 2)), flat
