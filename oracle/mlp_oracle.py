"""FP64 CPU restatement of the density surrogate's network (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's baseline leg use this
module, as the checker of the HIP kernels in bbrcit_kde_amd/csrc/bk_mlp.hip.
The surrogate has no counterpart in the reference (dchao34/bbrcit_kde), so
this file pins it by mathematics instead: tests/test_oracle_mlp.py checks
the analytic gradient below against central finite differences in FP64 and
the BF16 rounding against hand-computed round-to-nearest-even cases.

Network: y = W4 . s(W3 s(W2 s(W1 x + b1) + b2) + b3) + b4, s(z) = z sigmoid(z),
W[out][in]; the forward pass uses the BF16-rounded weights (biases FP32);
loss = mean_b (y_b - t_b)^2.  Parameter order as BK_MLP_NUM_PARAMS in
include/bbrcit_hip.h: W1[128][2], b1, W2[128][128], b2, W3[128][128], b3,
W4[128], b4.
"""
from __future__ import annotations

import numpy as np

WIDTH = 128
SIZES = [("W1", (WIDTH, 2)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
         ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH,)), ("b4", (1,))]
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16_round(a) -> np.ndarray:
    """Round FP32 values to the nearest BF16 value, ties to even (as float32)."""
    a = np.asarray(a, dtype=np.float32)
    u = np.ascontiguousarray(a.reshape(-1)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(theta) -> dict:
    theta = np.asarray(theta)
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = theta[off:off + n].reshape(shape)
        off += n
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).ravel() for name, _ in SIZES])


def quantize(theta) -> np.ndarray:
    """Working parameters: weights rounded to BF16, biases unchanged (FP64)."""
    p = unpack(np.asarray(theta, dtype=np.float32))
    q = {k: (bf16_round(v).astype(np.float64) if k.startswith("W") else v.astype(np.float64)) for k, v in p.items()}
    return pack(q)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(wq, x):
    """Outputs y (n,) and the cache of pre-activations for n inputs x (n, 2)."""
    p = unpack(np.asarray(wq, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = z1 * _sig(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = z2 * _sig(z2)
    z3 = h2 @ p["W3"].T + p["b3"]
    h3 = z3 * _sig(z3)
    y = h3 @ p["W4"] + p["b4"][0]
    return y, (x, z1, h1, z2, h2, z3, h3)


def _dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def loss_grad(wq, x, t):
    """mean (y - t)^2 and its gradient with respect to the parameters."""
    p = unpack(np.asarray(wq, dtype=np.float64))
    y, (x, z1, h1, z2, h2, z3, h3) = forward(wq, x)
    n = len(y)
    r = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {"W4": h3.T @ dy, "b4": np.array([dy.sum()])}
    dz3 = np.outer(dy, p["W4"]) * _dsilu(z3)
    g["W3"] = dz3.T @ h2
    g["b3"] = dz3.sum(axis=0)
    dz2 = (dz3 @ p["W3"]) * _dsilu(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(axis=0)
    dz1 = (dz2 @ p["W2"]) * _dsilu(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(axis=0)
    return loss, pack(g)


def density(wq, q, lo, hi, mu, sigma, eps):
    """Surrogate density max(exp(mu + sigma y) - eps, 0) at queries q (n, 2)."""
    c, s = 0.5 * (lo + hi), 0.5 * (hi - lo)
    y, _ = forward(wq, (np.asarray(q, dtype=np.float64) - c) / s)
    return np.maximum(np.exp(mu + sigma * y) - eps, 0.0)
