"""O2 -- CPU restatement of the surrogate MLP's forward and backward passes.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ to check the device
kernels (include/pmmlp.h); never by the product.

The reference has no learned model; this is the textbook dense network the
device kernels implement: hidden layers a = gelu(x W^T + b) with
gelu(z) = z (1 + erf(z / sqrt 2)) / 2, a linear output layer, weighted MSE,
Adam with decoupled weight decay.  With ``bf16=True`` the GEMM operands are
rounded to BF16 (nearest even) exactly as the device stages them, and the
sums are FP64, so the device (FP32 sums) can be compared tightly.  Pinned in
tests/test_oracle_mlp.py by finite differences of its own forward pass
(FP64, no rounding), GELU values from scipy's erf, and the closed form of
one Adam step.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(x) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (r.astype(np.uint32) << 16).view(np.float32)
    return np.where(np.isnan(a), a, out).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def _mm(a, b, rnd):
    if rnd:
        return bf16(a) @ bf16(b)
    return np.asarray(a, dtype=np.float64) @ np.asarray(b, dtype=np.float64)


def forward(params, X, bf16_ops=True):
    """params: list of (W [out, in], b [out]).  Returns (Y, cache)."""
    acts = [np.asarray(X, dtype=np.float64)]
    zs = []
    a = acts[0]
    for li, (W, b) in enumerate(params):
        z = _mm(a, np.asarray(W).T, bf16_ops) + np.asarray(b, dtype=np.float64)
        if li < len(params) - 1:
            zs.append(z)
            a = gelu(z)
        else:
            a = z
        acts.append(a)
    return a, (acts, zs)


def backward(params, cache, dY, bf16_ops=True):
    """Gradients [(dW, db)] of sum(dY * Y) through the network."""
    acts, zs = cache
    grads = [None] * len(params)
    g = np.asarray(dY, dtype=np.float64)
    for li in range(len(params) - 1, -1, -1):
        W, _ = params[li]
        dW = _mm(g.T, acts[li], bf16_ops)
        db = g.sum(axis=0)
        grads[li] = (dW, db)
        if li > 0:
            g = _mm(g, np.asarray(W), bf16_ops) * dgelu(zs[li - 1])
    return grads


def mse(Y, T, w=None):
    Y = np.asarray(Y, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    wt = np.ones(Y.shape[1]) if w is None else np.asarray(w, dtype=np.float64)
    n = Y.size
    d = Y - T
    return float(np.sum(wt * d * d) / n), 2.0 * wt * d / n


def adam(p, g, m, v, lr, b1, b2, eps, wd, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1**t)
    vh = v / (1 - b2**t)
    p = p - lr * (mh / (np.sqrt(vh) + eps) + wd * p)
    return p, m, v
