"""CPU restatement of the threshold-surrogate network (training and inference).

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by dendra_amd.

The network and its training step are defined by dendra_amd/csrc/dn_mlp.hip
(this build's own design, with no counterpart in the reference): members
128 -> 512 -> 512 -> 512 -> 1 with exact-erf GELU, BF16 weight copies,
float32 accumulation on the GPU, activations and backward signals rounded to
BF16 where they enter a matrix product, Adam on float32 masters.  This module
recomputes everything in float64 with the same BF16 rounding points
(``emulate=True``) so that GPU and CPU agree to accumulation-order error, or
exactly in float64 without rounding (``emulate=False``).

Pins (tests/test_mlp_ref.py): the float64 backward equals torch autograd of
the same float64 network; BF16 rounding equals torch's float32 -> bfloat16
conversion (round to nearest even); Adam equals torch.optim.Adam.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(
        2.0 * math.pi)


def forward(W, X, emulate=True):
    """W: dict of float64 arrays W1 (512,128), b1 (512,), ..., w4 (512,), b4 ().
    X: (B, 128).  Returns y (B,) and the cache."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Wq = {k: (r(v) if k[0] in "Ww" else np.asarray(v, np.float64)) for k, v in W.items()}
    H = r(X)
    cache = {"W": Wq, "H0": H}
    for l in (1, 2, 3):
        Z = H @ Wq[f"W{l}"].T + Wq[f"b{l}"]
        H = r(gelu(Z))
        cache[f"Z{l}"], cache[f"H{l}"] = Z, H
    y = H @ Wq["w4"] + Wq["b4"]
    cache["y"] = y
    return y, cache


def backward(cache, t, emulate=True):
    """Gradients of mean((y - t)^2) for every parameter (float64)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Wq, y = cache["W"], cache["y"]
    B = y.shape[0]
    dy = 2.0 * (y - t) / B
    g = {"w4": cache["H3"].T @ dy, "b4": dy.sum()}
    dZ = r(np.outer(dy, Wq["w4"]) * gelu_d(cache["Z3"]))
    for l in (3, 2, 1):
        g[f"b{l}"] = dZ.sum(0)
        g[f"W{l}"] = dZ.T @ cache[f"H{l - 1}"]
        if l > 1:
            dZ = r((dZ @ Wq[f"W{l}"]) * gelu_d(cache[f"Z{l - 1}"]))
    return g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (float64), returning (p, m, v); t counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    lr_t = lr * math.sqrt(1 - b2 ** t) / (1 - b1 ** t)
    return p - lr_t * m / (np.sqrt(v) + eps), m, v
