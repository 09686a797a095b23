"""TEST INFRASTRUCTURE ONLY: numpy restatement of the surrogate network's forward pass (with the
BF16 rounding of the deployed inference emulated) and of its straight-through gradient, the
checker for the GPU kernels (sdfibm_amd/csrc/surrogate.h, surrogate.hip)."""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 11, 32


def bf16(a):
    """Round float32 values to BF16 (nearest, ties to even) and back to float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + N_HID * N_IN].reshape(N_HID, N_IN); o += N_HID * N_IN
    b1 = p[o:o + N_HID]; o += N_HID
    W2 = p[o:o + N_HID * N_HID].reshape(N_HID, N_HID); o += N_HID * N_HID
    b2 = p[o:o + N_HID]; o += N_HID
    w3 = p[o:o + N_HID]; o += N_HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X, rounding=True):
    """Network output for feature rows X [n, >= 11]; intermediate values for the gradient."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if rounding:
        r = lambda a: bf16(np.asarray(a, dtype=np.float32))
        x = r(np.asarray(X, dtype=np.float32)[:, :N_IN])
    else:
        r = lambda a: np.asarray(a, dtype=np.float64)
        x = np.asarray(X, dtype=np.float64)[:, :N_IN]
    W1r, W2r = r(W1), r(W2)
    z1 = x @ W1r.T + b1
    a1 = r(np.maximum(z1, 0))
    z2 = a1 @ W2r.T + b2
    a2 = r(np.maximum(z2, 0))
    out = a2 @ w3 + b3
    return out, (x, z1, a1, z2, a2, W2r)


def grad(p, X, y, rounding=True):
    """Gradient of 0.5 sum (out - y)^2, straight through the rounding (as the training kernels)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    out, (x, z1, a1, z2, a2, W2r) = forward(p, X, rounding)
    e = out - np.asarray(y, dtype=np.float64)
    d2 = (e[:, None] * w3[None, :]) * (z2 > 0)
    d1 = (d2 @ W2r) * (z1 > 0)
    g = np.concatenate([(d1.T @ x).ravel(), d1.sum(0), (d2.T @ a1).ravel(), d2.sum(0), a2.T @ e, [e.sum()]])
    return g, float(np.sum(e * e))
