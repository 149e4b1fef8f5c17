"""FP64 reference of the JKR surrogate network (TEST INFRASTRUCTURE ONLY).

Forward and backward pass of the 4 -> 64 -> 64 -> 2 tanh MLP in numpy FP64.
With bf16=True the operands of every matrix product are rounded to BF16 at
the same points as the GPU kernels (inputs, weights, hidden activations,
output and hidden deltas), so the GPU's FP32-accumulated products can be
compared with it closely.  The backward pass itself is pinned by central
finite differences of the bf16=False loss (tests/test_surrogate.py).
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 4, 64, 2
SIZES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_backward(p, x, t, bf=False):
    """Returns (loss, grad packed, y) for scaled inputs x [m x 4] and targets t [m x 2]."""
    w = unpack(p)
    r = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, dtype=np.float64))
    X = r(x)
    W1, W2, W3 = r(w["W1"]), r(w["W2"]), r(w["W3"])
    H1 = np.tanh(X @ W1 + w["b1"])
    if bf:
        H1 = H1.astype(np.float32).astype(np.float64)
    H1b = r(H1)
    H2 = np.tanh(H1b @ W2 + w["b2"])
    if bf:
        H2 = H2.astype(np.float32).astype(np.float64)
    H2b = r(H2)
    Y = H2b @ W3 + w["b3"]
    m = x.shape[0]
    E = Y - t
    loss = float(np.mean(E ** 2))
    dY = 2.0 * E / (m * OUT)
    dYb = r(dY)
    g = {"W3": H2b.T @ dYb, "b3": dY.sum(0)}
    dH2 = dYb @ W3.T
    dZ2 = dH2 * (1.0 - H2 ** 2)
    dZ2b = r(dZ2)
    g["b2"] = dZ2.sum(0)
    g["W2"] = H1b.T @ dZ2b
    dH1 = dZ2b @ W2.T
    dZ1 = dH1 * (1.0 - H1 ** 2)
    dZ1b = r(dZ1)
    g["b1"] = dZ1.sum(0)
    g["W1"] = X.T @ dZ1b
    return loss, pack(g), Y


def forward(p, x, bf=False):
    return forward_backward(p, x, np.zeros((x.shape[0], OUT)), bf)[2]
