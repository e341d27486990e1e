"""CPU restatement of the radiance MLP (numpy). TEST INFRASTRUCTURE ONLY:
imported by tests/ and bench.py's CPU leg, never by the product.

It mirrors the GPU schedule: the encoding, each GEMM's BF16 operands (weights,
activations, back-propagated gradients rounded to nearest even) and the FP32
output, with FP64 accumulation. With emulate=False nothing is rounded (the FP64
reference whose backward pass is pinned by finite differences).
"""
from __future__ import annotations

import numpy as np

FREQS = 12
SHAPES = [(256, 256)] * 6 + [(16, 256)]


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def encode(f):
    """[n, 9] -> [n, 256]: raw features, then sin / cos(pi 2^k x) per
    feature and frequency, zero padding."""
    f = np.asarray(f, np.float32).astype(np.float64)
    n = f.shape[0]
    out = np.zeros((n, 256))
    out[:, :9] = f
    for d in range(9):
        for k in range(FREQS):
            x = np.pi * f[:, d] * (2.0 ** k)
            out[:, 9 + 24 * d + 2 * k] = np.sin(x)
            out[:, 9 + 24 * d + 2 * k + 1] = np.cos(x)
    return out


def split(w):
    mats, o = [], 0
    for out, inp in SHAPES:
        mats.append(np.asarray(w[o:o + out * inp]).reshape(out, inp))
        o += out * inp
    return mats


def forward(w, f, emulate=True):
    """Returns (pred [n, 3], list of layer inputs, pred16 [n, 16])."""
    r = bf16 if emulate else (lambda x: np.asarray(x, np.float64))
    mats = [r(m) for m in split(w)]
    x = r(encode(f))
    inputs = [x]
    for m in mats[:-1]:
        x = r(np.maximum(x @ m.T, 0.0))
        inputs.append(x)
    pred16 = x @ mats[-1].T
    if emulate:
        pred16 = pred16.astype(np.float32).astype(np.float64)
    return pred16[:, :3], inputs, pred16


def loss_and_grad(w, f, t, emulate=True):
    """Relative L2 loss mean((p - t)^2 / (p^2 + 0.01)), p constant in the
    denominator, and its gradient w.r.t. the flat weights."""
    r = bf16 if emulate else (lambda x: np.asarray(x, np.float64))
    mats = [r(m) for m in split(w)]
    p, inputs, _ = forward(w, f, emulate)
    t = np.asarray(t, np.float64)
    n = p.shape[0]
    den = p * p + 0.01
    loss = float(((p - t) ** 2 / den).sum() / (3 * n))
    g = np.zeros((n, 16))
    g[:, :3] = 2.0 * (p - t) / den / (3 * n)
    g = r(g)
    grads = [None] * len(mats)
    for l in range(len(mats) - 1, -1, -1):
        grads[l] = g.T @ inputs[l]
        if l == 0:
            break
        g = r((g @ mats[l]) * (inputs[l] > 0))
    return loss, np.concatenate([x.ravel() for x in grads])


def loss_only(w, f, t):
    p, _, _ = forward(w, f, emulate=False)
    t = np.asarray(t, np.float64)
    den = (p * p + 0.01)
    return ((p - t) ** 2 / den).sum() / (3 * p.shape[0]), den
