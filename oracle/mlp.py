"""Oracle (test infrastructure only): numpy forward/backward of the preconditioner MLP.

Network: 11 -> 64 -> 64 -> 64 -> 9, GELU (tanh form) on the hidden layers,
weights rounded to bfloat16 (round to nearest even) for the matrix products,
accumulation in float32.  Flat parameter order: W1[11x64], b1[64], W2[64x64], b2,
W3[64x64], b3, W4[64x9], b4[9] (row-major, W[in][out]).  Loss: mean over samples
and outputs of (y - t)^2.  Only tests/ may import this module.
"""
from __future__ import annotations

import numpy as np

SIZES = (11, 64, 64, 64, 9)
C_GELU = np.float32(np.sqrt(2.0 / np.pi))


def n_params(sizes=SIZES) -> int:
    return sum(a * b + b for a, b in zip(sizes[:-1], sizes[1:]))


def unpack(theta, sizes=SIZES):
    out, o = [], 0
    for a, b in zip(sizes[:-1], sizes[1:]):
        W = theta[o:o + a * b].reshape(a, b); o += a * b
        bb = theta[o:o + b]; o += b
        out.append((W, bb))
    return out


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(z, ft=np.float32):
    z = z.astype(ft)
    c = ft(np.sqrt(2.0 / np.pi))
    t = np.tanh(c * (z + ft(0.044715) * z * z * z))
    return ft(0.5) * z * (ft(1.0) + t)


def gelu_prime(z, ft=np.float32):
    z = z.astype(ft)
    c = ft(np.sqrt(2.0 / np.pi))
    u = c * (z + ft(0.044715) * z * z * z)
    t = np.tanh(u)
    du = c * (ft(1.0) + ft(3 * 0.044715) * z * z)
    return ft(0.5) * (ft(1.0) + t) + ft(0.5) * z * (ft(1.0) - t * t) * du


def forward(theta, X, sizes=SIZES, exact=False):
    """exact=True: no bfloat16 rounding, float64 throughout (for derivative checks)."""
    ft = np.float64 if exact else np.float32
    rnd = (lambda a: np.asarray(a, dtype=np.float64)) if exact else bf16
    layers = unpack(np.asarray(theta, dtype=ft), sizes)
    h = np.asarray(X, dtype=ft)
    zs, hs = [], [h]
    for k, (W, b) in enumerate(layers):
        z = (rnd(h).astype(np.float64) @ rnd(W).astype(np.float64)).astype(ft) + b
        if k < len(layers) - 1:
            zs.append(z)
            h = gelu(z, ft)
            hs.append(h)
        else:
            return z, zs, hs
    raise AssertionError


def loss_and_grad(theta, X, T, sizes=SIZES, exact=False):
    ft = np.float64 if exact else np.float32
    rnd = (lambda a: np.asarray(a, dtype=np.float64)) if exact else bf16
    theta = np.asarray(theta, dtype=ft)
    layers = unpack(theta, sizes)
    y, zs, hs = forward(theta, X, sizes, exact)
    n = X.shape[0]
    diff = y - T
    loss = float((diff.astype(np.float64) ** 2).mean())
    dz = (2.0 / (n * y.shape[1])) * diff.astype(np.float64)
    grads = []
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        hprev = rnd(hs[k]).astype(np.float64)
        grads.append((hprev.T @ dz, dz.sum(axis=0)))
        if k > 0:
            dh = dz @ rnd(W).astype(np.float64).T
            dz = dh * gelu_prime(zs[k - 1], ft).astype(np.float64)
    grads.reverse()
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads]).astype(ft)
    return loss, g
