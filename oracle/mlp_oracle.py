"""CPU FP64 oracle of the learned warm-start network — TEST INFRASTRUCTURE ONLY.

The network has no reference counterpart (it is this project's addition to the PBF
solver).  This module restates it in float64 numpy: an 8 -> 64 -> 64 -> 1 MLP with the
tanh form of GELU, inputs standardised as (x - mu) / sigma and the output scaled by
outScale, trained on the mean squared error between output / outScale and target /
outScale.  The parameter vector layout is the one of include/rtp.h:

    W1[8][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3

`bf16=True` rounds inputs, weights and layer-1 activations to bfloat16 (round to nearest
even) exactly where the GPU kernel feeds the matrix cores, so the remaining differences
are FP32 accumulation order.
"""
from __future__ import annotations

import numpy as np

N_PARAMS = 8 * 64 + 64 + 64 * 64 + 64 + 64 + 1
OFF = {"W1": 0, "b1": 512, "W2": 576, "b2": 4672, "w3": 4736, "b3": 4800}
A = np.sqrt(2.0 / np.pi)
B = 0.044715


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    return (p[0:512].reshape(8, 64), p[512:576], p[576:4672].reshape(64, 64), p[4672:4736],
            p[4736:4800], p[4800])


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(A * (z + B * z ** 3)))


def gelu_grad(z):
    t = np.tanh(A * (z + B * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * A * (1.0 + 3.0 * B * z * z)


def forward(p, feat, mu, sigma, out_scale, bf16_ops: bool = False):
    """Returns (output, cache) for feat (n x 8)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = (np.asarray(feat, np.float64) - np.asarray(mu, np.float64)) / np.asarray(sigma, np.float64)
    if bf16_ops:
        x = bf16(x.astype(np.float32))
        W1 = bf16(W1.astype(np.float32))
        W2 = bf16(W2.astype(np.float32))
    z1 = x @ W1 + b1
    h1 = gelu(z1)
    h1m = bf16(h1.astype(np.float32)) if bf16_ops else h1
    z2 = h1m @ W2 + b2
    h2 = gelu(z2)
    y = h2 @ w3 + b3
    return y * out_scale, (x, z1, h1m, z2, h2, y, W2)


def gradients(p, feat, target, mu, sigma, out_scale, bf16_ops: bool = False):
    """Gradient of mean((y - target / outScale)^2) w.r.t. the parameter vector, and the loss."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    _, (x, z1, h1m, z2, h2, y, W2op) = forward(p, feat, mu, sigma, out_scale, bf16_ops)
    n = x.shape[0]
    t = np.asarray(target, np.float64) / out_scale
    e = y - t
    dy = 2.0 * e / n
    g = np.zeros(N_PARAMS)
    g[OFF["b3"]] = dy.sum()
    g[OFF["w3"]:OFF["w3"] + 64] = h2.T @ dy
    dz2 = np.outer(dy, w3) * gelu_grad(z2)
    g[OFF["b2"]:OFF["b2"] + 64] = dz2.sum(0)
    g[OFF["W2"]:OFF["W2"] + 4096] = (h1m.T @ dz2).ravel()
    dh1 = dz2 @ W2op.T
    dz1 = dh1 * gelu_grad(z1)
    g[OFF["b1"]:OFF["b1"] + 64] = dz1.sum(0)
    g[OFF["W1"]:OFF["W1"] + 512] = (x.T @ dz1).ravel()
    return g, float(np.mean(e * e))
