"""CPU oracle of the rank-predictor network (test infrastructure only).

numpy FP64 restatement of the MLP the library trains and runs on the GPU:
8 -> 64 -> 64 -> 1, GELU(x) = 0.5 x (1 + erf(x / sqrt 2)), loss
mean((y - t)^2), parameters in the library's order (W1[64][8], b1[64],
W2[64][64], b2[64], W3[64], b3).  Also the bfloat16 rounding of the
inference path and the block features.  No reference counterpart: the
network has none in amiraa127/Dense_HODLR; this file pins the GPU kernels.
"""
from __future__ import annotations

import math

import numpy as np

IN, H = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1

_erf = np.vectorize(math.erf)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[OFF_W1:OFF_B1].reshape(H, IN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(H, H),
            P[OFF_B2:OFF_W3], P[OFF_W3:OFF_B3], P[OFF_B3])


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def forward(P, X, keep=False):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1.T + b1
    A1 = gelu(Z1)
    Z2 = A1 @ W2.T + b2
    A2 = gelu(Z2)
    y = A2 @ W3 + b3
    if keep:
        return y, (X, Z1, A1, Z2, A2)
    return y


def loss_and_grad(P, X, t):
    """mean((y - t)^2) and its gradient in the parameter order."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    y, (X, Z1, A1, Z2, A2) = forward(P, X, keep=True)
    n = len(y)
    d3 = 2.0 * (y - np.asarray(t, dtype=np.float64)) / n
    d2 = d3[:, None] * W3[None, :] * gelu_grad(Z2)
    d1 = (d2 @ W2) * gelu_grad(Z1)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (d1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ A1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = A2.T @ d3
    g[OFF_B3] = d3.sum()
    return float(np.mean((y - t) ** 2)), g


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward_bf16(P, X):
    """Inference with parameters and activations rounded to bfloat16 and
    FP64 accumulation (the products of two bfloat16 numbers are exact)."""
    Pb = bf16(np.asarray(P, dtype=np.float32)).astype(np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(Pb)
    Xb = bf16(np.asarray(X, dtype=np.float32)).astype(np.float64)
    A1 = bf16(gelu(Xb @ W1.T + b1).astype(np.float32)).astype(np.float64)
    A2 = bf16(gelu(A1 @ W2.T + b2).astype(np.float32)).astype(np.float64)
    return A2 @ W3 + b3


def block_features(points, r0, nr, c0, nc, level, length_scale, tol, family):
    """The 8 features of one block of rows [r0, r0+nr) x columns [c0, c0+nc)."""
    P = np.asarray(points, dtype=np.float64)
    if P.ndim == 1:
        P = P[:, None]
    R, Cc = P[r0:r0 + nr], P[c0:c0 + nc]
    cr, cc = R.mean(0), Cc.mean(0)
    dr = np.sqrt(((R - cr) ** 2).sum(1).max())
    dc = np.sqrt(((Cc - cc) ** 2).sum(1).max())
    sep = np.sqrt(((cr - cc) ** 2).sum())
    return np.array([math.log(nr + nc), sep / (dr + dc) if dr + dc > 0 else 0.0,
                     dr / length_scale, dc / length_scale, math.log(length_scale),
                     math.log10(tol), level, family])
