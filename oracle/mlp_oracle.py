"""Float64 CPU oracle of the learned warm-start MLP -- TEST INFRASTRUCTURE ONLY.

13 -> 64 -> 64 -> 64 -> 2 (inputs padded to 16, outputs to 16), SiLU, mean squared error over the two outputs.
Values are rounded to bf16 (round to nearest even) exactly where the HIP kernels feed the MFMA units: inputs,
weights, hidden activations and the backward errors; sums and everything else are float64. Used by
tests/test_mlp_gpu.py to check the GPU forward pass and gradients.
"""

from __future__ import annotations

import numpy as np

IN, H, OUT = 16, 64, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + OUT * H
N_PARAMS = OFF_B4 + OUT


def _bf16(x) -> np.ndarray:
    """Round float32 values to bf16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32)).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


bf16 = _bf16


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(H, H), p[OFF_B3:OFF_W4], p[OFF_W4:OFF_B4].reshape(OUT, H), p[OFF_B4:])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


IN_USED = 13  # the feature rows carry other data in columns 13..15; the network reads columns 0..12


def _exact(x) -> np.ndarray:
    return np.asarray(x, np.float64)


def forward(p: np.ndarray, X: np.ndarray, rounding: bool = True):
    """``rounding`` False: the same network without the bf16 roundings, all float64 (the smooth function whose
    derivative ``grad`` computes, for finite-difference checks)."""
    bf16 = _bf16 if rounding else _exact
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X = np.array(X, np.float32 if rounding else np.float64, copy=True)
    X[:, IN_USED:] = 0.0
    a0 = bf16(X)
    z1 = a0 @ bf16(W1).T + b1
    a1 = bf16(_silu(z1))
    z2 = a1 @ bf16(W2).T + b2
    a2 = bf16(_silu(z2))
    z3 = a2 @ bf16(W3).T + b3
    a3 = bf16(_silu(z3))
    y = a3 @ bf16(W4).T + b4
    return y, (a0, z1, a1, z2, a2, z3, a3)


def grad(p: np.ndarray, X: np.ndarray, T: np.ndarray, n_total: int | None = None, rounding: bool = True):
    """(gradient [N_PARAMS], squared-error sum) of sum((y[:, :2] - T)^2) / (2 n_total)."""
    bf16 = _bf16 if rounding else _exact
    n = X.shape[0]
    n_total = n_total or n
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (a0, z1, a1, z2, a2, z3, a3) = forward(p, X, rounding)
    d = np.zeros_like(y)
    d[:, :2] = y[:, :2] - np.asarray(T, np.float64)
    se = float(np.sum(d[:, :2] ** 2))
    dy = 2.0 * d / (2.0 * n_total)
    g = np.zeros(N_PARAMS)
    gy = bf16(dy)
    g[OFF_B4:] = dy.sum(0)
    g[OFF_W4:OFF_B4] = (gy.T @ a3).reshape(-1)
    dz3 = (gy @ bf16(W4)) * _silu_grad(z3)
    g3 = bf16(dz3)
    g[OFF_B3:OFF_W4] = dz3.sum(0)
    g[OFF_W3:OFF_B3] = (g3.T @ a2).reshape(-1)
    dz2 = (g3 @ bf16(W3)) * _silu_grad(z2)
    g2 = bf16(dz2)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (g2.T @ a1).reshape(-1)
    dz1 = (g2 @ bf16(W2)) * _silu_grad(z1)
    g1 = bf16(dz1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (g1.T @ a0).reshape(-1)
    return g, se
