"""FP64 numpy reference of the band-energy surrogate MLP (test infrastructure only).

The network is the project's own (the reference has no learned component): periodic Fourier features
of k -> Linear -> SiLU -> Linear -> SiLU -> Linear -> band energies. This module restates its forward
pass, loss and analytic gradients in FP64 so the HIP kernels can be checked; the gradients here are
themselves pinned by central finite differences in tests/test_surrogate_cpu.py.

Parameter layout (row-major, the same as the device buffers):
  W1 (h, f), b1 (h), W2 (h, h), b2 (h), W3 (o, h), b3 (o), with f = 2 * len(mvec).
"""
from __future__ import annotations

import numpy as np


def fourier_vectors(nf2: int) -> np.ndarray:
    """The nf2 shortest nonzero integer vectors m, one of each +-m pair, ordered by |m|^2 then
    lexicographically (features cos/sin(2 pi m.k) are periodic in k with period 1)."""
    out = []
    r = 4
    for a in range(-r, r + 1):
        for b in range(-r, r + 1):
            for c in range(-r, r + 1):
                m = (a, b, c)
                if m == (0, 0, 0):
                    continue
                if tuple(-x for x in m) in out:
                    continue
                out.append(m)
    out.sort(key=lambda m: (m[0] ** 2 + m[1] ** 2 + m[2] ** 2, m))
    return np.array(out[:nf2], dtype=np.int32)


def features(xk: np.ndarray, mvec: np.ndarray) -> np.ndarray:
    t = 2 * np.pi * (xk @ mvec.T.astype(np.float64))
    return np.concatenate([np.cos(t), np.sin(t)], axis=1)


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(p: dict, xk: np.ndarray, mvec: np.ndarray):
    f = features(xk, mvec)
    z1 = f @ p["W1"].T + p["b1"]
    a1 = silu(z1)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = silu(z2)
    y = a2 @ p["W3"].T + p["b3"]
    return y, (f, z1, a1, z2, a2)


def loss_and_grad(p: dict, xk: np.ndarray, target: np.ndarray, mvec: np.ndarray):
    """Mean squared error over points and bands, and its gradient w.r.t. every parameter."""
    y, (f, z1, a1, z2, a2) = forward(p, xk, mvec)
    n, o = y.shape
    d = y - target
    loss = float((d * d).sum() / (n * o))
    gy = 2 * d / (n * o)
    g = {"W3": gy.T @ a2, "b3": gy.sum(0)}
    gz2 = (gy @ p["W3"]) * dsilu(z2)
    g["W2"], g["b2"] = gz2.T @ a1, gz2.sum(0)
    gz1 = (gz2 @ p["W2"]) * dsilu(z1)
    g["W1"], g["b1"] = gz1.T @ f, gz1.sum(0)
    return loss, g


def init_params(rng: np.random.Generator, nfeat: int, hidden: int, nout: int) -> dict:
    def lin(o, i):
        return rng.uniform(-1, 1, size=(o, i)) * np.sqrt(3.0 / i)
    return {"W1": lin(hidden, nfeat), "b1": np.zeros(hidden), "W2": lin(hidden, hidden), "b2": np.zeros(hidden),
            "W3": lin(nout, hidden), "b3": np.zeros(nout)}
