"""CPU restatement of the rate surrogate (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module, as the checker of the GPU surrogate kernels
(flysim_amd/csrc/mlp_kernels.hpp); nothing in the product imports it.

The surrogate has no counterpart in the reference (it is a new part of the
sweep driver), so the oracle is pinned by mathematics instead of reference
vectors: central finite differences of the loss pin the backward pass, the
closed forms of GELU (gelu(0) = 0, gelu(x) - gelu(-x) = x, gelu'(0) = 1/2)
pin the activation, and the first Adam step (update = -lr * g / (|g| + eps))
pins the optimiser.  tests/test_mlp_oracle.py holds those checks.

Conventions shared with the GPU (include/flysim_mlp.h): flat parameters, for
each layer W[out][in] row-major then b[out]; GELU (erf form) after every layer
but the last; loss = mean over rows and outputs of (y - t)^2; Adam with bias
correction and decoupled L2 term g + wd * p.  With ``bf16=True`` the forward
and backward round to bfloat16 (nearest-even) exactly where the GPU does:
inputs, weights, hidden activations, and the output-side gradients fed to the
matrix products; everything else stays FP64.
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import numpy as np

_ERF = np.vectorize(math.erf, otypes=[np.float64])


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z: np.ndarray) -> np.ndarray:
    return 0.5 * z * (1.0 + _ERF(z / math.sqrt(2.0)))


def gelu_grad(z: np.ndarray) -> np.ndarray:
    return 0.5 * (1.0 + _ERF(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unflatten(p: np.ndarray, sizes: Sequence[int]) -> List[Tuple[np.ndarray, np.ndarray]]:
    out, o = [], 0
    for i in range(len(sizes) - 1):
        n_in, n_out = sizes[i], sizes[i + 1]
        W = p[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = p[o:o + n_out]
        o += n_out
        out.append((W, b))
    assert o == p.size
    return out


def n_params(sizes: Sequence[int]) -> int:
    return int(sum(sizes[i + 1] * sizes[i] + sizes[i + 1] for i in range(len(sizes) - 1)))


def forward(p: np.ndarray, sizes: Sequence[int], x: np.ndarray, bf16: bool = False):
    """Returns (y, cache) with cache = [(h_in, z)] per layer."""
    layers = unflatten(np.asarray(p, np.float64), sizes)
    h = bf16_round(x) if bf16 else np.asarray(x, np.float64)
    cache = []
    for i, (W, b) in enumerate(layers):
        Wm = bf16_round(W) if bf16 else W
        z = h @ Wm.T + b
        if bf16:
            z = z.astype(np.float32).astype(np.float64)  # f32 accumulator + bias
        cache.append((h, z))
        if i < len(layers) - 1:
            h = gelu(z)
            if bf16:
                h = bf16_round(h)
        else:
            h = z
    return h, cache


def loss_grad(p: np.ndarray, sizes: Sequence[int], x: np.ndarray, t: np.ndarray, bf16: bool = False):
    """Mean-squared-error loss and its gradient (flat, like p)."""
    y, cache = forward(p, sizes, x, bf16)
    t = np.asarray(t, np.float64)
    rows, n_out = t.shape
    diff = y - t
    loss = float(np.mean(diff * diff))
    dz = 2.0 * diff / (rows * n_out)
    layers = unflatten(np.asarray(p, np.float64), sizes)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        h_in, _ = cache[i]
        dzm = bf16_round(dz) if bf16 else dz
        dW = dzm.T @ h_in
        db = dz.sum(axis=0)
        grads[i] = (dW, db)
        if i > 0:
            Wm = bf16_round(W) if bf16 else W
            dh = dzm @ Wm
            dz = dh * gelu_grad(cache[i - 1][1])
            if bf16:
                dz = dz.astype(np.float32).astype(np.float64)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat


def adam_step(p, g, m1, m2, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (Kingma & Ba 2015, algorithm 1) at 1-based step t;
    returns (p, m1, m2)."""
    g = g + wd * p
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** t)
    vh = m2 / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2
