"""Pins of the surrogate oracle (oracle/mlp.py) by mathematics: central finite
differences of the loss for the backward pass, closed forms of GELU and of
the first Adam step, and the bfloat16 rounding rule.  CPU only."""
import math

import numpy as np
import pytest

from oracle import mlp as M

SIZES = [5, 7, 6, 3]


def _params(seed=0, sizes=SIZES):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.6, M.n_params(sizes))


def test_gelu_closed_forms():
    z = np.linspace(-6, 6, 101)
    np.testing.assert_allclose(M.gelu(np.array([0.0])), [0.0], atol=0)
    # gelu(x) - gelu(-x) = x  (x * Phi(x) - (-x) * Phi(-x) = x (Phi(x) + Phi(-x)))
    np.testing.assert_allclose(M.gelu(z) - M.gelu(-z), z, atol=1e-13)
    # gelu(1) = Phi(1) = 0.8413447460685429...
    np.testing.assert_allclose(M.gelu(np.array([1.0])), [0.8413447460685429], rtol=1e-14)
    np.testing.assert_allclose(M.gelu_grad(np.array([0.0])), [0.5], rtol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-8)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    x = rng.normal(size=(9, SIZES[0]))
    t = rng.normal(size=(9, SIZES[-1]))
    loss, g = M.loss_grad(p, SIZES, x, t)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (M.loss_grad(p + e, SIZES, x, t)[0] - M.loss_grad(p - e, SIZES, x, t)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    # a planted mistake (transposed weight gradient of the first layer) is caught
    W0 = g[: SIZES[1] * SIZES[0]].reshape(SIZES[1], SIZES[0])
    bad = g.copy()
    bad[: W0.size] = W0.T.reshape(-1)[: W0.size] if W0.shape[0] == W0.shape[1] else np.roll(W0.ravel(), 1)
    assert not np.allclose(bad, fd, rtol=1e-3, atol=1e-6)


def test_loss_is_mean_squared_error():
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.normal(size=(4, SIZES[0]))
    y, _ = M.forward(p, SIZES, x)
    t = y + 0.5
    loss, _ = M.loss_grad(p, SIZES, x, t)
    assert abs(loss - 0.25) < 1e-12


def test_first_adam_step_closed_form():
    rng = np.random.default_rng(5)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    lr, eps = 1e-3, 1e-8
    p1, m1, m2 = M.adam_step(p, g, np.zeros(50), np.zeros(50), 1, lr, eps=eps)
    np.testing.assert_allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(m1, 0.1 * g)
    np.testing.assert_allclose(m2, 0.001 * g * g)


def test_bf16_rounding_rule():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.14159265, 0.0],
                 dtype=np.float64)
    r = M.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    np.testing.assert_array_equal(r[:4], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6])
    assert r[4] == -2.5 and r[6] == 0.0
    assert abs(r[5] - math.pi) < 2 ** -7 * 4


def test_bf16_mode_is_close_to_fp64():
    rng = np.random.default_rng(6)
    sizes = [54, 64, 64, 50]
    p = rng.normal(0, 0.1, M.n_params(sizes))
    x = rng.normal(size=(32, 54))
    t = rng.normal(size=(32, 50))
    y64, _ = M.forward(p, sizes, x)
    y16, _ = M.forward(p, sizes, x, bf16=True)
    assert np.max(np.abs(y64 - y16)) < 0.05 * np.max(np.abs(y64))
    _, g64 = M.loss_grad(p, sizes, x, t)
    _, g16 = M.loss_grad(p, sizes, x, t, bf16=True)
    assert np.linalg.norm(g16 - g64) < 0.02 * np.linalg.norm(g64)
