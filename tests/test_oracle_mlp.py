"""Pins the FP64 oracle of the cost-model network (oracle/mlp.py) without
relying on its own back-propagation formulas: gradients against central
finite differences of the loss, planted mistakes that must fail that check,
and Adam's closed form on the first step."""
import numpy as np
import pytest

import oracle.mlp as M


def _data(seed=0, B=37):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, M.IN))
    t = rng.standard_normal(B)
    p = M.init_params(seed).astype(np.float64)
    p[M.OFF_B1:M.OFF_W2] = rng.standard_normal(M.H) * 0.1  # non-zero biases
    p[M.OFF_B2:M.OFF_W3] = rng.standard_normal(M.H) * 0.1
    p[M.OFF_B3] = 0.3
    return p, x, t


def _sample_indices():
    # a few parameters of every block
    rng = np.random.default_rng(1)
    blocks = [(M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
              (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.P)]
    idx = []
    for lo, hi in blocks:
        idx += list(rng.integers(lo, hi, min(6, hi - lo)))
    return idx


def test_gradient_matches_finite_differences():
    p, x, t = _data()
    L, g = M.loss_and_grad(p, x, t)
    assert abs(L - M.loss(p, x, t)) < 1e-15
    idx = _sample_indices()
    fd = M.finite_difference_grad(p, x, t, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("fault", ["no_tanh_derivative", "w2_not_transposed"])
def test_planted_gradient_mistake_is_caught(fault):
    p, x, t = _data(2)
    idx = _sample_indices()
    fd = M.finite_difference_grad(p, x, t, idx)
    M.fault = fault
    try:
        _, g = M.loss_and_grad(p, x, t)
    finally:
        M.fault = ""
    assert not np.allclose(g[idx], fd, rtol=1e-4, atol=1e-7)


def test_adam_first_step_closed_form():
    # step 1 from zero moments: mhat = g, vhat = g^2 -> update = lr * g/(|g|+eps)
    rng = np.random.default_rng(3)
    p = rng.standard_normal(M.P)
    g = rng.standard_normal(M.P)
    g[::7] = 0.0
    new, m, v = M.adam(p, np.zeros(M.P), np.zeros(M.P), 1, g, lr=0.01, weight_decay=0.0)
    step = p - new
    nz = g != 0
    np.testing.assert_allclose(np.abs(step[nz]), 0.01 * np.abs(g[nz]) / (np.abs(g[nz]) + 1e-8), rtol=1e-12)
    assert np.all(step[~nz] == 0)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_forward_shapes_and_tanh_bounds():
    p, x, t = _data(4, B=5)
    y, h1, h2 = M.forward(p, x)
    assert y.shape == (5,) and h1.shape == (5, M.H) and h2.shape == (5, M.H)
    assert np.all(np.abs(h1) < 1) and np.all(np.abs(h2) < 1)
    # a network with w3 = 0 predicts b3 everywhere
    q = p.copy()
    q[M.OFF_W3:M.OFF_B3] = 0
    assert np.allclose(M.forward(q, x)[0], q[M.OFF_B3])
