"""FP64 CPU oracle of the cost-model network -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the product library never calls it.

The network is the one the HIP kernels in opensycl_amd/csrc/mlp.hip
implement (it has no counterpart in the reference): 16 -> 64 -> 64 -> 1,
tanh hidden units, parameters W1[64][16], b1[64], W2[64][64], b2[64], w3[64],
b3 in one flat array, loss L = mean_b (y_b - t_b)^2, and Adam with
beta1 0.9, beta2 0.999, eps 1e-8 and L2 weight decay added to the gradient.

Pinning: ``loss_and_grad`` is checked against central finite differences of
``loss`` (tests/test_oracle_mlp.py), which is independent of the
back-propagation formulas; a planted mistake (``fault=`` below) must fail
that check.  ``adam`` is checked against its closed form on the first step
(|update| = lr for every parameter with a non-zero gradient).
"""
from __future__ import annotations

import numpy as np

IN, H = 16, 64
P = H * IN + H + H * H + H + H + 1
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
assert OFF_B3 + 1 == P == 5313

# Planted faults for the oracle's own tests: "" (none), "no_tanh_derivative"
# (drops the (1 - h^2) factor of the first layer), "w2_not_transposed"
# (back-propagates with W2^T in place of W2, i.e. sum_k d2[k] W2[j][k]).
fault = ""


def unpack(params: np.ndarray):
    p = np.asarray(params, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(H, IN)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(H, H)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (float32, as the kernels use)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    a1 = np.sqrt(6.0 / (IN + H))
    a2 = np.sqrt(6.0 / (H + H))
    a3 = np.sqrt(6.0 / (H + 1))
    p[OFF_W1:OFF_B1] = rng.uniform(-a1, a1, H * IN)
    p[OFF_W2:OFF_B2] = rng.uniform(-a2, a2, H * H)
    p[OFF_W3:OFF_B3] = rng.uniform(-a3, a3, H)
    return p


def forward(params: np.ndarray, x: np.ndarray):
    W1, b1, W2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ w3 + b3
    return y, h1, h2


def loss(params: np.ndarray, x: np.ndarray, t: np.ndarray) -> float:
    y, _, _ = forward(params, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def loss_and_grad(params: np.ndarray, x: np.ndarray, t: np.ndarray):
    W1, b1, W2, b2, w3, b3 = unpack(params)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    B = x.shape[0]
    y, h1, h2 = forward(params, x)
    e = y - t
    L = float(np.mean(e * e))
    dy = 2.0 * e / B                              # [B]
    d2 = dy[:, None] * w3[None, :] * (1.0 - h2 * h2)   # [B, H]
    # back[j] = sum_k d2[k] W2[k][j]
    back = d2 @ (W2.T if fault == "w2_not_transposed" else W2)
    d1 = back if fault == "no_tanh_derivative" else back * (1.0 - h1 * h1)
    g = np.zeros(P, dtype=np.float64)
    g[OFF_W1:OFF_B1] = (d1.T @ x).reshape(-1)
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).reshape(-1)
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    return L, g


def finite_difference_grad(params: np.ndarray, x: np.ndarray, t: np.ndarray, idx, h: float = 1e-6):
    """Central differences of ``loss`` for the parameter indices ``idx``."""
    p = np.asarray(params, dtype=np.float64).copy()
    out = []
    for i in idx:
        old = p[i]
        p[i] = old + h
        lp = loss(p, x, t)
        p[i] = old - h
        lm = loss(p, x, t)
        p[i] = old
        out.append((lp - lm) / (2 * h))
    return np.array(out)


def adam(params, m, v, step, grad, lr, weight_decay, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in FP64; returns new (params, m, v)."""
    p = np.asarray(params, dtype=np.float64)
    g = np.asarray(grad, dtype=np.float64) + weight_decay * p
    m = beta1 * np.asarray(m, dtype=np.float64) + (1 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1 - beta2) * g * g
    mhat = m / (1 - beta1 ** step)
    vhat = v / (1 - beta2 ** step)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v
