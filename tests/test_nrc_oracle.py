"""Pins of the NRC oracle (oracle/nrc.py): what the mathematics fixes, independent of its code paths."""
import math

import numpy as np

from oracle import nrc


def batch(n, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 12))
    raw[:, 0:3] = rng.random((n, 3))
    raw[:, 3] = rng.random(n)
    d = rng.normal(size=(n, 3))
    raw[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    raw[:, 8:11] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    target = rng.uniform(0.0, 2.0, size=(n, 3))
    return raw, target


def test_parameter_count():
    assert nrc.param_count() == 96 * 64 + 64 + 4 * (64 * 64 + 64) + 64 * 16 + 16 == 23888


def test_sh_basis_is_orthonormal():
    # Monte-Carlo-free check: a fine Fibonacci sphere quadrature of Y_i Y_j
    n = 200000
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    phi = math.pi * (1 + 5 ** 0.5) * k
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    Y = nrc.sh16(d)
    G = Y.T @ Y * (4 * math.pi / n)
    assert np.allclose(G, np.eye(16), atol=2e-3)


def test_encoding_layout():
    raw, _ = batch(4, 1)
    e = nrc.encode(raw)
    assert e.shape == (4, 96)
    assert np.allclose(e[:, 0], np.sin(math.pi * raw[:, 0]))
    assert np.allclose(e[:, 2 * 11 + 1], np.cos(2 ** 11 * math.pi * raw[:, 0]))
    assert np.allclose(e[:, 24], np.sin(math.pi * raw[:, 1]))
    assert np.allclose(e[:, 88:91], raw[:, 8:11]) and np.allclose(e[:, 91], raw[:, 3])
    assert np.all(e[:, 92:] == 0)


def test_backward_matches_central_finite_differences():
    raw, target = batch(8, 2)
    theta = nrc.init(3).astype(np.float64)
    g = nrc.backward(theta, raw, target)
    rng = np.random.default_rng(4)
    idx = rng.choice(theta.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        # the loss's denominator is held constant: differentiate with the denominator of theta
        y0, _, _ = nrc.forward(theta, raw)
        den = y0[:, :3] ** 2 + nrc.EPS_LOSS
        f = lambda t: float(np.mean(np.sum((nrc.forward(t, raw)[0][:, :3] - target) ** 2 / den, axis=1)))
        fd = (f(tp) - f(tm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_adam_first_step_closed_form():
    # after one step, m_hat = g and v_hat = g^2, so the update is lr * g / (|g| + eps) = lr * sign(g)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    th, m, v = nrc.adam(np.zeros(4), np.zeros(4), np.zeros(4), g, 1, lr=1e-2)
    assert np.allclose(th, -1e-2 * g / (np.abs(g) + nrc.EPS_ADAM))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.01 * g * g)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.140625, 65504.0], np.float32)
    r = nrc.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -3.140625
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_training_on_the_oracle_reduces_the_loss():
    # a smooth target the network can fit: gradient descent with Adam lowers the loss
    raw, _ = batch(512, 5)
    target = 0.5 + 0.4 * np.sin(3 * raw[:, 0:3]) * raw[:, 4:7]
    theta = nrc.init(6).astype(np.float64)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    l0 = nrc.loss(nrc.forward(theta, raw)[0], target)
    for step in range(1, 101):
        theta, m, v = nrc.adam(theta, m, v, nrc.backward(theta, raw, target), step, lr=3e-3)
    assert nrc.loss(nrc.forward(theta, raw)[0], target) < 0.2 * l0
