"""Pins of the cost model's FP64 oracle (oracle/mlp_ref.py): the
backpropagated gradient against central finite differences of the loss and
against closed forms, and Adam's first step against its known shape."""
import numpy as np
import pytest

from oracle import mlp_ref as M


def problem(seed, n=24, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAMS)
    p[M.W1:M.B1] = rng.uniform(-0.5, 0.5, M.H * M.F) * scale
    p[M.B1:M.W2] = rng.uniform(-0.2, 0.2, M.H) * scale
    p[M.W2:M.B2] = rng.uniform(-0.3, 0.3, M.H * M.H) * scale
    p[M.B2:M.W3] = rng.uniform(-0.2, 0.2, M.H) * scale
    p[M.W3:M.B3] = rng.uniform(-0.5, 0.5, M.H) * scale
    p[M.B3] = 0.3
    x = rng.normal(size=(n, M.F)) * 2.0 + 1.0
    t = rng.normal(size=n)
    mu = rng.normal(size=M.F)
    sigma = rng.uniform(0.5, 2.0, M.F)
    return p, x, t, mu, sigma


def test_gradient_matches_central_differences():
    p, x, t, mu, sigma = problem(3)
    _, g = M.grad(p, x, t, mu, sigma)
    rng = np.random.default_rng(5)
    # every block of the layout: W1, b1, W2, b2, w3, b3
    idx = np.concatenate([rng.choice(np.arange(a, b), size=min(40, b - a), replace=False)
                          for a, b in ((M.W1, M.B1), (M.B1, M.W2), (M.W2, M.B2), (M.B2, M.W3), (M.W3, M.B3),
                                       (M.B3, M.NPARAMS))])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss(pp, x, t, mu, sigma) - M.loss(pm, x, t, mu, sigma)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_zero_network_closed_form():
    # all weights zero, b3 = c: y = c everywhere; dL/db3 = 2 mean(c - t), the
    # rest vanishes (tanh(0) = 0 and w3 = 0 stop every other path)
    _, x, t, mu, sigma = problem(4)
    p = np.zeros(M.NPARAMS)
    p[M.B3] = 0.7
    loss, g = M.grad(p, x, t, mu, sigma)
    assert loss == pytest.approx(np.mean((0.7 - t) ** 2), rel=1e-14)
    assert g[M.B3] == pytest.approx(2 * np.mean(0.7 - t), rel=1e-14)
    assert np.all(g[:M.B3] == 0.0)


def test_w3_gradient_closed_form():
    # dL/dw3 = h2^T g with g = 2(y - t)/n: the last layer is linear in w3
    p, x, t, mu, sigma = problem(6)
    y, (_, _, h2) = M.forward(p, x, mu, sigma)
    _, g = M.grad(p, x, t, mu, sigma)
    assert np.allclose(g[M.W3:M.B3], h2.T @ (2 * (y - t) / len(t)), rtol=1e-13, atol=1e-15)


def test_standardisation_is_applied():
    # shifting x and mu together, or scaling (x - mu) and sigma together,
    # leaves the network unchanged
    p, x, t, mu, sigma = problem(7)
    y0, _ = M.forward(p, x, mu, sigma)
    y1, _ = M.forward(p, x + 5.0, mu + 5.0, sigma)
    y2, _ = M.forward(p, mu + 3.0 * (x - mu), mu, 3.0 * sigma)
    assert np.allclose(y0, y1, rtol=1e-12) and np.allclose(y0, y2, rtol=1e-12)


def test_adam_first_step_is_lr_sign():
    rng = np.random.default_rng(8)
    g = rng.normal(size=50)
    p = rng.normal(size=50)
    p1, m, v = M.adam_step(p, np.zeros(50), np.zeros(50), g, 1, lr=0.01, eps=0.0)
    assert np.allclose(p1, p - 0.01 * np.sign(g), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_planted_transpose_is_caught():
    # a W2 gradient computed with the transposed product differs from the
    # finite-difference value: the pin above would catch that mistake
    p, x, t, mu, sigma = problem(9)
    _, g = M.grad(p, x, t, mu, sigma)
    wrong = g[M.W2:M.B2].reshape(M.H, M.H).T.ravel()
    i = M.W2 + 1   # W2[0][1]
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (M.loss(pp, x, t, mu, sigma) - M.loss(pm, x, t, mu, sigma)) / (2 * h)
    assert abs(fd - g[i]) < 1e-7 + 1e-6 * abs(fd)
    assert abs(fd - wrong[1]) > 1e-4 * max(abs(fd), 1e-6)
