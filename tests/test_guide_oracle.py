"""Pins the FP64 guiding-network oracle: finite differences of the loss
against its analytic gradient, closed forms of the vMF lobe, and Adam
against its definition."""
import math

import numpy as np
import pytest

from oracle import guide_oracle as G


def _batch(B, seed):
    rng = np.random.default_rng(seed)
    pos = rng.random((B, 3))
    n = rng.normal(size=(B, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.normal(size=(B, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    om = rng.normal(size=(B, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    pb = rng.uniform(0.05, 1.0, B)
    w = rng.uniform(0.0, 2.0, B)
    return G.features(pos, n, wo), om, pb, w


def test_parameter_count():
    assert G.N_PARAMS == 16624
    assert len(G.init_params(0)) == G.N_PARAMS


def test_gradient_matches_finite_differences():
    x, om, pb, w = _batch(24, 1)
    rng = np.random.default_rng(2)
    theta = G.init_params(3).astype(np.float64) + rng.normal(0, 0.05, G.N_PARAMS)
    L, g, _ = G.loss_and_grad(theta, x, om, pb, w)
    idx = rng.choice(G.N_PARAMS, 300, replace=False)
    for i in idx:
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (G.loss_and_grad(tp, x, om, pb, w)[0] - G.loss_and_grad(tm, x, om, pb, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)) + 1e-4 * abs(fd), (i, fd, g[i])


def test_head_gradient_matches_finite_differences():
    x, om, pb, w = _batch(16, 4)
    theta = G.init_params(5).astype(np.float64)
    y, _ = G.forward(theta, x)
    y = y + np.random.default_rng(6).normal(0, 1.0, y.shape)  # spread kappa / logits
    L, dy = G.loss_and_grad_y(y, om, pb, w)
    for b in range(16):
        for j in range(41):
            h = 1e-6
            yp, ym = y.copy(), y.copy()
            yp[b, j] += h
            ym[b, j] -= h
            fd = (G.loss_and_grad_y(yp, om, pb, w)[0] - G.loss_and_grad_y(ym, om, pb, w)[0]) / (2 * h)
            assert abs(fd - dy[b, j]) <= 1e-7 + 1e-4 * abs(fd), (b, j, fd, dy[b, j])


def test_vmf_is_a_density_and_reduces_to_uniform():
    rng = np.random.default_rng(7)
    v = rng.normal(size=(200000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    mu = np.array([[[0.0, 0.6, 0.8]]])
    for kappa in (0.5, 5.0, 50.0):
        dens = G.vmf(np.repeat(mu, len(v), 0), np.full((len(v), 1), kappa), v)[:, 0]
        assert dens.mean() * 4 * math.pi == pytest.approx(1.0, rel=0.05 if kappa < 20 else 0.15)
    tiny = G.vmf(mu, np.array([[1e-6]]), np.array([[1.0, 0.0, 0.0]]))[0, 0]
    assert tiny == pytest.approx(1 / (4 * math.pi), rel=1e-4)


def test_adam_step_definition():
    th = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 0.0])
    t1, m, v = G.adam_step(th, np.zeros(3), np.zeros(3), g, 1, lr=0.01)
    # first bias-corrected step moves by lr * g / (|g| + eps)
    assert np.allclose(t1, th - 0.01 * g / (np.abs(g) + 1e-8))
    t2, m2, v2 = G.adam_step(t1, m, v, g, 2, lr=0.01)
    assert np.allclose(m2, 0.9 * m + 0.1 * g) and np.allclose(v2, 0.999 * v + 0.001 * g * g)


def test_mixture_alpha_range():
    y = np.zeros((4, 48))
    y[:, 40] = [-50, -1, 1, 50]
    a = G.head(y)["alpha"]
    assert np.all(a >= 0.05 - 1e-12) and np.all(a <= 0.95 + 1e-12)
