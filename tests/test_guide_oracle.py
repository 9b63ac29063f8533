"""Pins the path-guiding oracle (oracle/guide_oracle.py) by what the mathematics fixes:
finite differences of the loss, normalisation of the vMF mixture, agreement of the
sampler with the density, and Adam against its closed form on a quadratic."""
import numpy as np
import pytest

from oracle import guide_oracle as G


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, G.IN))
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    weight = rng.exponential(size=n)
    return x, w, weight


def test_layout():
    assert G.NPARAMS == 29480
    assert G.OFFSETS[0] == (0, 14336) and G.OFFSETS[4] == (26880, 29440)


def test_encoding_is_one_blob():
    x = np.array([[0.0, 0.5, 1.0, 0.25, 0.015625, 0.3, 0.7]])
    e = G.encode(x).reshape(G.IN, G.BINS)
    # the peak bin holds the value, neighbours fall off as a unit-variance Gaussian in bins
    assert np.argmax(e[1]) in (15, 16) and abs(e[1][15] - np.exp(-0.125)) < 1e-12
    assert abs(e[4][0] - 1.0) < 1e-12 and abs(e[4][1] - np.exp(-0.5)) < 1e-12
    assert abs(e[0][0] - np.exp(-0.125)) < 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(100 + seed)
    th = G.init_params(seed)
    # larger output layer so every lobe parameter matters
    th[G.OFFSETS[4][0]:G.OFFSETS[4][1]] *= 4.0
    th[G.OFFSETS[4][1]:] = rng.normal(scale=0.5, size=G.OUT)
    x, w, weight = _batch(24, seed)
    loss, g = G.loss_grad(th, x, w, weight)
    assert abs(loss - G.loss_value(th, x, w, weight)) < 1e-12
    idx = np.concatenate([rng.choice(np.arange(G.OFFSETS[i][0], G.OFFSETS[i][1] + G.SHAPES[i][0]), 12, replace=False)
                          for i in range(5)])
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (G.loss_value(tp, x, w, weight) - G.loss_value(tm, x, w, weight)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_emulated_gradient_close_to_exact():
    th = G.init_params(3)
    x, w, weight = _batch(512, 3)
    _, g = G.loss_grad(th, x, w, weight)
    _, ge = G.loss_grad(th, x, w, weight, emulate=True)
    # bf16 weights change the function itself by ~1%; the gradient follows within a few %
    assert np.linalg.norm(ge - g) / np.linalg.norm(g) < 0.1


def _sphere_quadrature(n_theta=400, n_phi=800):
    ct = np.polynomial.legendre.leggauss(n_theta)
    z, wz = ct
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    Z, P = np.meshgrid(z, ph, indexing="ij")
    s = np.sqrt(1 - Z ** 2)
    d = np.stack([s * np.cos(P), s * np.sin(P), Z], -1).reshape(-1, 3)
    wt = (wz[:, None] * np.full(n_phi, 2 * np.pi / n_phi)[None, :]).ravel()
    return d, wt


def test_mixture_density_integrates_to_one():
    rng = np.random.default_rng(5)
    o = rng.normal(size=(3, G.OUT))
    o[:, 4::5] = np.array([[-3.0], [1.0], [5.0]])  # broad, medium and narrow lobes (kappa ~ 0.05 .. 150)
    d, wt = _sphere_quadrature()
    for i in range(3):
        q = np.exp(G.log_pdf(np.repeat(o[i:i + 1], len(d), 0), d))
        assert abs((q * wt).sum() - 1.0) < 1e-6


def test_single_lobe_closed_form():
    # one dominant lobe: q reduces to the vMF density kappa e^{kappa(cos-1)} / (2 pi (1 - e^{-2 kappa}))
    o = np.zeros((1, G.OUT))
    o[0, 0::5] = -60.0
    o[0, 0] = 0.0
    o[0, 1:4] = [0.0, 0.0, 2.0]
    o[0, 4] = np.log(10.0)
    for c in (1.0, 0.3, -1.0):
        w = np.array([[np.sqrt(1 - c * c), 0.0, c]])
        ref = 10.0 * np.exp(10.0 * (c - 1)) / (2 * np.pi * (1 - np.exp(-20.0)))
        assert abs(np.exp(G.log_pdf(o, w))[0] - ref) < 1e-12 * max(1.0, ref)


def test_sampler_matches_density():
    rng = np.random.default_rng(9)
    o = rng.normal(size=(1, G.OUT))
    o[0, 4::5] = rng.uniform(0.0, 3.0, G.LOBES)
    n = 400000
    d, pdf = G.sample(np.repeat(o, n, 0), rng.random((n, 3)))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert np.allclose(pdf, np.exp(G.log_pdf(np.repeat(o, n, 0), d)))
    # histogram over (cos theta, phi) cells against the integrated density
    nz, nph = 8, 16
    cz = np.clip(((d[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    cp = np.clip(((np.arctan2(d[:, 1], d[:, 0]) / (2 * np.pi) + 0.5) * nph).astype(int), 0, nph - 1)
    hist = np.bincount(cz * nph + cp, minlength=nz * nph) / n
    qd, wt = _sphere_quadrature(200, 400)
    q = np.exp(G.log_pdf(np.repeat(o, len(qd), 0), qd)) * wt
    qz = np.clip(((qd[:, 2] + 1) / 2 * nz).astype(int), 0, nz - 1)
    qp = np.clip(((np.arctan2(qd[:, 1], qd[:, 0]) / (2 * np.pi) + 0.5) * nph).astype(int), 0, nph - 1)
    exp = np.bincount(qz * nph + qp, weights=q, minlength=nz * nph)
    chi2 = (n * (hist - exp) ** 2 / np.maximum(exp, 1e-12)).sum()
    assert chi2 < 200.0, chi2  # 127 degrees of freedom


def test_frame_is_orthonormal():
    rng = np.random.default_rng(2)
    n = rng.normal(size=(1000, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    n[0] = [0, 0, -1]
    n[1] = [0, 0, 1]
    t1, t2 = G.frame(n)
    for a, b in ((t1, t2), (t1, n), (t2, n)):
        assert np.abs((a * b).sum(1)).max() < 1e-12
    for a in (t1, t2):
        assert np.abs(np.linalg.norm(a, axis=1) - 1).max() < 1e-12


def test_adam_closed_form_on_quadratic():
    # first step moves every coordinate by lr * sign(g) (bias-corrected moments)
    th = np.array([1.0, -2.0, 0.5])
    g = 2 * th
    th1, m, v = G.adam(th, np.zeros(3), np.zeros(3), 1, g, lr=0.1)
    assert np.allclose(th1, th - 0.1 * np.sign(g), atol=1e-7)
    # many steps converge to the minimum of |th|^2
    t, m, v = th.copy(), np.zeros(3), np.zeros(3)
    for k in range(1, 2001):
        t, m, v = G.adam(t, m, v, k, 2 * t, lr=0.01)
    assert np.abs(t).max() < 1e-2


def test_training_reduces_loss():
    # fit a fixed target direction field: the loss must fall
    rng = np.random.default_rng(4)
    th = G.init_params(4)
    x = rng.random((256, G.IN))
    w = np.tile([[0.0, 0.0, 1.0]], (256, 1))
    w[:, 0] = 0.3 * (x[:, 0] - 0.5)
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    weight = np.ones(256)
    l0 = G.loss_value(th, x, w, weight)
    m, v = np.zeros_like(th), np.zeros_like(th)
    for k in range(1, 61):
        _, g = G.loss_grad(th, x, w, weight)
        th, m, v = G.adam(th, m, v, k, g, lr=1e-2)
    assert G.loss_value(th, x, w, weight) < l0 - 1.0
