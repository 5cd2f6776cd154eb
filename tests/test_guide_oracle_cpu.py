"""The oracle of the learned directional proposal (oracle/guide.py), pinned
by mathematics: bfloat16 rounding known answers, vMF normalisation by
quadrature, a chi-square test of the Jakob 2012 sampler against the
density, analytic gradients against central finite differences, and the
first Adam step in closed form."""
import math

import numpy as np
import pytest

from oracle import guide as G


def test_bf16_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40, 0.0, 65504.0], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, -2.5, 2.9975715e38, 1e-40, 0.0, 65536.0], np.float32)
    got = G.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0          # tie to even
    assert got[2] == np.float32(1 + 2 ** -6)        # tie to even (odd neighbour below)
    assert got[3] == -2.5 and got[6] == 0.0 and got[7] == 65536.0
    assert (got.view(np.uint32) & 0xFFFF).max() == 0


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 100.0, 1000.0])
def test_vmf_normalised(kappa):
    # integrate in (t = mu.w, phi) with nodes dense near t = 1
    u = np.linspace(0, 1, 200001)
    t = 1.0 - 2.0 * u ** 4
    dt = np.gradient(t)
    mu = np.array([0.3, -0.2, 0.9])
    mu /= np.linalg.norm(mu)
    tan, bit = G.frame_of(mu[None])
    d = t[:, None] * mu + np.sqrt(np.maximum(0, 1 - t * t))[:, None] * tan[0]
    f = G.vmf_pdf(kappa, mu, d)
    integral = abs(np.sum(f * dt)) * 2 * np.pi
    assert integral == pytest.approx(1.0, rel=1e-4)


@pytest.mark.parametrize("kappa", [0.5, 8.0, 200.0])
def test_vmf_sampler_chi_square(kappa):
    from scipy import stats
    rng = np.random.default_rng(int(kappa * 10))
    n = 200_000
    mu = np.tile(np.array([[0.48, 0.6, 0.64]]), (n, 1))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    d = G.vmf_sample(np.full(n, kappa), mu, rng.random(n), rng.random(n))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    t = d @ mu[0]
    # bins of equal probability in t: CDF(t) = (exp(k(t-1)) - exp(-2k)) / (1 - exp(-2k))
    nb = 40
    edges_p = np.linspace(0, 1, nb + 1)
    edges_t = 1.0 + np.log(edges_p + (1 - edges_p) * np.exp(-2 * kappa)) / kappa
    counts, _ = np.histogram(t, bins=np.clip(edges_t, -1, 1))
    chi2 = np.sum((counts - n / nb) ** 2 / (n / nb))
    assert stats.chi2.sf(chi2, nb - 1) > 1e-3
    # azimuth uniform around mu
    tan, bit = G.frame_of(mu[:1])
    phi = np.arctan2(d @ bit[0], d @ tan[0])
    counts, _ = np.histogram(phi, bins=36, range=(-np.pi, np.pi))
    chi2 = np.sum((counts - n / 36) ** 2 / (n / 36))
    assert stats.chi2.sf(chi2, 35) > 1e-3


def test_frame_is_orthonormal():
    rng = np.random.default_rng(1)
    mu = rng.normal(size=(1000, 3))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    mu[0] = [0, 0, -1]
    mu[1] = [0, 0, 1]
    t, b = G.frame_of(mu)
    for a, c in ((t, b), (t, mu), (b, mu)):
        assert np.abs(np.sum(a * c, axis=1)).max() < 1e-12
    for a in (t, b):
        assert np.allclose(np.linalg.norm(a, axis=1), 1.0)


def _batch(seed, n=64):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    x = G.features(pos, wo, nn, rng.random(n))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.random(n) * 2
    return x, d, c


def test_gradients_match_finite_differences():
    theta = G.init_params(3).astype(np.float64)
    theta += np.random.default_rng(4).normal(scale=0.05, size=theta.size)   # non-zero biases / larger last layer
    x, d, c = _batch(5)
    _, g = G.loss_and_grad(theta, x, d, c, exact=True)
    rng = np.random.default_rng(6)
    idx = np.concatenate([rng.choice(theta.size, 60, replace=False), [theta.size - 1, theta.size - 49]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(theta[i]))
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        lp, _ = G.loss_and_grad(tp, x, d, c, exact=True)
        lm, _ = G.loss_and_grad(tm, x, d, c, exact=True)
        fd = (lp - lm) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=2e-4, abs=1e-8), i


def test_loss_is_nll_of_mixture():
    theta = G.init_params(7)
    x, d, c = _batch(8, 16)
    loss, _ = G.loss_and_grad(theta, x, d, c)
    w, k, mu, _, _ = G.decode(G.forward(theta, x))
    q = G.mixture_pdf(w, k, mu, d)
    assert loss == pytest.approx(np.sum(-c * np.log(q)) / c.sum(), rel=1e-12)
    assert np.allclose(w.sum(axis=1), 1.0)


def test_adam_first_step_closed_form():
    theta = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -4.0, 0.0])
    opt = G.Adam(3, lr=1e-2)
    t1 = opt.step(theta, g)
    # first bias-corrected step: m_hat = g, v_hat = g^2 -> theta - lr * sign(g)
    assert t1[0] == pytest.approx(1.0 - 1e-2, abs=1e-7)
    assert t1[1] == pytest.approx(-2.0 + 1e-2, abs=1e-7)
    assert t1[2] == np.float32(0.5)


def test_features_layout():
    x = G.features(np.array([[0.25, 0.5, 0.0]]), np.array([[0, 0, 1.0]]), np.array([[1.0, 0, 0]]), np.array([0.3]))
    assert x.shape == (1, 96)
    r = math.sqrt(0.5)
    assert x[0, 0] == pytest.approx(r) and x[0, 1] == pytest.approx(r)       # x: sin, cos of pi/4
    assert x[0, 2] == pytest.approx(1.0) and abs(x[0, 3]) < 1e-7              # k = 1: pi/2
    assert x[0, 24] == pytest.approx(1.0) and abs(x[0, 25]) < 1e-7            # y: pi/2
    assert x[0, 48] == 0.0 and x[0, 49] == 1.0                                # z = 0
    assert list(x[0, 72:79]) == [0, 0, 1, 1, 0, 0, np.float32(0.3)]
    assert not x[0, 79:].any()
