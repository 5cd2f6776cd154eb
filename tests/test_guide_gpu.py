"""The learned proposal's network on the matrix cores against the oracle
(oracle/guide.py): forward outputs and mixtures, the loss and its gradient
with respect to all 21,808 parameters, the Adam update, and that online
training fits a known directional distribution."""
import numpy as np
import pytest

from oracle import guide as G

pytestmark = pytest.mark.gpu


def _inputs(seed, n):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3))
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1, keepdims=True)
    return G.features(pos, wo, nn, rng.random(n)), rng


def _theta(seed):
    th = G.init_params(seed)
    th = th + np.random.default_rng(seed + 1).normal(scale=0.05, size=th.size).astype(np.float32)
    return th.astype(np.float32)


def _close(got, ref, q=0.999, tol=2e-2):
    scale = np.sqrt(np.mean(ref ** 2)) + 1e-12
    err = np.abs(got - ref) / (np.abs(ref) + 0.05 * scale)
    return float(np.quantile(err, q)) < tol and float(err.max()) < 0.5, float(np.quantile(err, q)), float(err.max())


def test_forward_matches_oracle(backend):
    th = _theta(11)
    x, _ = _inputs(12, 1000)            # not a multiple of the 64-row tile
    backend.guide_init(params=th)
    out, mix = backend.guide_eval(x)
    ref = G.forward(th, x)
    ok, q, mx = _close(out, ref)
    print("forward rel err q999 %.3e max %.3e" % (q, mx), flush=True)
    assert ok
    w, k, mu, _, _ = G.decode(out.astype(np.float64))
    m = mix.reshape(-1, 8, 5)
    assert np.allclose(m[:, :, 0], w, rtol=1e-5, atol=1e-7)
    assert np.allclose(m[:, :, 1], k, rtol=1e-5)
    assert np.allclose(m[:, :, 2:5], mu, atol=1e-6)


def test_gradient_matches_oracle(backend):
    th = _theta(21)
    x, rng = _inputs(22, 4096)
    d = rng.normal(size=(4096, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.random(4096) * 3
    backend.guide_init(params=th)
    loss = backend.guide_train_step(x, d, c, apply=False)
    _, g = backend.guide_params()
    rl, rg = G.loss_and_grad(th, x, d, c)
    print("loss gpu %.6f oracle %.6f" % (loss, rl), flush=True)
    assert loss == pytest.approx(rl, rel=2e-3)
    cos = float(np.dot(g, rg) / (np.linalg.norm(g) * np.linalg.norm(rg)))
    print("gradient cosine %.6f" % cos, flush=True)
    assert cos > 0.999
    k = 0
    for i, o in G.LAYERS:
        for size in (o * i, o):
            a, b = g[k:k + size], rg[k:k + size]
            rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
            assert rel < 0.05, (i, o, size, rel)
            k += size


def test_adam_step_matches_oracle(backend):
    th = _theta(31)
    x, rng = _inputs(32, 2048)
    d = rng.normal(size=(2048, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.random(2048)
    backend.guide_init(params=th, lr=1e-2)
    backend.guide_train_step(x, d, c, apply=True)
    th1, gres = backend.guide_params()
    assert not gres.any()                 # gradient buffer cleared by the update
    _, rg = G.loss_and_grad(th, x, d, c)
    ref = G.Adam(th.size, lr=1e-2).step(th, rg)
    moved = np.abs(rg) > 1e-6
    agree = np.isclose(th1, ref, atol=2e-4)
    print("adam agreement %.5f" % agree[moved].mean(), flush=True)
    assert agree[moved].mean() > 0.99


def test_training_fits_known_distribution(backend):
    """Directions drawn from a vMF whose mean depends on position: online
    training must raise the likelihood of held-out samples well above the
    uniform density 1/(4 pi)."""
    rng = np.random.default_rng(41)

    def batch(n):
        pos = rng.random((n, 3))
        mu = np.stack([np.cos(2 * np.pi * pos[:, 0]), np.sin(2 * np.pi * pos[:, 0]), np.zeros(n)], 1)
        x = G.features(pos, np.tile([0, 0, 1.0], (n, 1)), np.tile([0, 1.0, 0], (n, 1)), np.full(n, 0.5))
        d = G.vmf_sample(np.full(n, 20.0), mu, rng.random(n), rng.random(n))
        return x, d
    backend.guide_init(seed=5, lr=1e-2)
    losses = []
    for it in range(300):
        x, d = batch(4096)
        losses.append(backend.guide_train_step(x, d, np.ones(4096)))
    xt, dt = batch(4096)
    _, mix = backend.guide_eval(xt)
    m = mix.reshape(-1, 8, 5).astype(np.float64)
    q = G.mixture_pdf(m[:, :, 0], m[:, :, 1], m[:, :, 2:5], dt)
    print("loss first %.3f last %.3f  mean held-out log q %.3f" % (losses[0], losses[-1], np.log(q).mean()),
          flush=True)
    assert losses[-1] < losses[0] - 1.0
    assert np.log(q).mean() > np.log(1 / (4 * np.pi)) + 1.5
