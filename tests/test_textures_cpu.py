"""Bitmap-texture sampler of the oracle (the reference GPU path's
normalized-coordinate, bilinear, wrap-addressed lookup,
src/optix/host/cuda_texture.cpp:57-82): texel centres return the texel,
midpoints the average of neighbours, wrap addressing repeats with period
1, and random lookups equal a float64 closed form of bilinear filtering."""
import numpy as np

import oracle
from drawlab_amd import scenes


def _bilinear(rgb, uv):
    h, w, _ = rgb.shape
    x = uv[:, 0].astype(np.float64) * w - 0.5
    y = uv[:, 1].astype(np.float64) * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    ax, ay = (x - x0)[:, None], (y - y0)[:, None]
    xi, yi = x0.astype(np.int64) % w, y0.astype(np.int64) % h
    xj, yj = (xi + 1) % w, (yi + 1) % h
    t = rgb.astype(np.float64)
    return ((t[yi, xi] * (1 - ax) + t[yi, xj] * ax) * (1 - ay) + (t[yj, xi] * (1 - ax) + t[yj, xj] * ax) * ay)


def test_texel_centres_and_midpoints():
    rgb = scenes.checker_texture(6, 4, seed=5, n=2)
    h, w, _ = rgb.shape
    jj, ii = np.mgrid[0:h, 0:w]
    uv = np.stack([(ii + 0.5) / w, (jj + 0.5) / h], -1).reshape(-1, 2).astype(np.float32)
    got = oracle.texture_sample(rgb, uv).reshape(h, w, 3)
    assert np.array_equal(got, rgb)
    # halfway between texel (i, j) and (i + 1, j), wrapping at the right edge
    uv = np.stack([(ii + 1.0) / w, (jj + 0.5) / h], -1).reshape(-1, 2).astype(np.float32)
    got = oracle.texture_sample(rgb, uv).reshape(h, w, 3)
    want = 0.5 * rgb + 0.5 * np.roll(rgb, -1, axis=1)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-7)


def test_wrap_period_and_closed_form():
    rgb = scenes.checker_texture(13, 7, seed=6, n=3)
    rng = np.random.default_rng(0)
    uv = rng.uniform(-3, 4, size=(4000, 2)).astype(np.float32)
    got = oracle.texture_sample(rgb, uv)
    assert np.allclose(got, _bilinear(rgb, uv), rtol=2e-5, atol=2e-6)
    # dyadic coordinates shifted by whole periods land on the same texels
    base = (rng.integers(0, 64, size=(500, 2)) / 64.0).astype(np.float32)
    a = oracle.texture_sample(rgb, base)
    b = oracle.texture_sample(rgb, base + np.float32(2.0))
    c = oracle.texture_sample(rgb, base - np.float32(1.0))
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_non_finite_coordinates_give_zero():
    rgb = scenes.checker_texture(4, 4, seed=7)
    uv = np.array([[np.nan, 0.5], [0.5, np.inf], [-np.inf, 0.1]], np.float32)
    assert not oracle.texture_sample(rgb, uv).any()


def test_textured_scene_renders_on_oracle():
    d = scenes.textured_box(24, 24)
    img, _ = oracle.from_desc(d).render(2, seed=3, max_depth=4, rr_depth=3)
    assert np.isfinite(img).all() and img.mean() > 0
