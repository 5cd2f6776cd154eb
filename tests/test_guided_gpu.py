"""Rendering with the learned proposal (DL_RENDER_GUIDED): one-sample MIS of
the vMF mixture and BSDF sampling is unbiased -- per 16x16 tile and channel
the guided image agrees with the plain one within its sampling error, both
with online training and with a frozen random network -- and a frozen
proposal renders deterministically."""
import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import scenes
from helpers import assert_unbiased, tile_z

pytestmark = pytest.mark.gpu


def _batches(backend, k, spp, seed0, **kw):
    out = []
    for b in range(k):
        img, st = backend.render(spp=spp, seed=seed0 + b, **kw)
        out.append(img)
    return out, st


def test_guided_unbiased_with_training(backend):
    desc = scenes.cornell_box(128, 128, spp=64, max_depth=8, rr_depth=3)
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    plain, _ = _batches(backend, 24, 64, 100)
    backend.guide_init(seed=3)
    guided, st = _batches(backend, 24, 64, 200, guided=True)
    print("guided stats", st, flush=True)
    assert st["ms_guide"] > 0
    for img in guided:
        assert np.isfinite(img).all()
    assert_unbiased(tile_z(guided, plain))


def test_guided_frozen_random_network_unbiased(backend):
    """An untrained (random) proposal is a poor but valid sampling density:
    the estimate stays unbiased."""
    desc = scenes.mis_plates(64, 36, max_depth=6)
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    plain, _ = _batches(backend, 16, 32, 300)
    backend.guide_init(seed=9)
    guided, _ = _batches(backend, 16, 32, 400, guided=True, guide_frozen=True)
    assert_unbiased(tile_z(guided, plain))
    a, _ = backend.render(spp=8, seed=5, guided=True, guide_frozen=True)
    b, _ = backend.render(spp=8, seed=5, guided=True, guide_frozen=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_guided_env_scene_trains(backend):
    d = scenes.displaced_sphere(width=96, height=54, grid=128, glass_level=3, env_size=(128, 64), max_depth=8)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    backend.guide_init(seed=1)
    th0, _ = backend.guide_params()
    img, st = backend.render(spp=16, seed=1, guided=True, spp_per_pass=4)
    th1, _ = backend.guide_params()
    assert np.isfinite(img).all() and np.isfinite(th1).all()
    assert not np.array_equal(th0, th1)          # four passes -> four training steps
