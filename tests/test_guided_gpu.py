"""Rendering with the learned proposal (DL_RENDER_GUIDED): one-sample MIS of
the vMF mixture and BSDF sampling is unbiased -- per 16x16 tile and channel
the guided image agrees with the plain one within its sampling error, both
with online training and with a frozen random network -- a frozen
proposal renders deterministically, and a trained network lowers the error
at equal samples where incident light is concentrated."""
import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import scenes
from drawlab_amd.scene import Bsdf
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
    img, st = backend.render(spp=16, seed=1, guided=True, spp_per_pass=4, guide_steps=3)
    th1, _ = backend.guide_params()
    assert np.isfinite(img).all() and np.isfinite(th1).all()
    assert not np.array_equal(th0, th1)
    assert st["train_steps"] == 12               # four passes, three steps after each


def _rough_materials_box():
    """Cornell box whose walls and boxes use every guided BSDF type: rough
    conductor, anisotropic GGX with a diffuse part, microfacet with kd."""
    d = scenes.cornell_box(64, 64, spp=32, max_depth=6, rr_depth=3)
    cond = d.add_bsdf(Bsdf.conductor(0.5))
    aniso = d.add_bsdf(Bsdf.aniso_ggx((0.3, 0.3, 0.3), (0.6, 0.6, 0.6), (0.5, 0.3)))
    micro = d.add_bsdf(Bsdf.microfacet(0.2, kd=(0.4, 0.3, 0.2)))
    light_free = [m for m in d.meshes if m.radiance is None]
    for i, m in enumerate(light_free):
        m.bsdf = (cond, aniso, micro)[i % 3]
    return d


def test_guided_rough_materials_unbiased(backend):
    """Proposal sampling on conductor / anisotropic GGX / microfacet
    surfaces (a frozen random network and a trained one) agrees with plain
    rendering per tile."""
    desc = _rough_materials_box()
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    plain, _ = _batches(backend, 16, 32, 500)
    backend.guide_init(seed=4)
    frozen, st = _batches(backend, 16, 32, 600, guided=True, guide_frozen=True)
    assert st["ms_guide"] > 0
    assert_unbiased(tile_z(frozen, plain))
    backend.guide_init(seed=5)
    trained, st = _batches(backend, 16, 32, 700, guided=True)
    assert st["train_steps"] > 0
    assert_unbiased(tile_z(trained, plain))


def _relmse(x, y):
    return float(np.mean((x.astype(np.float64) - y) ** 2 / (y.astype(np.float64) ** 2 + 1e-2)))


def test_trained_network_gains_over_plain_sampling(backend):
    """What the learned proposal buys where directions worth sampling are
    concentrated (a room lit only through a small ceiling patch): at equal
    samples per pixel, relMSE against a 16384-spp plain reference drops by
    more than 2x with a network trained on one earlier render, and by more
    than 1.5x when the render trains its own network from scratch."""
    d = scenes.chimney_room(64, 64, max_depth=6, rr_depth=3)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    ref, _ = backend.render(spp=16384, seed=999)
    plain = np.mean([_relmse(backend.render(spp=256, seed=s)[0], ref) for s in range(4)])
    backend.guide_init(seed=1)
    backend.render(spp=1024, seed=50, guided=True)             # trains the network
    frozen = np.mean([_relmse(backend.render(spp=256, seed=100 + s, guided=True, guide_frozen=True)[0], ref)
                      for s in range(4)])
    online = []
    for s in range(2):
        backend.guide_init(seed=2 + s)
        online.append(_relmse(backend.render(spp=256, seed=200 + s, guided=True)[0], ref))
    online = np.mean(online)
    print("relMSE at 256 spp: plain %.4g, trained network %.4g, trained online %.4g" % (plain, frozen, online),
          flush=True)
    assert frozen < plain / 2
    assert online < plain / 1.5


def test_guided_on_the_displaced_sphere_workload_unbiased(backend):
    """The learned proposal on the 2,097,152-triangle displaced-sphere scene
    (conductor, glass, ground under the sky and sun; reduced resolution):
    independent guided batches, each training the network online as a
    render does, agree per tile with independent plain batches."""
    d = scenes.displaced_sphere(width=96, height=64, max_depth=16)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    plain, _ = _batches(backend, 16, 256, 800)
    backend.guide_init(seed=7)
    guided, st = _batches(backend, 16, 256, 900, guided=True)
    print("guided stats", st, flush=True)
    assert st["train_steps"] > 0 and st["ms_guide"] > 0
    for img in guided:
        assert np.isfinite(img).all()
    assert_unbiased(tile_z(guided, plain))
