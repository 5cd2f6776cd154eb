"""Parity of the HIP wavefront integrator with the CPU oracle.

Per-path: with identical per-(pixel, sample) pcg32 streams every path's
radiance must match the oracle's bit-for-bit (the same binary32 arithmetic,
binary64 transcendentals); a path may only diverge where binary64 libm and
the device's binary64 routines round differently and that difference
crosses a binary32 rounding boundary, so the required bit-exact fraction is
0.999.  Per-image: relMSE against the oracle render (same streams)."""
import numpy as np
import pytest

import drawlab_amd
import oracle
from drawlab_amd import scenes

pytestmark = pytest.mark.gpu


from helpers import assert_unbiased, paths_parity as _paths, relmse, tile_z


def test_cornell_box_paths_bitexact(backend):
    _paths(backend, scenes.cornell_box(64, 64, spp=4, max_depth=8, rr_depth=3), range(4), seed=11)


def test_cornell_box_unlimited_depth_paths(backend):
    _paths(backend, scenes.cornell_box(48, 48, spp=2, max_depth=-1, rr_depth=4), range(2), seed=5,
           builder=drawlab_amd.BUILD_HOST_SAH)


def test_mis_plates_paths_bitexact(backend):
    _paths(backend, scenes.mis_plates(64, 36, max_depth=6), range(3), seed=3)


def test_envmap_sphere_paths_bitexact(backend):
    d = scenes.displaced_sphere(width=64, height=36, grid=128, glass_level=3, env_size=(128, 64), max_depth=16)
    _paths(backend, d, range(3), seed=9)


def test_random_scene_paths_bitexact(backend):
    _paths(backend, scenes.random_scene(7, 20000, width=48, height=48), range(2), seed=2)


def test_cornell_box_image_relmse(backend):
    desc = scenes.cornell_box(64, 64, spp=256, max_depth=8, rr_depth=3)
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    img, st = backend.render(spp=256, seed=1)
    osc = oracle.from_desc(desc)
    ref, ost = osc.render(256, seed=1, max_depth=8, rr_depth=3)
    e = relmse(img, ref)
    print("relMSE %.3e  gpu %.1f ms  stats %s" % (e, st["ms_total"], st), flush=True)
    assert e <= 1e-4
    assert st["paths"] == 64 * 64 * 256


def test_progressive_equals_single_pass(backend):
    """Splitting the samples over passes changes nothing but summation order."""
    desc = scenes.cornell_box(32, 32, spp=16, max_depth=8, rr_depth=3)
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    a, _ = backend.render(spp=16, seed=4, spp_per_pass=16)
    b, _ = backend.render(spp=16, seed=4, spp_per_pass=3)
    assert relmse(a, b) < 1e-10


def test_tile_bias_vs_oracle(backend):
    """Per 16x16 tile and channel, the GPU image agrees with the oracle's
    within the sampling error estimated from independent batches (disjoint
    seeds, so the two sides share no random streams)."""
    desc = scenes.cornell_box(64, 64, spp=64, max_depth=8, rr_depth=3)
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    gpu = [backend.render(spp=256, seed=1000 + b)[0] for b in range(16)]
    osc = oracle.from_desc(desc)
    cpu = [osc.render(32, seed=2000 + b, max_depth=8, rr_depth=3)[0] for b in range(16)]
    assert_unbiased(tile_z(gpu, cpu))


_FILTERS = [(drawlab_amd.FILTER_BOX, 0.5, 0.0, 0.0), (drawlab_amd.FILTER_TENT, 1.0, 0.0, 0.0),
            (drawlab_amd.FILTER_GAUSSIAN, 2.0, 0.5, 0.0), (drawlab_amd.FILTER_MITCHELL, 2.0, 1 / 3, 1 / 3),
            (drawlab_amd.FILTER_GAUSSIAN, 3.3, 1.0, 0.0), (drawlab_amd.FILTER_GAUSSIAN, 4.2, 1.2, 0.0),
            (drawlab_amd.FILTER_GAUSSIAN, 0.3, 0.2, 0.0),
            (drawlab_amd.FILTER_BOX, float(np.float32(1.4999999)), 0.0, 0.0)]


@pytest.mark.parametrize("filt", _FILTERS, ids=lambda f: "t%d_r%.9g" % (f[0], f[1]))
def test_reconstruction_filter_image_vs_oracle(backend, filt):
    """Every footprint size of the tiled film gather (filter reach 0 to 4
    pixels, a radius just below r + 1/2 = 2, and the Mitchell filter's
    negative lobes) on an image that is not a whole number of tiles: the
    image equals the oracle's ImageBlock accumulation of the same paths up to
    summation order."""
    desc = scenes.cornell_box(41, 29, spp=4, max_depth=5, rr_depth=3)
    desc.rfilter = scenes.RFilter(*filt)
    backend.upload_scene(desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    img, _ = backend.render(spp=4, seed=6)
    osc = oracle.from_desc(desc)
    ref, _ = osc.render(4, seed=6, max_depth=5, rr_depth=3)
    e = relmse(img, ref)
    print("filter %s relMSE %.3e" % (filt, e), flush=True)
    assert e <= 1e-8
