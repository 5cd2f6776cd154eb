"""Error paths and edge cases of the C ABI (include/dl_hip.h) on the GPU:
calls out of order fail with DL_ERR_STATE, malformed scenes and parameters
with DL_ERR_INVALID and the message the header promises -- the context stays
usable afterwards -- and degenerate inputs (zero-area triangles, a
one-triangle scene, empty and inverted ray intervals) give the oracle's
answers."""
import dataclasses

import numpy as np
import pytest

import drawlab_amd
import oracle
from drawlab_amd.scene import Bsdf, Camera, RFilter, SceneDesc, FILTER_BOX
from helpers import compare_hits, paths_parity, planes

pytestmark = pytest.mark.gpu


def _one_triangle():
    d = SceneDesc()
    b = d.add_bsdf(Bsdf.diffuse())
    d.add_mesh(np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.uint32([[0, 1, 2]]), b)
    d.point_lights.append(((0.2, 0.2, 2.0), (5.0, 5.0, 5.0)))
    d.camera = Camera(eye=(0.3, 0.3, 3), lookat=(0.3, 0.3, 0), up=(0, 1, 0), fov=30, width=16, height=16)
    d.rfilter = RFilter(FILTER_BOX)
    return d


def test_calls_out_of_order():
    be = drawlab_amd.Backend(0)
    try:
        with pytest.raises(drawlab_amd.DlError, match="build_accel before upload_scene"):
            be.build_accel(drawlab_amd.BUILD_GPU)
        d = _one_triangle()
        be.upload_scene(d)
        with pytest.raises(drawlab_amd.DlError, match="render before build_accel"):
            be.render(spp=1)
        with pytest.raises(drawlab_amd.DlError, match="trace before build_accel"):
            be.trace_closest(planes(np.zeros((1, 3), np.float32), np.float32([[0, 0, 1]]), 0.0, np.inf))
        be.build_accel(drawlab_amd.BUILD_GPU)
        img, _ = be.render(spp=2)
        assert img.shape == (16, 16, 3) and np.isfinite(img).all()
    finally:
        be.close()


def test_malformed_scenes_and_parameters(backend):
    d = _one_triangle()
    for mutate, msg in [
        (lambda x: setattr(x, "meshes", []), "scene has no meshes"),
        (lambda x: setattr(x.bsdfs[0], "type", 17), "unknown material type"),
        (lambda x: setattr(x.meshes[0], "bsdf", 5), "mesh material index out of range"),
        (lambda x: setattr(x.meshes[0], "F", np.uint32([[0, 1, 7]])), "vertex index out of range"),
        (lambda x: setattr(x, "camera", dataclasses.replace(x.camera, width=0)), "bad camera size"),
        (lambda x: setattr(x, "camera", dataclasses.replace(x.camera, width=65536, height=65537)),
         "more than 2\\^27 pixels"),
        (lambda x: setattr(x, "rfilter", RFilter(drawlab_amd.FILTER_GAUSSIAN, float("nan"), 0.5)),
         "bad reconstruction filter"),
        (lambda x: setattr(x, "rfilter", RFilter(drawlab_amd.FILTER_GAUSSIAN, 100.0, 0.5)),
         "bad reconstruction filter"),
        (lambda x: setattr(x, "rfilter", RFilter(drawlab_amd.FILTER_GAUSSIAN, -1.0, 0.5)),
         "bad reconstruction filter"),
    ]:
        bad = _one_triangle()
        mutate(bad)
        with pytest.raises(drawlab_amd.DlError, match=msg):
            backend.upload_scene(bad)
    # the context still renders a good scene
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    # a camera whose pixel count would wrap 32-bit pixel indices is refused
    # before any kernel runs
    with pytest.raises(drawlab_amd.DlError, match="more than 2\\^27 pixels"):
        backend.set_camera(dataclasses.replace(d.camera, width=65536, height=65537))
    backend.set_camera(d.camera)
    with pytest.raises(drawlab_amd.DlError, match="spp must be positive"):
        backend.render(spp=0)
    with pytest.raises(drawlab_amd.DlError, match="pass too large"):
        backend.render(spp=2400000, spp_per_pass=2400000)
    img, _ = backend.render(spp=2)
    assert img.mean() > 0
    # no emitter: the reference's CPU renderer message (cpu_renderer.cpp:63-66)
    dark = _one_triangle()
    dark.point_lights = []
    backend.upload_scene(dark)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    with pytest.raises(drawlab_amd.DlError, match="There is not emitter in the scene."):
        backend.render(spp=1)


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_GPU, drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU_LBVH],
                         ids=["gpu", "host_sah", "lbvh"])
def test_degenerate_geometry_and_rays(backend, builder):
    """Zero-area triangles (repeated vertices, collinear points) among
    regular ones, and a one-triangle scene: hits equal the oracle's brute
    force bit for bit, including rays whose interval is empty or inverted;
    renders stay finite and per-path equal to the oracle."""
    rng = np.random.default_rng(3)
    d = _one_triangle()
    V = rng.random((60, 3)).astype(np.float32)
    F = rng.integers(0, 60, (80, 3)).astype(np.uint32)
    F[::7, 1] = F[::7, 0]                               # repeated vertex
    V[10:13] = np.float32([[0.1, 0.1, 0.1], [0.2, 0.2, 0.2], [0.4, 0.4, 0.4]])
    F[3] = (10, 11, 12)                                 # collinear
    d.add_mesh(V, F, 0)
    for desc in (_one_triangle(), d):
        backend.upload_scene(desc)
        backend.build_accel(builder)
        osc = oracle.from_desc(desc)
        n = 4000
        o = (rng.random((n, 3)) * 1.6 - 0.3).astype(np.float32)
        t = rng.random((n, 3)).astype(np.float32)
        dirs = (t - o).astype(np.float32)
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        tmin = np.where(np.arange(n) % 5 == 0, 0.5, 0.0).astype(np.float32)
        tmax = np.where(np.arange(n) % 5 == 0, 0.25, np.inf).astype(np.float32)   # every fifth: tmin > tmax
        tmax[1::5] = 0.0                                                          # empty interval at 0
        rays = planes(o, dirs, tmin, tmax)
        g = backend.trace_closest(rays)
        ref = osc.trace_closest(rays, mode=oracle.ACCEL_BRUTE)
        nbad, idx = compare_hits(g, ref)
        assert nbad == 0, idx[:8]
        prim = np.asarray(g[0])
        assert np.all(prim[::5] == 0xFFFFFFFF) and np.all(prim[1::5] == 0xFFFFFFFF)
        occ = backend.trace_shadow(rays)
        assert np.array_equal(occ, osc.trace_shadow(rays, mode=oracle.ACCEL_BRUTE))
        img, _ = backend.render(spp=4, seed=1)
        assert np.isfinite(img).all()
    paths_parity(backend, d, samples=[0, 1], seed=2, builder=builder)
