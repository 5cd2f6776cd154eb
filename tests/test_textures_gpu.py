"""Bitmap textures on the HIP path (albedo, anisotropic GGX pd / ps / axay,
tangent- and object-space normal and tangent maps): every camera path of a
scene that uses them all equals the oracle's path, also for a scene whose
textures the plugin layer loaded from PNG / HDR files, and the HipRenderer
path renders that scene like the Python binding."""
import os

import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import host, scenes
from helpers import paths_parity

pytestmark = pytest.mark.gpu


def test_textured_box_paths_bitexact(backend):
    paths_parity(backend, scenes.textured_box(64, 64), samples=[0, 1, 5], seed=11)


def test_textured_box_image_matches_oracle(backend):
    import oracle
    d = scenes.textured_box(48, 48)
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    img, _ = backend.render(spp=16, seed=4, max_depth=6, rr_depth=3)
    ref, _ = oracle.from_desc(d).render(16, seed=4, max_depth=6, rr_depth=3)
    assert np.isfinite(img).all()
    err = float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 1e-2)))
    print("relMSE vs oracle", err, flush=True)
    assert err <= 1e-4


def test_textures_invalid_index_rejected(backend):
    d = scenes.textured_box(16, 16)
    d.bsdfs[2].textures = {0: 99}
    with pytest.raises(drawlab_amd.DlError, match="texture index"):
        backend.upload_scene(d)


def test_xml_textured_scene_paths_bitexact(backend):
    """Textures loaded by the plugin layer (PNG rows, palette, Radiance HDR,
    tangent- and object-space normal maps) render per path like the
    oracle over the same exported scene."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes", "textured.xml")
    xs = host.load_xml(path)
    print("scene", xs.path, len(xs.desc.textures), "textures", flush=True)
    paths_parity(backend, xs.desc, samples=[0, 1, 2], seed=5)


def test_xml_textured_hip_renderer_matches_python_binding(backend):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes", "textured.xml")
    xs = host.load_xml(path)
    img, st = xs.render(device=0, builder=drawlab_amd.BUILD_GPU, spp=8)
    print("dlx_render", st["paths"], flush=True)
    backend.upload_scene(xs.desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    ref, _ = backend.render(spp=8, seed=xs.params.seed, max_depth=xs.params.max_depth, rr_depth=xs.params.rr_depth)
    assert st["paths"] == 48 * 36 * 8
    assert np.isfinite(img).all() and img.mean() > 0
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
