"""Bitmap textures on the HIP path (albedo, anisotropic GGX pd / ps / axay,
tangent- and object-space normal and tangent maps): every camera path of a
scene that uses them all equals the oracle's path, and a textured scene
kept resident in the render service renders like a standalone one."""
import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import scenes
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
