"""Scenes loaded from XML by the C++ plugin layer, rendered on the GPU:
per-path parity with the oracle (virtual and perspective cameras), the
HipRenderer path (dlx_render) against the Python binding, and the drawlab
command line writing its OpenEXR and PNG images."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import host
from helpers import paths_parity

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")


def _cbox_copy(tmp_path, camera_xml=None):
    shutil.copy(os.path.join(SCENES, "unit_cube.obj"), tmp_path / "unit_cube.obj")
    text = open(os.path.join(SCENES, "cbox.xml")).read()
    if camera_xml:
        a, b = text.index("<camera"), text.index("</camera>") + len("</camera>")
        text = text[:a] + camera_xml + text[b:]
    p = tmp_path / "cbox.xml"
    p.write_text(text)
    return str(p)


def test_xml_cbox_paths_bitexact(backend):
    xs = host.load_xml(os.path.join(SCENES, "cbox.xml"))
    print("scene", xs.path, flush=True)
    paths_parity(backend, xs.desc, range(3), seed=7)


def test_xml_perspective_paths_bitexact(backend, tmp_path):
    cam = """<camera type="perspective">
      <float name="fov" value="45"/><integer name="width" value="48"/><integer name="height" value="32"/>
      <transform name="toWorld"><lookat origin="0.278, 0.273, -0.8" target="0.278, 0.273, 0" up="0, 1, 0"/></transform>
    </camera>"""
    xs = host.load_xml(_cbox_copy(tmp_path, cam))
    cd = xs.raw.camera
    assert cd.type == 1
    mats = (np.array(cd.sample_to_camera[:], np.float32), np.array(cd.camera_to_world[:], np.float32))
    print("perspective scene", xs.path, flush=True)
    paths_parity(backend, xs.desc, range(3), seed=4, camera_matrices=mats)


def test_hip_renderer_matches_python_binding(backend):
    xs = host.load_xml(os.path.join(SCENES, "cbox.xml"))
    img, st = xs.render(device=0, builder=drawlab_amd.BUILD_GPU, spp=8)
    print("dlx_render", st["paths"], st["ms_total"], flush=True)
    backend.upload_scene(xs.desc)
    backend.build_accel(drawlab_amd.BUILD_GPU)
    ref, _ = backend.render(spp=8, seed=xs.params.seed, max_depth=xs.params.max_depth, rr_depth=xs.params.rr_depth)
    assert st["paths"] == 64 * 64 * 8
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_cli_writes_exr(tmp_path):
    scene = _cbox_copy(tmp_path)
    print("cli", scene, flush=True)
    r = subprocess.run([host.CLI_PATH, "-s", scene, "--spp", "4"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    out = host.load_image(tmp_path / "cbox.exr")
    pfm = host.load_image(tmp_path / "cbox.pfm")
    xs = host.load_xml(scene)
    img, _ = xs.render(spp=4)
    assert out.shape == (64, 64, 3)
    assert np.array_equal(out, img) and np.array_equal(pfm, img)
    assert np.isfinite(out).all() and out.mean() > 0.01
    # and the 8-bit sRGB PNG the reference writes beside the EXR
    host.save_image(tmp_path / "again.png", img)
    assert (tmp_path / "cbox.png").read_bytes() == (tmp_path / "again.png").read_bytes()
