"""CPU checks of the BVH builder's output structure (run before any GPU use)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from drawlab_amd import scenes
from helpers import scene_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("bvh") / "bvh_check.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", out,
                    os.path.join(HERE, "native", "bvh_check.cpp"),
                    os.path.join(ROOT, "drawlab_amd", "csrc", "host", "bvh_build.cpp"), "-lpthread"], check=True)
    L = ctypes.CDLL(out)
    L.bvh_check.restype = ctypes.c_int
    L.bvh_check.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    return L


def _run(L, desc):
    V, F = scene_arrays(desc)
    tv = np.ascontiguousarray(V[F].reshape(-1, 9), dtype=np.float32)
    st = np.zeros(3, np.uint32)
    rc = L.bvh_check(tv.ctypes.data_as(ctypes.c_void_p), tv.shape[0], st.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0, "bvh structure check failed with code %d" % rc
    return st


@pytest.mark.parametrize("make", [lambda: scenes.cornell_box(8, 8), lambda: scenes.mis_plates(8, 8),
                                  lambda: scenes.displaced_sphere(8, 8, grid=128, glass_level=3, env_size=(8, 4)),
                                  lambda: scenes.random_scene(3, 20000)])
def test_bvh_structure(checker, make):
    st = _run(checker, make())
    assert st[2] <= 48


def test_bvh_structure_tiny(checker):
    s = scenes.SceneDesc() if hasattr(scenes, "SceneDesc") else None
    from drawlab_amd.scene import SceneDesc, Bsdf
    d = SceneDesc()
    b = d.add_bsdf(Bsdf.diffuse())
    V, F, _ = scenes.quad((0, 0, 0), (1, 0, 0), (0, 1, 0))
    d.add_mesh(V, F, b)
    _run(checker, d)


@pytest.mark.slow
def test_bvh_structure_2m(checker):
    st = _run(checker, scenes.displaced_sphere(8, 8, env_size=(8, 4)))
    print("2M-tri BVH nodes/leaves/depth", st)
