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
    L.bvh_check.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int]
    return L


def _run(L, desc, sah_depth=40):
    V, F = scene_arrays(desc)
    tv = np.ascontiguousarray(V[F].reshape(-1, 9), dtype=np.float32)
    st = np.zeros(3, np.uint32)
    rc = L.bvh_check(tv.ctypes.data_as(ctypes.c_void_p), tv.shape[0], st.ctypes.data_as(ctypes.c_void_p),
                     int(sah_depth))
    assert rc == 0, "bvh structure check failed with code %d" % rc
    return st


@pytest.mark.parametrize("make", [lambda: scenes.cornell_box(8, 8), lambda: scenes.mis_plates(8, 8),
                                  lambda: scenes.displaced_sphere(8, 8, grid=128, glass_level=3, env_size=(8, 4)),
                                  lambda: scenes.random_scene(3, 20000)])
def test_bvh_structure(checker, make):
    st = _run(checker, make())
    assert st[2] <= 40


def _soup(V, F):
    from drawlab_amd.scene import SceneDesc, Bsdf
    d = SceneDesc()
    d.add_mesh(np.asarray(V, np.float32), np.asarray(F, np.uint32), d.add_bsdf(Bsdf.diffuse()))
    return d


@pytest.mark.parametrize("sah_depth", [2, 6, 12])
def test_bvh_median_levels_below_the_sah_cap(checker, sah_depth):
    """Below the SAH depth cap the builder splits at the object median until
    every leaf fits the 4-wide node's 8-bit triangle count: with the cap
    lowered, a 50k-triangle scene still gives leaves of at most 255
    triangles, every triangle once, within the traversal stack's depth."""
    d = scenes.random_scene(5, 50000)
    st = _run(checker, d, sah_depth=sah_depth)
    print("sah depth %d, %d triangles: nodes/leaves/depth" % (sah_depth, d.num_triangles), st)
    assert st[1] >= -(-d.num_triangles // 255) and st[2] <= 64


def test_bvh_coincident_triangles(checker):
    """5000 copies of one triangle (no split separates them): leaves of at
    most 255 triangles."""
    n = 5000
    V = np.tile(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]), (n, 1))
    st = _run(checker, _soup(V, np.arange(3 * n).reshape(n, 3)))
    assert st[1] >= n // 255


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
