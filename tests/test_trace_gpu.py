"""Bit-exact closest-hit and shadow parity of the HIP traversal kernels
against the oracle's brute-force Moller-Trumbore argmin (lowest primitive id
on equal t)."""
import numpy as np
import pytest

import drawlab_amd
import oracle
from drawlab_amd import scenes

from helpers import compare_hits, ray_batches

pytestmark = pytest.mark.gpu


def _check(be, desc, builder, n, seed, brute_n=None):
    be.upload_scene(desc)
    st = be.build_accel(builder)
    print("accel", st, flush=True)
    osc = oracle.from_desc(desc)
    for name, rays in ray_batches(desc, osc, n, seed).items():
        rays = np.ascontiguousarray(rays)
        if brute_n is not None:
            rays = np.ascontiguousarray(rays[:, :brute_n])
        print("batch", name, rays.shape[1], flush=True)
        if name == "shadow":
            g = be.trace_shadow(rays)
            o = osc.trace_shadow(rays, oracle.ACCEL_BRUTE)
            assert int((g != o).sum()) == 0, name
            continue
        g = be.trace_closest(rays)
        o = osc.trace_closest(rays, oracle.ACCEL_BRUTE)
        nbad, idx = compare_hits(g, o)
        assert nbad == 0, "%s: %d of %d rays differ (first %s)" % (name, nbad, rays.shape[1], idx[:8])
        assert (o[0] != 0xFFFFFFFF).mean() > 0.05 or name in ("random", "axis")


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU], ids=["host_sah", "gpu"])
def test_cornell_box_hits_bitexact(backend, builder):
    _check(backend, scenes.cornell_box(64, 64), builder, 20000, seed=1)


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU], ids=["host_sah", "gpu"])
def test_mis_plates_hits_bitexact(backend, builder):
    _check(backend, scenes.mis_plates(64, 36), builder, 20000, seed=2)


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU], ids=["host_sah", "gpu"])
def test_displaced_sphere_hits_bitexact(backend, builder):
    desc = scenes.displaced_sphere(width=64, height=36, grid=256, glass_level=4, env_size=(64, 32))
    _check(backend, desc, builder, 4000, seed=3)


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU], ids=["host_sah", "gpu"])
def test_displaced_sphere_full_size_hits(backend, builder):
    """2,097,152-triangle sphere: brute force on a bounded batch, plus the
    robust octree on a larger one."""
    desc = scenes.displaced_sphere(width=64, height=36, env_size=(64, 32))
    _check(backend, desc, builder, 600, seed=4)
    osc = oracle.from_desc(desc)
    for name, rays in ray_batches(desc, osc, 50000, seed=5).items():
        rays = np.ascontiguousarray(rays)
        if name == "shadow":
            assert int((backend.trace_shadow(rays) != osc.trace_shadow(rays, oracle.ACCEL_OCTREE_ROBUST)).sum()) == 0
            continue
        nbad, idx = compare_hits(backend.trace_closest(rays), osc.trace_closest(rays, oracle.ACCEL_OCTREE_ROBUST))
        assert nbad == 0, "%s: %d differ" % (name, nbad)


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_GPU, drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU_LBVH],
                         ids=["gpu", "host_sah", "lbvh"])
def test_traversal_counts_match_cpu_replay(backend, builder):
    """The node visits and triangle tests of every ray, closest-hit and
    shadow -- the counts behind the rooflines' algorithmic bytes -- and the
    hits themselves equal a CPU
    replay of the traversal over the exported 4-wide BVH (oracle/replay.cpp),
    including rays with NaN components, which finish with no visit."""
    desc = scenes.displaced_sphere(width=64, height=36, grid=128, glass_level=3, env_size=(64, 32))
    backend.upload_scene(desc)
    st = backend.build_accel(builder)
    nodes, tris = backend.accel_export()
    assert nodes.size == st["num_nodes"] * 16 and tris.size == st["num_prims"] * 12
    osc = oracle.from_desc(desc)
    batches = ray_batches(desc, osc, 3000, seed=7)
    nan_rays = np.array(batches["random"][:, :64])
    nan_rays[np.arange(64) % 8, np.arange(64)] = np.nan
    batches["nan"] = nan_rays
    total = 0
    for name, rays in batches.items():
        rays = np.ascontiguousarray(rays, np.float32)
        if name == "shadow":
            r = oracle.bvh4_replay(nodes, tris, rays, any_hit=True)
            g = backend.trace_shadow(rays, counted=True)
            assert np.array_equal(backend.trace_shadow(rays), r[0]), name
            for a, b, what in zip(g, r, ("occlusion", "node visits", "triangle tests")):
                assert np.array_equal(a, b), "shadow %s differ" % what
            continue
        g = backend.trace_closest(rays, counted=True)
        r = oracle.bvh4_replay(nodes, tris, rays)
        for a, b, what in zip(g, r, ("prim", "t", "u", "v", "node visits", "triangle tests")):
            a, b = np.asarray(a), np.asarray(b)
            bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
            assert bad.size == 0, "%s %s: %d rays differ (first %s)" % (name, what, bad.size, bad[:8])
        total += int(r[4].sum())
        if name == "nan":
            assert not r[4].any()
    print("replayed node visits", total, flush=True)
    assert total > 0


@pytest.mark.parametrize("builder", [drawlab_amd.BUILD_GPU, drawlab_amd.BUILD_HOST_SAH, drawlab_amd.BUILD_GPU_LBVH],
                         ids=["gpu", "host_sah", "lbvh"])
def test_quantised_boxes_contain_their_triangles(backend, builder):
    """Every decoded child box of the exported 4-wide BVH (origin + q * step
    in binary32, as the traversal decodes it) contains the vertices of every
    triangle below it; the boxes are not degenerate padding of the whole
    scene (the mean leaf box is far smaller than the scene)."""
    desc = scenes.displaced_sphere(width=32, height=18, grid=64, glass_level=2, env_size=(32, 16))
    backend.upload_scene(desc)
    st = backend.build_accel(builder)
    nodes, tris = backend.accel_export()
    N = st["num_nodes"]
    w = nodes.view(np.uint32).reshape(N, 16)
    origin = nodes.reshape(N, 16)[:, :3]
    step = ((w[:, 3:4] >> np.array([0, 8, 16], np.uint32)) & 0xFF) << 23
    step = step.astype(np.uint32).view(np.float32)                          # (N, 3)
    q = ((w[:, 4:10, None] >> (8 * np.arange(4, dtype=np.uint32))) & 0xFF).astype(np.float32)  # (N, 6, 4)
    planes = np.empty((N, 6, 4), np.float32)
    for k in range(6):
        ax = k // 2
        planes[:, k] = (origin[:, ax:ax + 1] + q[:, k] * step[:, ax:ax + 1]).astype(np.float32)
    cnt = ((w[:, 10, None] >> (8 * np.arange(4, dtype=np.uint32))) & 0xFF).astype(np.int64)
    ref = nodes.view(np.int32).reshape(N, 16)[:, 12:16].astype(np.int64)
    t = tris.reshape(-1, 12)
    p0 = t[:, 0:3]
    e1 = np.stack([t[:, 3], t[:, 4], t[:, 5]], 1)
    e2 = np.stack([t[:, 6], t[:, 7], t[:, 8]], 1)
    verts = np.stack([p0, (p0 + e1).astype(np.float32), (p0 + e2).astype(np.float32)], 1)   # (T, 3, 3)
    tlo, thi = verts.min(axis=1), verts.max(axis=1)
    # subtree bounds, children after parents (breadth-first order): fold from the end
    lo = np.full((N, 3), np.inf, np.float32)
    hi = np.full((N, 3), -np.inf, np.float32)
    slot_lo = np.full((N, 4, 3), np.inf, np.float32)
    slot_hi = np.full((N, 4, 3), -np.inf, np.float32)
    for i in range(N - 1, -1, -1):
        for c in range(4):
            r = ref[i, c]
            if r >= 0:
                assert r > i
                slot_lo[i, c], slot_hi[i, c] = lo[r], hi[r]
            elif cnt[i, c] > 0:
                f = ~r
                slot_lo[i, c] = tlo[f:f + cnt[i, c]].min(axis=0)
                slot_hi[i, c] = thi[f:f + cnt[i, c]].max(axis=0)
        lo[i] = slot_lo[i].min(axis=0)
        hi[i] = slot_hi[i].max(axis=0)
    used = (ref >= 0) | (cnt > 0)
    box_lo = np.stack([planes[:, 0], planes[:, 2], planes[:, 4]], -1)        # (N, 4, 3)
    box_hi = np.stack([planes[:, 1], planes[:, 3], planes[:, 5]], -1)
    assert np.all(box_lo[used] <= slot_lo[used]) and np.all(box_hi[used] >= slot_hi[used])
    leaf = (ref < 0) & (cnt > 0)
    ext = (box_hi - box_lo)[leaf].max(axis=-1)
    scene = float((hi[0] - lo[0]).max())
    print("leaf box extent mean / scene", float(ext.mean()) / scene, flush=True)
    assert ext.mean() < 0.05 * scene


def _exponential_chain(n=100):
    """n triangles at distances and sizes doubling along x: agglomerative
    clustering merges them one at a time into a chain n - 1 levels deep."""
    from drawlab_amd.scene import Bsdf, SceneDesc
    s = 2.0 ** np.arange(n)
    base = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.05]])
    V = (base[None] * s[:, None, None] + np.stack([s, 0.3 * s, 0.0 * s], 1)[:, None, :]).reshape(-1, 3)
    d = SceneDesc()
    d.add_mesh(V.astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(n, 3), d.add_bsdf(Bsdf.diffuse()))
    return d, V.reshape(n, 3, 3)


def test_too_deep_gpu_tree_is_rebuilt_on_the_host(backend):
    """A GPU-built tree deeper than the traversal stack allows is rebuilt by
    the host builder within dl_hip_build_accel: the result is the host
    builder's BVH, and hits stay bit-exact."""
    d, tris = _exponential_chain()
    backend.upload_scene(d)
    backend.build_accel(drawlab_amd.BUILD_HOST_SAH)
    host_nodes, host_tris = backend.accel_export()
    st = backend.build_accel(drawlab_amd.BUILD_GPU)
    print("accel", st, flush=True)
    gpu_nodes, gpu_tris = backend.accel_export()
    assert np.array_equal(gpu_nodes.view(np.uint32), host_nodes.view(np.uint32))
    assert np.array_equal(gpu_tris.view(np.uint32), host_tris.view(np.uint32))
    # rays straight down onto each triangle's centroid and its vertices
    c = tris.mean(1)
    scale = np.abs(tris).max(axis=(1, 2))
    targets = np.concatenate([c, tris[:, 0], tris[:, 1], tris[:, 2]]).astype(np.float32)
    up = np.tile(scale, 4)[:, None] * np.array([[0.0, 0.0, 1.0]])
    o = (targets + up).astype(np.float32)
    rays = np.zeros((8, len(o)), np.float32)
    rays[0:3] = o.T
    rays[3:6] = np.array([0.0, 0.0, -1.0], np.float32)[:, None]
    rays[6], rays[7] = 0.0, np.inf
    osc = oracle.from_desc(d)
    g = backend.trace_closest(rays)
    ref = osc.trace_closest(rays, oracle.ACCEL_BRUTE)
    nbad, idx = compare_hits(g, ref)
    assert nbad == 0, idx[:8]
    assert (ref[0] != 0xFFFFFFFF).mean() > 0.2
