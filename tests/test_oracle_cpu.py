"""Pins of the CPU oracle: published known answers, the reference's own
compiled math (oracle/_ref), closed forms and invariants."""
import math

import numpy as np
import pytest

import oracle
from drawlab_amd import scenes
from helpers import bits, ray_batches, compare_hits


def test_pcg32_published_known_answers():
    # pcg32-demo (pcg-random.org), initstate 42, initseq 54
    got = oracle.pcg32(42, 54, 6)
    assert [hex(x) for x in got] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b",
                                     "0xcbed606e"]


def _pcg_py(initstate, initseq, n):
    M = (1 << 64) - 1

    def step(st, inc):
        old = st
        st = (old * 0x5851F42D4C957F2D + inc) & M
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return st, ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    inc = ((initseq << 1) | 1) & M
    st, _ = step(0, inc)
    st = (st + initstate) & M
    st, _ = step(st, inc)
    out = []
    for _ in range(n):
        st, o = step(st, inc)
        out.append(o)
    return out


@pytest.mark.parametrize("state,seq", [(0, 0), (12345, 678), ((7 << 32) + 99, 1023), (2**63 + 5, 2**40)])
def test_pcg32_matches_bigint_restatement(state, seq):
    assert list(oracle.pcg32(state, seq, 64)) == _pcg_py(state, seq, 64)
    f = oracle.pcg32_float(state, seq, 64)
    exp = np.array([(x >> 9) * 2.0 ** -23 for x in _pcg_py(state, seq, 64)], np.float32)
    assert np.array_equal(f, exp)


REF = oracle.ref_lib()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (needs the reference sources)")


@needs_ref
@pytest.mark.parametrize("which,alpha", [(0, 0), (1, 0), (2, 0.05), (2, 0.3), (2, 1.0), (3, 0)])
def test_warps_match_reference_build(which, alpha):
    rng = np.random.default_rng(which)
    inp = rng.random((50000, 3), dtype=np.float32)
    inp[:8, :2] = [[0, 0], [0.5, 0.5], [1 - 2**-23, 0.25], [0.75, 1 - 2**-23], [2**-20, 0.1], [0.3, 0], [0.9, 0.9],
                   [0.001, 0.999]]
    o = oracle.warp(which, inp, alpha)
    r = np.zeros(inp.size, np.float32)
    REF.ref_warp(which, np.ascontiguousarray(inp.reshape(-1)), inp.shape[0], alpha, 0.0, r)
    assert np.array_equal(bits(o), bits(r.reshape(-1, 3)))


@needs_ref
@pytest.mark.parametrize("which,alpha", [(4, 0.1), (4, 0.7), (5, 0)])
def test_warp_pdfs_match_reference_build(which, alpha):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(50000, 3)).astype(np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    o = oracle.warp(which, v, alpha)
    r = np.zeros(v.size, np.float32)
    REF.ref_warp(which, np.ascontiguousarray(v.reshape(-1)), v.shape[0], alpha, 0.0, r)
    assert np.array_equal(bits(o), bits(r.reshape(-1, 3)))


@needs_ref
def test_optics_match_reference_build():
    rng = np.random.default_rng(2)
    n = 40000
    c = rng.uniform(-1, 1, n).astype(np.float32)
    ext = rng.uniform(1.0, 1.6, n).astype(np.float32)
    it = rng.uniform(1.0, 2.4, n).astype(np.float32)
    inp = np.zeros((n, 6), np.float32)
    inp[:, 0], inp[:, 1], inp[:, 2] = c, ext, it
    r = np.zeros(3 * n, np.float32)
    REF.ref_optics(0, np.ascontiguousarray(inp.reshape(-1)), n, r)
    assert np.array_equal(bits(oracle.optics(0, inp)[:, 0]), bits(r.reshape(n, 3)[:, 0]))
    wi = rng.normal(size=(n, 3)).astype(np.float32)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    inp[:, :3] = wi
    inp[:, 3] = rng.uniform(0.5, 1.0, n)
    REF.ref_optics(1, np.ascontiguousarray(inp.reshape(-1)), n, r)
    assert np.array_equal(bits(oracle.optics(1, inp)), bits(r.reshape(n, 3)))
    inp[:, 3:] = rng.normal(size=(n, 3))
    REF.ref_optics(2, np.ascontiguousarray(inp.reshape(-1)), n, r)
    assert np.array_equal(bits(oracle.optics(2, inp)), bits(r.reshape(n, 3)))


def test_fresnel_closed_forms():
    # normal incidence: ((n1 - n2) / (n1 + n2))^2; equal indices: 0; TIR: 1
    inp = np.array([[1.0, 1.0, 1.5, 0, 0, 0], [1.0, 1.5, 1.5, 0, 0, 0], [-0.1, 1.0, 1.5, 0, 0, 0]], np.float32)
    f = oracle.optics(0, inp)[:, 0]
    assert abs(f[0] - 0.04) < 1e-7
    assert f[1] == 0.0
    assert f[2] == 1.0
    # conductor at normal incidence: ((eta-1)^2 + k^2) / ((eta+1)^2 + k^2)
    eta, k = 0.2, 3.912
    fc = oracle.optics(3, np.array([[1.0, eta, k, 0, 0, 0]], np.float32))[0, 0]
    assert abs(fc - ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k)) < 2e-6


def test_vector_known_answers_from_reference_tests():
    # tests/core/math/test_vector.cpp:109-115: (1,2,3) x (10,-2,20) = (46,10,-22)
    if REF is None:
        pytest.skip("oracle/_ref not built")
    inp = np.array([[1, 2, 3, 10, -2, 20]], np.float32)
    out = np.zeros(3, np.float32)
    REF.ref_vec(1, inp.reshape(-1), 1, out)
    assert list(out) == [46, 10, -22]
    REF.ref_vec(0, inp.reshape(-1), 1, out)
    assert out[0] == 1 * 10 + 2 * -2 + 3 * 20


def test_gaussian_filter_table_closed_form():
    s = oracle.Scene()
    s.set_filter(oracle.FILTER_GAUSSIAN, 2.0, 0.5)
    t = s.filter_table()
    x = np.arange(32, dtype=np.float64) * 2.0 / 32
    ref = np.maximum(0, np.exp(-2 * x * x) - np.exp(-8.0))
    assert np.allclose(t[:32], ref, atol=1e-6) and t[32] == 0
    s.set_filter(oracle.FILTER_BOX)
    assert np.all(s.filter_table()[:32] == 1.0)


def test_box_filter_image_is_per_pixel_mean():
    d = scenes.cornell_box(16, 16, spp=4)
    s = oracle.from_desc(d)
    s.set_filter(oracle.FILTER_BOX)
    img, _ = s.render(4, seed=3, max_depth=8, rr_depth=3)
    xys = np.array([(x, y, k) for y in range(16) for x in range(16) for k in range(4)], np.int32)
    p = s.render_samples(xys, seed=3, max_depth=8, rr_depth=3)
    mean = p[:, 2:].reshape(16, 16, 4, 3).mean(2)
    assert np.allclose(img, mean, rtol=1e-5, atol=1e-6)


def test_envmap_alias_invariant():
    env = scenes.sky_envmap(64, 32)
    s = oracle.Scene()
    s.add_bsdf(oracle.BSDF_DIFFUSE, [0.5, 0.5, 0.5])
    s.add_envmap(env)
    q, alias, pdf = s.env_tables(64, 32)
    n = q.size
    # probability of picking pixel i through the alias table equals its importance
    prob = q / n
    np.add.at(prob, alias, (1 - q) / n)
    theta = (np.arange(33, dtype=np.float64)) * math.pi / 32
    area = (np.cos(theta[:-1]) - np.cos(theta[1:]))[:, None] * (2 * math.pi / 64)
    imp = (area * env.max(axis=2)).reshape(-1)
    # exact up to the float32 accumulation residue of the in-place construction,
    # which leaves the last leftover entry's q slightly below 1
    r = np.abs(prob - imp / imp.sum()) / (imp / imp.sum())
    assert np.sort(r)[-2] < 1e-5 and r.max() < 1e-3
    # solid-angle pdf integrates to one
    assert abs((pdf.reshape(32, 64) * area).sum() - 1.0) < 1e-4


def test_spiral_block_order():
    b = oracle.spiral_blocks(96, 64)  # 3 x 2 blocks, starts at the centre block (1, 1)
    assert tuple(b[0]) == (32, 32)
    assert len(b) == 6 and len({tuple(x) for x in b}) == 6


@pytest.mark.parametrize("make", [lambda: scenes.cornell_box(32, 32), lambda: scenes.mis_plates(32, 18),
                                  lambda: scenes.random_scene(4, 5000, 32, 32)])
def test_robust_octree_matches_brute_force(make):
    d = make()
    s = oracle.from_desc(d)
    for name, rays in ray_batches(d, s, 3000, seed=8).items():
        if name == "shadow":
            assert np.array_equal(s.trace_shadow(rays, oracle.ACCEL_OCTREE_ROBUST),
                                  s.trace_shadow(rays, oracle.ACCEL_BRUTE))
            continue
        a = s.trace_closest(rays, oracle.ACCEL_OCTREE_ROBUST)
        b = s.trace_closest(rays, oracle.ACCEL_BRUTE)
        nbad, _ = compare_hits(a, b)
        assert nbad == 0, name


def test_reference_octree_divergence_is_confined_to_edge_rays():
    """The reference octree (lowest-index tie rule applied) equals brute
    force on generic rays; its unpadded slabs only lose hits of rays aimed
    exactly at triangle edges and vertices (reported, not hidden)."""
    d = scenes.cornell_box(32, 32)
    s = oracle.from_desc(d)
    for name, rays in ray_batches(d, s, 4000, seed=8).items():
        if name == "shadow":
            continue
        a = s.trace_closest(rays, oracle.ACCEL_OCTREE_LOWEST)
        b = s.trace_closest(rays, oracle.ACCEL_BRUTE)
        nbad, _ = compare_hits(a, b)
        if name == "edges":
            print("reference octree diverges on %d of %d edge-aimed rays" % (nbad, rays.shape[1]))
            assert nbad < 0.05 * rays.shape[1]
        else:
            assert nbad == 0, name


def test_reference_octree_never_finds_a_closer_hit():
    """The reference octree keeps the last equal-t hit and can lose edge
    hits: where it disagrees with brute force its hit is never closer."""
    d = scenes.cornell_box(32, 32)
    s = oracle.from_desc(d)
    rays = np.ascontiguousarray(ray_batches(d, s, 20000, seed=9)["edges"])
    a = s.trace_closest(rays, oracle.ACCEL_OCTREE)
    b = s.trace_closest(rays, oracle.ACCEL_BRUTE)
    diff = a[0] != b[0]
    assert np.all(a[1][diff] >= b[1][diff])
    ties = diff & (a[1] == b[1])
    print("edge rays: %d differ, %d of them equal-t ties" % (int(diff.sum()), int(ties.sum())))


def test_mt_hand_cases():
    # axis-aligned unit square at z = 0 hit head-on at dyadic points
    from drawlab_amd.scene import SceneDesc, Bsdf
    d = SceneDesc()
    b = d.add_bsdf(Bsdf.diffuse())
    V, F, _ = scenes.quad((0, 0, 0), (1, 0, 0), (0, 1, 0))
    d.add_mesh(V, F, b)
    d.camera = scenes.Camera(width=4, height=4)
    s = oracle.from_desc(d)
    o = np.array([[0.25, 0.25, 1.0], [0.75, 0.5, 2.0], [0.5, 0.5, 1.0], [1.5, 0.5, 1.0]], np.float32)
    dirs = np.tile(np.array([[0, 0, -1]], np.float32), (4, 1))
    from helpers import planes
    prim, t, u, v = s.trace_closest(planes(o, dirs, np.float32(1e-4), np.float32(np.inf)))
    assert list(prim[:3]) == [0, 1, 0] and prim[3] == 0xFFFFFFFF
    assert list(t[:3]) == [1.0, 2.0, 1.0]
    # (0.25, 0.25) in triangle 0 (p0=(0,0), e1=(1,0), e2=(0,1)): u = 0.25, v = 0.25
    assert (u[0], v[0]) == (0.25, 0.25)
    # the diagonal point (0.5, 0.5) lies on both triangles: lowest index wins
    assert prim[2] == 0 and (u[2], v[2]) == (0.5, 0.5)
