"""Ray batches and comparison helpers for the parity tests."""
import numpy as np

import oracle


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def scene_arrays(desc):
    V = np.concatenate([m.V for m in desc.meshes]).astype(np.float32)
    offs = np.cumsum([0] + [m.V.shape[0] for m in desc.meshes])[:-1]
    F = np.concatenate([m.F + o for m, o in zip(desc.meshes, offs)]).astype(np.int64)
    return V, F


def planes(o, d, tmin, tmax):
    n = o.shape[0]
    r = np.empty((8, n), np.float32)
    r[0:3] = o.T
    r[3:6] = d.T
    r[6] = tmin
    r[7] = tmax
    return r


def unit(v):
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def surface_points(desc, n, rng):
    V, F = scene_arrays(desc)
    f = rng.integers(0, F.shape[0], n)
    s = rng.random((n, 2))
    t = np.sqrt(1 - s[:, 0])
    b = np.stack([1 - t, t * s[:, 1], t - t * s[:, 1]], 1)
    P = (V[F[f, 0]] * b[:, :1] + V[F[f, 1]] * b[:, 1:2] + V[F[f, 2]] * b[:, 2:3]).astype(np.float32)
    return P, f


def ray_batches(desc, osc, n, seed=0):
    """Camera, bounce (surface-to-random), random-in-box, axis-aligned,
    edge/vertex-grazing and shadow rays."""
    rng = np.random.default_rng(seed)
    V, F = scene_arrays(desc)
    lo, hi = V.min(0), V.max(0)
    W, H = desc.camera.width, desc.camera.height
    out = {}
    pos = np.stack([rng.random(n) * W, rng.random(n) * H], 1).astype(np.float32)
    out["camera"] = osc.camera_rays(pos)
    P, _ = surface_points(desc, n, rng)
    d = unit(rng.normal(size=(n, 3)))
    out["bounce"] = planes(P, d, np.float32(1e-4), np.float32(np.inf))
    o = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    out["random"] = planes(o, unit(rng.normal(size=(n, 3))), np.float32(1e-4), np.float32(np.inf))
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    o2 = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    out["axis"] = planes(o2, axes[rng.integers(0, 6, n)], np.float32(1e-4), np.float32(np.inf))
    # aim exactly at vertices and edge midpoints (ties between neighbours)
    f = rng.integers(0, F.shape[0], n)
    k = rng.integers(0, 3, n)
    tgt = V[F[f, k]].astype(np.float32)
    half = rng.random(n) < 0.5
    tgt[half] = ((V[F[f[half], k[half]]] + V[F[f[half], (k[half] + 1) % 3]]) * 0.5).astype(np.float32)
    o3 = (lo + (hi - lo) * rng.random((n, 3)) * 1.5 - (hi - lo) * 0.25).astype(np.float32)
    out["edges"] = planes(o3, unit(tgt - o3), np.float32(1e-4), np.float32(np.inf))
    # shadow-style segments between two surface points
    P1, _ = surface_points(desc, n, rng)
    P2, _ = surface_points(desc, n, rng)
    v = (P2 - P1).astype(np.float32)
    dist = np.linalg.norm(v, axis=1).astype(np.float32)
    ok = dist > 1e-3
    out["shadow"] = planes(P1[ok], unit(v[ok]), np.float32(1e-4), (dist[ok] - np.float32(1e-4)).astype(np.float32))
    return out


def compare_hits(g, o):
    """Returns the number of rays whose (prim, t, u, v) differ bitwise."""
    gp, gt, gu, gv = g
    op, ot, ou, ov = o
    miss = op == 0xFFFFFFFF
    bad = gp != op
    hit = ~miss
    bad |= hit & ((bits(gt) != bits(ot)) | (bits(gu) != bits(ou)) | (bits(gv) != bits(ov)))
    return int(bad.sum()), np.nonzero(bad)[0]


def relmse(x, y):
    """mean over pixels and channels of (x - y)^2 / (y^2 + 1e-2)."""
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))


def paths_parity(backend, desc, samples, seed, builder=0, camera_matrices=None, normals=False, view=None):
    """Per-path radiance of the HIP integrator against the oracle for every
    pixel of the given sample indices: at least 99.9% of paths bit-exact, the
    rest within 1e-3 relative (binary64 libm vs device rounding).
    normals: the normals integrator; view: keyword arguments of
    Backend.set_view_options (bg_color, epsilon, hide_lights), applied to
    both sides."""
    import oracle
    backend.upload_scene(desc)
    backend.build_accel(builder)
    osc = oracle.from_desc(desc, camera_matrices=camera_matrices)
    view = dict(view or {})
    if view:
        backend.set_view_options(**view)
        if view.get("bg_color") is not None:
            osc.set_background(view["bg_color"])
    osc.set_options(normals=normals, hide_lights=view.get("hide_lights", False), eps=view.get("epsilon", 0.0))
    H, W = desc.camera.height, desc.camera.width
    tot, same_all = 0, 0
    for s in samples:
        print("paths sample", s, flush=True)
        g = backend.render_paths(sample=s, seed=seed, normals=normals)
        xys = np.array([(x, y, s) for y in range(H) for x in range(W)], np.int32)
        o = osc.render_samples(xys, seed=seed, max_depth=desc.max_depth, rr_depth=desc.rr_depth)
        same = (g.view(np.uint32) == o.view(np.uint32)).all(axis=1)
        tot += same.size
        same_all += int(same.sum())
        close = np.allclose(g, o, rtol=1e-3, atol=1e-5)
        if not same.all():
            bad = np.nonzero(~same)[0][:5]
            print("differing paths:", bad, g[bad], o[bad], flush=True)
        assert close or same.mean() > 0.999
    frac = same_all / tot
    print("bit-exact paths: %d / %d (%.5f)" % (same_all, tot, frac), flush=True)
    assert frac >= 0.999
    return frac


def tile_means(img, t=16):
    """Per t x t tile and channel means of an (H, W, 3) image."""
    H, W, _ = img.shape
    return img[:H - H % t, :W - W % t].reshape(H // t, t, W // t, t, 3).mean(axis=(1, 3))


def tile_z(batches_a, batches_b, t=16):
    """z = (mean_a - mean_b) / sqrt(var_a / Ka + var_b / Kb) per tile and
    channel, with the variances taken across independent batches; returns
    (z, degrees of freedom of the smaller side)."""
    A = np.stack([tile_means(b, t) for b in batches_a])
    B = np.stack([tile_means(b, t) for b in batches_b])
    return tile_z_means(A, B)


def tile_z_means(A, B):
    """tile_z from per-batch tile means (K, tiles_y, tiles_x, 3)."""
    A, B = np.asarray(A, np.float64), np.asarray(B, np.float64)
    va = A.var(axis=0, ddof=1) / len(A)
    vb = B.var(axis=0, ddof=1) / len(B)
    return (A.mean(axis=0) - B.mean(axis=0)) / np.sqrt(va + vb + 1e-30), min(len(A), len(B)) - 1


def assert_unbiased(zd, alpha=0.01):
    """z from batch means follows a Student t with the batches' degrees of
    freedom under the null.  Bonferroni over all tiles and channels for the
    largest |z|, and a binomial tail test for the count beyond 3."""
    from scipy import stats
    z, dof = zd
    m = z.size
    zmax = stats.t.isf(alpha / (2 * m), dof)
    p3 = 2 * stats.t.sf(3.0, dof)
    k3 = int(np.sum(np.abs(z) > 3))
    tail = stats.binom.sf(k3 - 1, m, p3)
    print("tiles %d  dof %d  max |z| %.2f (limit %.2f)  beyond 3: %d (P = %.3g)" % (m, dof, np.abs(z).max(), zmax, k3,
                                                                                    tail), flush=True)
    assert np.abs(z).max() < zmax
    assert tail > alpha
