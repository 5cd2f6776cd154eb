"""Seeded procedural scenes for the benchmark workloads.

The reference ships no scenes, so every workload is generated in code:
  cornell_box       Cornell box, 5 walls + 2 axis-aligned boxes + ceiling light
                    (36 triangles), 512x512, 64 spp, max depth 8, RR from 3
  displaced_sphere  2,097,152-triangle octahedral sphere displaced by 4-octave
                    value noise (GGX conductor), 20,480-triangle glass
                    icosphere, diffuse ground, 1024x512 sky + sun env map,
                    1920x1080, 256 spp, max depth 16
  mis_plates        4 GGX plates (last anisotropic), 4 icosphere lights of equal
                    power, dark backdrop, 1280x720, 512 spp, max depth 6
  random_scene      displaced spheres and boxes with random materials and
                    1-8 area lights (render-service job mix)
"""
import math

import numpy as np

from .scene import (BSDF_ANISO_GGX, BSDF_DIFFUSE, FILTER_GAUSSIAN, TEX_ALBEDO, TEX_AXAY, TEX_NORMAL, TEX_PD, TEX_TANGENT,
                    Bsdf, Camera, EnvMap, RFilter, SceneDesc)


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
def quad(p, e0, e1, with_uv=False):
    """Two triangles (p, p+e0, p+e1), (p+e0+e1, p+e1, p+e0); normal = e0 x e1
    (the reference's `rectangle` mesh, src/tracer/mesh/rectangle.cpp:11-40)."""
    p, e0, e1 = (np.asarray(a, np.float64) for a in (p, e0, e1))
    V = np.array([p, p + e0, p + e1, p + e0 + e1], np.float32)
    F = np.array([[0, 1, 2], [3, 2, 1]], np.uint32)
    UV = np.array([[1, 0], [1, 1], [0, 0], [0, 1]], np.float32) if with_uv else None
    return V, F, UV


def merge(parts):
    Vs, Fs, off = [], [], 0
    for V, F in parts:
        Vs.append(V)
        Fs.append(F + off)
        off += V.shape[0]
    return np.concatenate(Vs).astype(np.float32), np.concatenate(Fs).astype(np.uint32)


def box(lo, hi):
    """Axis-aligned box as 6 outward-facing quads (12 triangles)."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    d = hi - lo
    X, Y, Z = np.array([d[0], 0, 0]), np.array([0, d[1], 0]), np.array([0, 0, d[2]])
    faces = [
        (lo, X, Z),            # bottom, -y
        (lo + Y, Z, X),        # top, +y
        (lo, Y, X),            # front (z = lo), -z
        (lo + Z, X, Y),        # back (z = hi), +z
        (lo, Z, Y),            # x = lo, -x
        (lo + X, Y, Z),        # x = hi, +x
    ]
    return merge([quad(p, a, b)[:2] for p, a, b in faces])


def icosphere(level, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron: 20 * 4**level triangles, outward winding."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, np.float64)
    F = np.array(faces, np.uint32)
    return _orient_outward(V.astype(np.float32), F, center)


def _orient_outward(V, F, center):
    p0, p1, p2 = V[F[:, 0]].astype(np.float64), V[F[:, 1]].astype(np.float64), V[F[:, 2]].astype(np.float64)
    n = np.cross(p1 - p0, p2 - p0)
    c = (p0 + p1 + p2) / 3.0 - np.asarray(center, np.float64)
    flip = (n * c).sum(1) < 0
    F = F.copy()
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def value_noise(p, seed, octaves=4, base_freq=2.0):
    """Seeded 3-D value noise (smoothstep-interpolated lattice values in
    [-1, 1]) summed over octaves with halving amplitude, normalised to [-1, 1]."""
    p = np.asarray(p, np.float64)
    total = np.zeros(p.shape[0])
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        f = base_freq * (2 ** o)
        q = p * f
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3.0 - 2.0 * fr)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    h = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed + 7919 * o)
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += wx * wy * wz * h
        total += amp * acc
        norm += amp
        amp *= 0.5
    return total / norm


def _hash3(x, y, z, seed):
    h = (x * 73856093) ^ (y * 19349663) ^ (z * 83492791) ^ (seed * 2654435761)
    h = h & 0xFFFFFFFF
    h = (h ^ (h >> 16)) * 0x7FEB352D & 0xFFFFFFFF
    h = (h ^ (h >> 15)) * 0x846CA68B & 0xFFFFFFFF
    h = h ^ (h >> 16)
    return (h.astype(np.float64) / 4294967295.0) * 2.0 - 1.0


def octa_sphere(n, radius=1.0, center=(0.0, 0.0, 0.0), displace=0.0, seed=0):
    """Octahedral-parametrised sphere on an n x n grid: 2 n^2 triangles
    (n = 1024 gives 2,097,152).  Quads are split along the diagonal parallel
    to the octahedron's edges so no triangle straddles one.  Optional radial
    displacement radius * (1 + displace * noise)."""
    g = np.linspace(-1.0, 1.0, n + 1)
    u, v = np.meshgrid(g, g, indexing="ij")
    u, v = u.reshape(-1), v.reshape(-1)
    z = 1.0 - np.abs(u) - np.abs(v)
    x, y = u.copy(), v.copy()
    neg = z < 0
    x[neg] = (1.0 - np.abs(v[neg])) * np.sign(u[neg])
    y[neg] = (1.0 - np.abs(u[neg])) * np.sign(v[neg])
    # sign(0) = 0 would collapse points on the axes; those lie on |u|+|v|=... only when z<0 with u or v = 0
    x[neg & (u == 0)] = 0.0
    y[neg & (v == 0)] = 0.0
    P = np.stack([x, y, z], 1)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    # boundary points of the square map to the same sphere point: snap them
    # through the same formula so duplicates are bit-identical
    r = np.full(P.shape[0], radius, np.float64)
    if displace:
        r = radius * (1.0 + displace * value_noise(P, seed))
    V = (P * r[:, None] + np.asarray(center, np.float64)).astype(np.float32)
    # canonicalise duplicates (boundary of the octahedral map)
    key = np.round(P * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    V = V[first[inv.reshape(-1)]]
    i = np.arange(n)
    I, J = np.meshgrid(i, i, indexing="ij")
    I, J = I.reshape(-1), J.reshape(-1)
    a = I * (n + 1) + J
    b = (I + 1) * (n + 1) + J
    c = I * (n + 1) + J + 1
    d = (I + 1) * (n + 1) + J + 1
    uc = (I + 0.5) / n * 2 - 1
    vc = (J + 0.5) / n * 2 - 1
    same = (uc * vc) > 0  # quadrants where the edge line has slope -1 in (u, v)
    t1 = np.where(same[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t2 = np.where(same[:, None], np.stack([b, d, c], 1), np.stack([a, d, c], 1))
    F = np.concatenate([t1, t2]).astype(np.uint32)
    return _orient_outward(V, F, center)


# ---------------------------------------------------------------------------
# scenes
# ---------------------------------------------------------------------------
def cornell_box(width=512, height=512, spp=64, max_depth=8, rr_depth=3):
    """Cornell box in metres (x to the left of the camera, y up, z into the
    scene), the classic camera (eye (0.278, 0.273, -0.8), fov 39.3)."""
    X, Y, Z = 0.556, 0.549, 0.559
    s = SceneDesc(name="cornell_box")
    white = s.add_bsdf(Bsdf.diffuse((0.73, 0.73, 0.73)))
    red = s.add_bsdf(Bsdf.diffuse((0.63, 0.06, 0.05)))
    green = s.add_bsdf(Bsdf.diffuse((0.14, 0.45, 0.09)))
    walls = [
        ((0, 0, 0), (0, 0, Z), (X, 0, 0), white),   # floor
        ((0, Y, 0), (X, 0, 0), (0, 0, Z), white),   # ceiling
        ((0, 0, Z), (0, Y, 0), (X, 0, 0), white),   # back wall
        ((X, 0, 0), (0, 0, Z), (0, Y, 0), red),     # left wall (x = X)
        ((0, 0, 0), (0, Y, 0), (0, 0, Z), green),   # right wall (x = 0)
    ]
    for p, e0, e1, m in walls:
        V, F, _ = quad(p, e0, e1)
        s.add_mesh(V, F, m, name="wall")
    V, F = box((0.130, 0.0, 0.065), (0.295, 0.165, 0.230))
    s.add_mesh(V, F, white, name="short_box")
    V, F = box((0.265, 0.0, 0.295), (0.430, 0.330, 0.460))
    s.add_mesh(V, F, white, name="tall_box")
    lw, ld = 0.130, 0.105
    V, F, _ = quad((X / 2 - lw / 2, Y - 1e-4, Z / 2 - ld / 2), (lw, 0, 0), (0, 0, ld))
    s.add_mesh(V, F, white, radiance=(17.0, 12.0, 4.0), name="light")
    s.camera = Camera(eye=(0.278, 0.273, -0.8), lookat=(0.278, 0.273, 0.0), up=(0, 1, 0), fov=39.3,
                      width=width, height=height)
    s.rfilter = RFilter(FILTER_GAUSSIAN, 2.0, 0.5)
    s.spp, s.max_depth, s.rr_depth = spp, max_depth, rr_depth
    return s


def opencv_camera(eye, lookat, up, fov, width, height):
    """An "opencv" camera (OpencvCamera) looking from eye at lookat: the
    world-to-camera extrinsic with OpenCV axes (x right, y down, z forward)
    and square pixels whose vertical field of view is fov degrees, the
    principal point at the image centre."""
    eye, lookat, up = (np.asarray(a, np.float64) for a in (eye, lookat, up))
    z = lookat - eye
    z /= np.linalg.norm(z)
    x = np.cross(z, up)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    ext = np.eye(4)
    ext[:3, :3] = R
    ext[:3, 3] = -R @ eye
    f = 0.5 * height / np.tan(np.radians(fov) / 2)
    return Camera(type="opencv", width=width, height=height, extrinsic=ext.astype(np.float32), cx=width / 2,
                  cy=height / 2, fx=float(f), fy=float(f))


def chimney_room(width=64, height=64, spp=64, max_depth=6, rr_depth=3):
    """A closed diffuse room (unit cube, walls facing in) whose only light
    sits at the bottom of an open-topped chimney in the middle of the floor:
    it lights a small patch of the ceiling, and everything the camera sees is
    lit indirectly from that patch -- a scene where directions worth
    sampling are concentrated (used to show what the learned proposal
    gains)."""
    s = SceneDesc(name="chimney_room")
    wall = s.add_bsdf(Bsdf.diffuse((0.75, 0.75, 0.75)))
    dark = s.add_bsdf(Bsdf.diffuse((0.2, 0.2, 0.2)))
    for p, e0, e1 in [((0, 0, 0), (0, 0, 1), (1, 0, 0)),    # floor, +y
                      ((0, 1, 0), (1, 0, 0), (0, 0, 1)),    # ceiling, -y
                      ((0, 0, 1), (0, 1, 0), (1, 0, 0)),    # back, -z
                      ((0, 0, 0), (1, 0, 0), (0, 1, 0)),    # front, +z
                      ((0, 0, 0), (0, 1, 0), (0, 0, 1)),    # x = 0, +x
                      ((1, 0, 0), (0, 0, 1), (0, 1, 0))]:   # x = 1, -x
        V, F, _ = quad(p, e0, e1)
        s.add_mesh(V, F, wall, name="wall")
    # chimney: four outward- and inward-facing plates around the light
    lo, hi, top = 0.42, 0.58, 0.8
    for p, e0, e1 in [((lo, 0, lo), (hi - lo, 0, 0), (0, top, 0)), ((lo, 0, hi), (0, top, 0), (hi - lo, 0, 0)),
                      ((lo, 0, lo), (0, top, 0), (0, 0, hi - lo)), ((hi, 0, lo), (0, 0, hi - lo), (0, top, 0))]:
        V, F, _ = quad(p, e0, e1)
        s.add_mesh(np.concatenate([V, V]), np.concatenate([F, F[:, ::-1] + 4]), dark, name="chimney")
    V, F, _ = quad((lo + 0.01, 0.05, lo + 0.01), (0, 0, hi - lo - 0.02), (hi - lo - 0.02, 0, 0))  # facing +y
    s.add_mesh(V, F, dark, radiance=(3000.0, 2750.0, 2250.0), name="light")
    s.camera = Camera(eye=(0.5, 0.5, 0.02), lookat=(0.5, 0.3, 1.0), up=(0, 1, 0), fov=70, width=width,
                      height=height)
    s.rfilter = RFilter(FILTER_GAUSSIAN, 2.0, 0.5)
    s.spp, s.max_depth, s.rr_depth = spp, max_depth, rr_depth
    return s


def sky_envmap(width=1024, height=512, sun_dir=(0.5, 0.7071, 0.5), sun_peak=2e4, sun_sigma_deg=0.5):
    """Lat-long sky: gradient 0.1 (horizon and below) to 1.0 (zenith) plus a
    Gaussian sun lobe (peak sun_peak, sigma sun_sigma_deg)."""
    y = (np.arange(height) + 0.5) / height
    x = (np.arange(width) + 0.5) / width
    theta = y * math.pi
    phi = x * 2 * math.pi - math.pi
    T, P = np.meshgrid(theta, phi, indexing="ij")
    d = np.stack([np.cos(P) * np.sin(T), np.cos(T), np.sin(P) * np.sin(T)], -1)
    elev = np.clip(d[..., 1], 0.0, 1.0)
    g = 0.1 + 0.9 * elev
    rgb = np.stack([g * 0.55, g * 0.75, g * 1.0], -1)
    sd = np.asarray(sun_dir, np.float64)
    sd /= np.linalg.norm(sd)
    ang = np.arccos(np.clip((d * sd).sum(-1), -1.0, 1.0))
    sig = math.radians(sun_sigma_deg)
    sun = sun_peak * np.exp(-0.5 * (ang / sig) ** 2)
    rgb += sun[..., None] * np.array([1.0, 0.95, 0.85])
    return rgb.astype(np.float32)


def displaced_sphere(width=1920, height=1080, spp=256, max_depth=16, grid=1024, glass_level=5, env_size=(1024, 512),
                     seed=0x5EED):
    """2,097,152-triangle displaced sphere (grid=1024) under a sky env map."""
    s = SceneDesc(name="displaced_sphere")
    cond = s.add_bsdf(Bsdf.conductor(0.2, (0.2, 0.92, 1.1), (3.912, 2.452, 2.142)))
    glass = s.add_bsdf(Bsdf.dielectric(1.5, 1.0))
    ground = s.add_bsdf(Bsdf.diffuse((0.5, 0.5, 0.5)))
    V, F = octa_sphere(grid, radius=1.0, displace=0.1, seed=seed)
    s.add_mesh(V, F, cond, name="displaced_sphere")
    V, F = icosphere(glass_level, radius=0.5, center=(1.7, -0.6, 0.9))
    s.add_mesh(V, F, glass, name="glass_sphere")
    V, F, _ = quad((-10.0, -1.1, -10.0), (0, 0, 20.0), (20.0, 0, 0))
    s.add_mesh(V, F, ground, name="ground")
    s.envmap = EnvMap(sky_envmap(*env_size), 1.0, True)
    s.camera = Camera(eye=(0.4, 0.9, 5.0), lookat=(0.4, -0.1, 0.0), up=(0, 1, 0), fov=38.0, width=width,
                      height=height)
    s.rfilter = RFilter(FILTER_GAUSSIAN, 2.0, 0.5)
    s.spp, s.max_depth, s.rr_depth = spp, max_depth, 4
    return s


def checker_texture(w, h, seed, n=8):
    """(h, w, 3) float texture: an n x n checkerboard of seeded colours times
    a smooth ramp (exercises bilinear weights and wrap addressing)."""
    rng = np.random.default_rng(seed)
    cols = rng.uniform(0.1, 0.9, size=(n, n, 3)).astype(np.float32)
    y, x = np.mgrid[0:h, 0:w]
    c = cols[(y * n) // h, (x * n) // w]
    ramp = (0.6 + 0.4 * np.sin(2 * np.pi * x / w) * np.cos(2 * np.pi * y / h)).astype(np.float32)
    return (c * ramp[..., None]).astype(np.float32)


def normal_map(w, h, seed, amp=0.35):
    """(h, w, 3) tangent-space normal map (stored as n * 0.5 + 0.5) of
    seeded bumps."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w] / np.array([h, w])[:, None, None]
    gx = np.zeros((h, w))
    gy = np.zeros((h, w))
    for _ in range(6):
        fx, fy, ph = rng.uniform(1, 4), rng.uniform(1, 4), rng.uniform(0, 2 * np.pi)
        gx += amp * np.cos(2 * np.pi * (fx * x + fy * y) + ph) * fx / 4
        gy += amp * np.cos(2 * np.pi * (fx * x + fy * y) + ph) * fy / 4
    n = np.stack([-gx, -gy, np.ones_like(gx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return (n * 0.5 + 0.5).astype(np.float32)


def textured_box(width=64, height=64, spp=16, max_depth=6, rr_depth=3):
    """Cornell-like box whose surfaces use every texture input: a floor with
    an albedo texture (texture coordinates 0..3, wrapped), a back wall of
    anisotropic GGX with pd / axay textures and a tangent-space normal map,
    a left wall of anisotropic GGX with object-space normal and tangent maps."""
    s = SceneDesc(name="textured_box")
    t_alb = s.add_texture(checker_texture(48, 40, 1))
    t_pd = s.add_texture(checker_texture(32, 32, 2, n=4))
    axay = np.stack(np.meshgrid(np.linspace(0.05, 0.6, 16), np.linspace(0.6, 0.05, 16)), -1)
    t_axay = s.add_texture(np.concatenate([axay, np.zeros((16, 16, 1))], -1))
    t_nrm = s.add_texture(normal_map(64, 64, 3))
    on = np.array([0.3, 0.2, 1.0]) / np.linalg.norm([0.3, 0.2, 1.0])
    ot = np.cross(on, [0.0, 1.0, 0.0])
    ot /= np.linalg.norm(ot)
    t_on = s.add_texture(np.tile((on * 0.5 + 0.5).astype(np.float32), (8, 8, 1)))
    t_ot = s.add_texture(np.tile((ot * 0.5 + 0.5).astype(np.float32), (8, 8, 1)))
    white = s.add_bsdf(Bsdf.diffuse((0.73, 0.73, 0.73)))
    red = s.add_bsdf(Bsdf.diffuse((0.63, 0.06, 0.05)))
    floor = s.add_bsdf(Bsdf(BSDF_DIFFUSE, (0.5, 0.5, 0.5), textures={TEX_ALBEDO: t_alb}))
    back = s.add_bsdf(Bsdf(BSDF_ANISO_GGX, (0.4, 0.4, 0.4, 0.5, 0.5, 0.5, 0.3, 0.1),
                           textures={TEX_PD: t_pd, TEX_AXAY: t_axay, TEX_NORMAL: t_nrm}))
    left = s.add_bsdf(Bsdf(BSDF_ANISO_GGX, (0.3, 0.5, 0.2, 0.4, 0.4, 0.4, 0.2, 0.05),
                           textures={TEX_NORMAL: t_on, TEX_TANGENT: t_ot}, maps_object_space=True))
    lamp = s.add_bsdf(Bsdf.diffuse((0.0, 0.0, 0.0)))

    def textured_quad(p, e0, e1, bsdf, uv_scale=1.0):
        V, F, UV = quad(p, e0, e1, with_uv=True)
        n = np.cross(np.asarray(e0, np.float64), np.asarray(e1, np.float64))
        N = np.tile((n / np.linalg.norm(n)).astype(np.float32), (4, 1))
        s.add_mesh(V, F, bsdf, N=N, UV=UV * np.float32(uv_scale))

    textured_quad((-1, -1, -1), (0, 0, 2), (2, 0, 0), floor, uv_scale=3.0)
    textured_quad((-1, -1, -1), (2, 0, 0), (0, 2, 0), back)
    textured_quad((-1, -1, 1), (0, 0, -2), (0, 2, 0), left)
    V, F, _ = quad((1, -1, -1), (0, 0, 2), (0, 2, 0))
    s.add_mesh(V, F, red)
    V, F, _ = quad((-1, 1, -1), (2, 0, 0), (0, 0, 2))
    s.add_mesh(V, F, white)
    V, F, _ = quad((-0.25, 0.999, -0.25), (0.5, 0, 0), (0, 0, 0.5))
    s.add_mesh(V, F, lamp, radiance=(12.0, 10.0, 8.0))
    V, F = box((0.1, -1.0, -0.3), (0.6, -0.2, 0.2))
    s.add_mesh(V, F, white)
    s.camera = Camera(eye=(0.0, 0.0, 3.8), lookat=(0.0, 0.0, 0.0), up=(0, 1, 0), fov=39.3, width=width, height=height)
    s.rfilter = RFilter(FILTER_GAUSSIAN, 2.0, 0.5)
    s.spp, s.max_depth, s.rr_depth = spp, max_depth, rr_depth
    return s


def mis_plates(width=1280, height=720, spp=512, max_depth=6):
    """Veach-style MIS scene: glossy plates of increasing roughness lit by
    four spherical lights of equal power."""
    s = SceneDesc(name="mis_plates")
    back = s.add_bsdf(Bsdf.diffuse((0.1, 0.1, 0.1)))
    lamp = s.add_bsdf(Bsdf.diffuse((0.0, 0.0, 0.0)))
    rough = [(0.005, 0.005), (0.02, 0.02), (0.05, 0.05), (0.1, 0.01)]
    plates = [s.add_bsdf(Bsdf.aniso_ggx((0, 0, 0), (1.0, 1.0, 1.0), r)) for r in rough]
    # plates: 8 x 0.9 strips stepping down and toward the camera, tilted up
    for i, m in enumerate(plates):
        y = -0.7 - 0.65 * i
        z = 0.9 * i
        ang = math.radians(38.0 - 7.0 * i)
        e1 = (0.0, 0.9 * math.sin(ang), 0.9 * math.cos(ang))
        V, F, UV = quad((4.0, y, z), (-8.0, 0.0, 0.0), e1, with_uv=True)
        s.add_mesh(V, F, m, UV=UV, name="plate%d" % i)
    V, F, _ = quad((-20.0, -10.0, -3.0), (40.0, 0, 0), (0, 30.0, 0))
    s.add_mesh(V, F, back, name="backdrop")
    V, F, _ = quad((-20.0, -5.0, -3.0), (0, 0, 25.0), (40.0, 0, 0))
    s.add_mesh(V, F, back, name="floor")
    radii = [0.03, 0.1, 0.3, 0.9]
    for i, r in enumerate(radii):
        V, F = icosphere(2, radius=r, center=(-3.75 + 2.5 * i, 1.6, -1.2))
        L = 0.8 / (r * r)
        s.add_mesh(V, F, lamp, radiance=(L, L, L), name="light%d" % i)
    s.camera = Camera(eye=(0.0, 2.2, 9.0), lookat=(0.0, -1.6, 1.0), up=(0, 1, 0), fov=38.0, width=width,
                      height=height)
    s.rfilter = RFilter(FILTER_GAUSSIAN, 2.0, 0.5)
    s.spp, s.max_depth, s.rr_depth = spp, max_depth, 4
    return s


def material_zoo(width=64, height=48, spp=16, max_depth=10, filter_radius=2.0, filter_sigma=0.5):
    """Every BSDF of the reference side by side: a mirror sphere, Beckmann
    microfacet spheres (the default kd 0.5 and a rough one with a coloured
    kd and extIOR 1.2), a microfacet box with kd 0 (specular only), a glass
    sphere with intIOR 1.7 inside a medium of extIOR 1.33, a GGX conductor
    sphere, a diffuse floor and a default-microfacet back wall, lit by an
    area light and a point light."""
    s = SceneDesc(name="material_zoo")
    floor = s.add_bsdf(Bsdf.diffuse((0.6, 0.55, 0.5)))
    wall = s.add_bsdf(Bsdf.microfacet())
    mirror = s.add_bsdf(Bsdf.mirror())
    mf_rough = s.add_bsdf(Bsdf.microfacet(0.4, 1.7, 1.2, (0.2, 0.3, 0.1)))
    mf_spec = s.add_bsdf(Bsdf.microfacet(0.25, 1.5, 1.0, (0.0, 0.0, 0.0)))
    glass = s.add_bsdf(Bsdf.dielectric(1.7, 1.33))
    cond = s.add_bsdf(Bsdf.conductor(0.15))
    lamp = s.add_bsdf(Bsdf.diffuse((0.0, 0.0, 0.0)))
    V, F, _ = quad((-3.0, 0.0, -3.0), (0, 0, 6.0), (6.0, 0, 0))
    s.add_mesh(V, F, floor, name="floor")
    V, F, _ = quad((-3.0, 0.0, -2.0), (6.0, 0, 0), (0, 4.0, 0))
    s.add_mesh(V, F, wall, name="wall")
    spheres = [(mirror, (-1.6, 0.5, -0.6)), (mf_rough, (-0.5, 0.5, -0.9)), (glass, (0.6, 0.5, -0.4)),
               (cond, (1.7, 0.5, -0.8))]
    for m, c in spheres:
        V, F = icosphere(3, radius=0.45, center=c)
        s.add_mesh(V, F, m, name="sphere")
    V, F = box((-0.9, 0.0, 0.4), (-0.3, 0.5, 1.0))
    s.add_mesh(V, F, mf_spec, name="box")
    V, F = icosphere(2, radius=0.3, center=(0.4, 0.35, 0.8))
    s.add_mesh(V, F, wall, name="default_microfacet_sphere")
    V, F, _ = quad((-0.6, 3.2, -0.8), (1.2, 0, 0), (0, 0, 0.9))
    s.add_mesh(V, F, lamp, radiance=(14.0, 12.0, 9.0), name="light")
    s.point_lights.append(((2.0, 2.5, 2.0), (6.0, 7.0, 9.0)))
    s.camera = Camera(eye=(0.0, 2.0, 5.0), lookat=(0.0, 0.6, 0.0), up=(0, 1, 0), fov=45.0, width=width,
                      height=height)
    s.rfilter = RFilter(FILTER_GAUSSIAN, float(filter_radius), float(filter_sigma))
    s.spp, s.max_depth, s.rr_depth = spp, max_depth, 3
    return s


def random_scene(seed, target_tris=50_000, width=512, height=512, spp=64):
    """Procedural job scene: displaced spheres and boxes with random diffuse /
    GGX-conductor / dielectric materials and 1-8 spherical area lights."""
    rng = np.random.default_rng(seed)
    s = SceneDesc(name="random_%d" % seed)
    floor = s.add_bsdf(Bsdf.diffuse((0.4, 0.4, 0.4)))
    V, F, _ = quad((-8.0, 0.0, -8.0), (0, 0, 16.0), (16.0, 0, 0))
    s.add_mesh(V, F, floor, name="floor")
    budget = int(target_tris)
    nobj = int(rng.integers(3, 9))
    per = max(200, budget // nobj)
    for k in range(nobj):
        kind = rng.integers(0, 3)
        if kind == 0:
            m = s.add_bsdf(Bsdf.diffuse(tuple(rng.uniform(0.1, 0.9, 3))))
        elif kind == 1:
            m = s.add_bsdf(Bsdf.conductor(float(rng.uniform(0.05, 0.5))))
        else:
            m = s.add_bsdf(Bsdf.dielectric(float(rng.uniform(1.3, 1.8)), 1.0))
        c = (float(rng.uniform(-3, 3)), float(rng.uniform(0.5, 1.5)), float(rng.uniform(-3, 3)))
        if rng.random() < 0.7:
            n = max(4, int(math.sqrt(per / 2)))
            V, F = octa_sphere(n, radius=float(rng.uniform(0.3, 0.8)), center=c, displace=float(rng.uniform(0, 0.15)),
                               seed=int(rng.integers(1 << 30)))
        else:
            h = rng.uniform(0.2, 0.7, 3)
            V, F = box(np.array(c) - h, np.array(c) + h)
        s.add_mesh(V, F, m, name="obj%d" % k)
    lamp = s.add_bsdf(Bsdf.diffuse((0.5, 0.5, 0.5)))
    for k in range(int(rng.integers(1, 9))):
        r = float(rng.uniform(0.1, 0.4))
        V, F = icosphere(2, radius=r, center=(float(rng.uniform(-4, 4)), float(rng.uniform(3, 5)),
                                               float(rng.uniform(-4, 4))))
        L = tuple(float(x) for x in rng.uniform(2.0, 10.0, 3) / (r * r) * 0.05)
        s.add_mesh(V, F, lamp, radiance=L, name="light%d" % k)
    s.camera = Camera(eye=(0.0, 4.0, 10.0), lookat=(0.0, 0.8, 0.0), up=(0, 1, 0), fov=40.0, width=width,
                      height=height)
    s.spp, s.max_depth, s.rr_depth = spp, 8, 4
    return s


def from_name(name, **kw):
    return {"cornell_box": cornell_box, "displaced_sphere": displaced_sphere, "mis_plates": mis_plates}[name](**kw)


SERVICE_RESOLUTIONS = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))


def service_scene_pool():
    """Scenes a render service job draws from: the Cornell box, the MIS
    plates, the 2,097,152-triangle displaced sphere and procedural scenes of
    10k to 500k triangles."""
    pool = [cornell_box(), mis_plates(), displaced_sphere()]
    for k, tris in enumerate((10_000, 50_000, 200_000, 500_000)):
        pool.append(random_scene(100 + k, tris))
    return pool


def service_jobs(n, seed=0x5EED, pool_size=7):
    """n jobs (scene index, width, height, spp): resolution uniform over
    SERVICE_RESOLUTIONS, spp = floor(16 * 64^U) with U uniform on [0, 1)
    (16 to 1024, mean about 242), scene uniform over the pool."""
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        w, h = SERVICE_RESOLUTIONS[int(rng.integers(0, len(SERVICE_RESOLUTIONS)))]
        spp = int(16 * 64 ** rng.random())
        jobs.append((int(rng.integers(0, pool_size)), w, h, spp))
    return jobs
