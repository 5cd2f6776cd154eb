"""CPU oracle for the DrawLab hot path -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference CPU path (see oracle/oracle.cpp).  Only the
test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this package, and only as the checker or as the timed CPU
baseline; the product (``drawlab_amd``) never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

BSDF_DIFFUSE, BSDF_MIRROR, BSDF_DIELECTRIC, BSDF_MICROFACET, BSDF_ANISO_GGX, BSDF_CONDUCTOR = range(6)
FILTER_GAUSSIAN, FILTER_BOX, FILTER_TENT, FILTER_MITCHELL = range(4)
ACCEL_BRUTE, ACCEL_OCTREE, ACCEL_OCTREE_LOWEST, ACCEL_OCTREE_ROBUST = range(4)

_lib = None
_ref = None

_P = ctypes.c_void_p
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        u32, i32, f32 = ctypes.c_uint32, ctypes.c_int, ctypes.c_float
        sig = {
            "dlo_scene_new": (_P, []),
            "dlo_scene_free": (None, [_P]),
            "dlo_scene_add_bsdf": (i32, [_P, i32, _f32p, i32]),
            "dlo_scene_add_texture": (i32, [_P, _f32p, i32, i32]),
            "dlo_scene_set_bsdf_textures": (None, [_P, i32, _i32p, i32]),
            "dlo_texture_sample": (None, [_f32p, i32, i32, _f32p, u32, _f32p]),
            "dlo_scene_add_mesh": (i32, [_P, _f32p, u32, _u32p, u32, _P, _P, i32, _P]),
            "dlo_scene_add_point_light": (i32, [_P, _f32p, _f32p]),
            "dlo_scene_add_envmap": (i32, [_P, _f32p, i32, i32, f32, i32]),
            "dlo_scene_set_camera_virtual": (None, [_P, _f32p, _f32p, _f32p, f32, i32, i32]),
            "dlo_scene_set_camera_perspective": (None, [_P, _f32p, _f32p, i32, i32, f32, f32]),
            "dlo_scene_set_camera_basis": (None, [_P, _f32p, i32, i32]),
            "dlo_scene_set_filter_table": (None, [_P, f32, _f32p]),
            "dlo_scene_get_camera": (None, [_P, _f32p]),
            "dlo_scene_set_filter": (None, [_P, i32, f32, f32, f32]),
            "dlo_scene_get_filter_table": (None, [_P, _f32p]),
            "dlo_scene_set_background": (None, [_P, _f32p]),
            "dlo_scene_set_camera_opencv": (None, [_P, _f32p, _f32p, _f32p, i32, i32]),
            "dlo_scene_set_options": (None, [_P, i32, i32, f32]),
            "dlo_scene_build": (i32, [_P]),
            "dlo_scene_num_prims": (u32, [_P]),
            "dlo_trace_closest": (None, [_P, i32, _f32p, u32, _u32p, _f32p, _f32p, _f32p]),
            "dlo_trace_shadow": (None, [_P, i32, _f32p, u32, _u8p]),
            "dlo_camera_rays": (None, [_P, _f32p, u32, _f32p]),
            "dlo_render_samples": (None, [_P, _i32p, u32, u32, i32, i32, i32, _f32p]),
            "dlo_render": (i32, [_P, i32, u32, i32, i32, i32, i32, i32, _P, _P]),
            "dlo_pcg32": (None, [ctypes.c_uint64, ctypes.c_uint64, u32, _u32p]),
            "dlo_pcg32_float": (None, [ctypes.c_uint64, ctypes.c_uint64, u32, _f32p]),
            "dlo_warp": (None, [i32, _f32p, u32, f32, f32, _f32p]),
            "dlo_optics": (None, [i32, _f32p, u32, _f32p]),
            "dlo_bsdf_query": (None, [_P, i32, _f32p, u32, _f32p]),
            "dlo_emitter_query": (None, [_P, _f32p, u32, _f32p]),
            "dlo_set_mutation": (None, [i32]),
            "dlo_env_tables": (None, [_P, _f32p, _u32p, _f32p]),
            "dlo_spiral_blocks": (u32, [i32, i32, _P]),
            "dlo_bvh4_replay": (None, [_f32p, _f32p, _f32p, u32, i32, _P, _P, _P, _P, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's own wrap.cpp/math.cpp (oracle/_ref), or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = ctypes.CDLL(REF_PATH)
        for name in ("ref_warp", "ref_optics", "ref_vec"):
            getattr(L, name).restype = None
        L.ref_warp.argtypes = [ctypes.c_int, _f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _f32p]
        L.ref_optics.argtypes = [ctypes.c_int, _f32p, ctypes.c_uint32, _f32p]
        L.ref_vec.argtypes = [ctypes.c_int, _f32p, ctypes.c_uint32, _f32p]
        _ref = L
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
    if n is not None:
        assert a.size == n
    return a


class Scene:
    """Oracle scene built from a plain scene description (drawlab_amd.scene.SceneDesc-like)."""

    def __init__(self):
        self.L = lib()
        self.h = self.L.dlo_scene_new()
        self.width = self.height = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.dlo_scene_free(self.h)
            self.h = None

    def add_bsdf(self, btype, params=()):
        p = _f(list(params) + [0.0] * (16 - len(params)))
        return self.L.dlo_scene_add_bsdf(self.h, int(btype), p, len(params))

    def add_texture(self, rgb):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        return self.L.dlo_scene_add_texture(self.h, rgb.reshape(-1), int(rgb.shape[1]), int(rgb.shape[0]))

    def set_bsdf_textures(self, bsdf, slots, object_space=False):
        t = np.zeros(5, np.int32)
        for slot, idx in slots.items():
            t[slot] = int(idx) + 1
        self.L.dlo_scene_set_bsdf_textures(self.h, int(bsdf), t, int(bool(object_space)))

    def add_mesh(self, V, F, bsdf, N=None, UV=None, radiance=None):
        V = _f(V)
        F = np.ascontiguousarray(np.asarray(F, dtype=np.uint32).reshape(-1))
        N = None if N is None else _f(N)
        UV = None if UV is None else _f(UV)
        rad = None if radiance is None else _f(radiance, 3)
        return self.L.dlo_scene_add_mesh(self.h, V, V.size // 3, F, F.size // 3, _ptr(N), _ptr(UV), int(bsdf), _ptr(rad))

    def add_point_light(self, position, intensity):
        return self.L.dlo_scene_add_point_light(self.h, _f(position, 3), _f(intensity, 3))

    def add_envmap(self, rgb, scale=1.0, visual=True):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        h, w = rgb.shape[:2]
        return self.L.dlo_scene_add_envmap(self.h, rgb.reshape(-1), w, h, float(scale), int(bool(visual)))

    def set_camera(self, eye, lookat, up, fov, width, height):
        self.width, self.height = int(width), int(height)
        self.L.dlo_scene_set_camera_virtual(self.h, _f(eye, 3), _f(lookat, 3), _f(up, 3), float(fov), int(width), int(height))

    def set_camera_perspective(self, sample_to_camera, camera_to_world, width, height, near, far):
        """PerspectiveCamera after activate(): the two row-major matrices are
        inputs."""
        self.width, self.height = int(width), int(height)
        self.L.dlo_scene_set_camera_perspective(self.h, _f(sample_to_camera, 16), _f(camera_to_world, 16),
                                                int(width), int(height), float(near), float(far))

    def set_camera_opencv(self, camera_to_world, eye, intrinsics, width, height):
        """OpencvCamera: the inverse extrinsic matrix (row-major), the eye and
        (cx, cy, fx, fy) are inputs."""
        self.width, self.height = int(width), int(height)
        self.L.dlo_scene_set_camera_opencv(self.h, _f(camera_to_world, 16), _f(eye, 3), _f(intrinsics, 4),
                                           int(width), int(height))

    def set_options(self, normals=False, hide_lights=False, eps=0.0):
        """normals: the NormalIntegrator; hide_lights: area-light surfaces
        transparent to path rays; eps: ray offset (0 = 1e-4)."""
        self.L.dlo_scene_set_options(self.h, int(bool(normals)), int(bool(hide_lights)), float(eps))

    def camera(self):
        out = np.zeros(12, np.float32)
        self.L.dlo_scene_get_camera(self.h, out)
        return out.reshape(4, 3)

    def set_filter(self, ftype=FILTER_GAUSSIAN, radius=2.0, p0=0.5, p1=0.0):
        self.L.dlo_scene_set_filter(self.h, int(ftype), float(radius), float(p0), float(p1))

    def filter_table(self):
        out = np.zeros(33, np.float32)
        self.L.dlo_scene_get_filter_table(self.h, out)
        return out

    def set_background(self, rgb):
        self.L.dlo_scene_set_background(self.h, _f(rgb, 3))

    def build(self):
        return self.L.dlo_scene_build(self.h)

    def num_prims(self):
        return self.L.dlo_scene_num_prims(self.h)

    def trace_closest(self, rays, mode=ACCEL_BRUTE):
        rays = _f(rays)
        n = rays.size // 8
        prim = np.zeros(n, np.uint32)
        t, u, v = (np.zeros(n, np.float32) for _ in range(3))
        self.L.dlo_trace_closest(self.h, int(mode), rays, n, prim, t, u, v)
        return prim, t, u, v

    def trace_shadow(self, rays, mode=ACCEL_BRUTE):
        rays = _f(rays)
        n = rays.size // 8
        occ = np.zeros(n, np.uint8)
        self.L.dlo_trace_shadow(self.h, int(mode), rays, n, occ)
        return occ

    def camera_rays(self, pos):
        pos = _f(pos)
        n = pos.size // 2
        out = np.zeros(8 * n, np.float32)
        self.L.dlo_camera_rays(self.h, pos, n, out)
        return out.reshape(8, n)

    def render_samples(self, xys, seed=0, max_depth=-1, rr_depth=4, accel=ACCEL_OCTREE_ROBUST):
        xys = np.ascontiguousarray(np.asarray(xys, dtype=np.int32).reshape(-1))
        n = xys.size // 3
        out = np.zeros(5 * n, np.float32)
        self.L.dlo_render_samples(self.h, xys, n, int(seed), int(max_depth), int(rr_depth), int(accel), out)
        return out.reshape(n, 5)

    def render(self, spp, seed=0, max_depth=-1, rr_depth=4, accel=ACCEL_OCTREE_ROBUST, threads=0, max_blocks=0,
               want_image=True):
        img = np.zeros((self.height, self.width, 3), np.float32) if want_image else None
        stats = np.zeros(8, np.float64)
        rc = self.L.dlo_render(self.h, int(spp), int(seed), int(max_depth), int(rr_depth), int(accel), int(threads),
                               int(max_blocks), _ptr(img), _ptr(stats))
        if rc != 0:
            raise RuntimeError("oracle render failed (%d)" % rc)
        return img, {"paths": stats[0], "closest": stats[1], "shadow": stats[2], "seconds": stats[3],
                     "threads": int(stats[4])}

    def bsdf_query(self, bsdf, wi, wo, s):
        inp = np.concatenate([np.asarray(wi, np.float32).reshape(-1, 3), np.asarray(wo, np.float32).reshape(-1, 3),
                              np.asarray(s, np.float32).reshape(-1, 2)], axis=1)
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        n = inp.shape[0]
        out = np.zeros(12 * n, np.float32)
        self.L.dlo_bsdf_query(self.h, int(bsdf), inp.reshape(-1), n, out)
        return out.reshape(n, 12)

    def emitter_query(self, p, n, u):
        """Scene::sampleEmitterDirection at shading points p with shading
        normals n and uniforms u (n, 3); returns a dict of spectrum (n, 3),
        d (n, 3), dist, pdf (the sample's), pdf_dir (pdfEmitterDirection of
        the sampled direction), emitter, delta."""
        inp = np.concatenate([np.asarray(p, np.float32).reshape(-1, 3), np.asarray(n, np.float32).reshape(-1, 3),
                              np.asarray(u, np.float32).reshape(-1, 3)], axis=1)
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        k = inp.shape[0]
        out = np.zeros(11 * k, np.float32)
        self.L.dlo_emitter_query(self.h, inp.reshape(-1), k, out)
        o = out.reshape(k, 11)
        return {"spectrum": o[:, 0:3], "d": o[:, 3:6], "dist": o[:, 6], "pdf": o[:, 7], "pdf_dir": o[:, 8],
                "emitter": o[:, 9].astype(np.int32), "delta": o[:, 10] != 0}

    def env_tables(self, w, h):
        q = np.zeros(w * h, np.float32)
        a = np.zeros(w * h, np.uint32)
        p = np.zeros(w * h, np.float32)
        self.L.dlo_env_tables(self.h, q, a, p)
        return q, a, p


# Planted mistakes (oracle.h DLO_MUT_*): only the tests that show a pin of
# the oracle fails on a wrong oracle set them.
MUT_NONE, MUT_GGX_PDF_NO_WHZ, MUT_DIELECTRIC_SWAP_IOR, MUT_DIFFUSE_UNIFORM_PDF, MUT_EMITTER_HIT_NO_MIS, \
    MUT_AREA_PDF_NO_DIST2, MUT_CONDUCTOR_SWAP_ETA_K, MUT_RR_NO_DIVIDE, MUT_MICROFACET_G_WI_TWICE, MUT_POINT_INV_DIST, \
    MUT_ENV_PDF_TIMES_AREA, MUT_DIELECTRIC_NO_FRESNEL = range(12)


class mutation:
    """Context manager: the oracle with a planted mistake."""

    def __init__(self, m):
        self.m = int(m)

    def __enter__(self):
        lib().dlo_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().dlo_set_mutation(MUT_NONE)
        return False


def perspective_matrices(fov, near, far, width, height, to_world=None):
    """PerspectiveCamera::activate (src/tracer/camera/perspective.cpp:42-70):
    sampleToCamera = (scale(-0.5, -0.5 aspect, 1) * translate(-1, -1/aspect, 0)
    * perspective(cot, near, far))^-1 and cameraToWorld, both row-major.
    Evaluated in float32 with a float64 inverse, so it agrees with the
    product's set-up to rounding (checked by the tests), not bit for bit."""
    f = np.float32
    aspect = f(width) / f(height)
    recip = f(1.0) / (f(far) - f(near))
    cot = f(1.0) / f(np.tan(f(fov) / f(2.0) * f(np.pi / 180.0)))
    P = np.array([[cot, 0, 0, 0], [0, cot, 0, 0], [0, 0, f(far) * recip, -f(near) * f(far) * recip], [0, 0, 1, 0]],
                 dtype=np.float32)
    S = np.diag(np.array([-0.5, -0.5 * aspect, 1.0, 1.0], dtype=np.float32))
    T = np.eye(4, dtype=np.float32)
    T[0, 3], T[1, 3] = -1.0, f(-1.0) / aspect
    M = (S @ T @ P).astype(np.float32)
    inv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    C = np.eye(4, dtype=np.float32) if to_world is None else np.asarray(to_world, np.float32).reshape(4, 4)
    return inv.reshape(16), C.reshape(16)


def from_desc(desc, build=True, camera_matrices=None):
    """Build an oracle Scene from a drawlab_amd scene description.
    camera_matrices: (sampleToCamera, cameraToWorld) of a perspective camera,
    used instead of perspective_matrices() (e.g. the values the product's
    set-up computed, so both sides trace identical rays)."""
    s = Scene()
    for rgb in getattr(desc, "textures", []):
        s.add_texture(rgb)
    for b in desc.bsdfs:
        bi = s.add_bsdf(b.type, b.params)
        if b.textures:
            s.set_bsdf_textures(bi, b.textures, b.maps_object_space)
    # Emitters are registered in the scene's emitter order: an area light
    # when its mesh is added (meshes keep their order), point lights and the
    # environment map where the order places them.
    order = desc.emitter_order() if hasattr(desc, "emitter_order") else None
    if order is None:
        order = [("area", i) for i, m in enumerate(desc.meshes) if m.radiance is not None]
        order += [("point", i) for i in range(len(desc.point_lights))]
        order += [("env", 0)] if desc.envmap is not None else []
    pending = list(order)

    def add_other(kind, idx):
        if kind == "point":
            s.add_point_light(desc.point_lights[idx][0], desc.point_lights[idx][1])
        else:
            s.add_envmap(desc.envmap.rgb, desc.envmap.scale, desc.envmap.visual)

    for mi, m in enumerate(desc.meshes):
        if m.radiance is not None:
            while pending and pending[0] != ("area", mi):
                kind, idx = pending.pop(0)
                assert kind != "area", "area emitters must follow mesh order"
                add_other(kind, idx)
            assert pending and pending[0] == ("area", mi)
            pending.pop(0)
        s.add_mesh(m.V, m.F, m.bsdf, m.N, m.UV, m.radiance)
    for kind, idx in pending:
        assert kind != "area"
        add_other(kind, idx)
    c = desc.camera
    cd = getattr(desc, "camera_desc", None)  # an exported camera (XML scenes)
    if cd is not None and cd.type == 1:
        s.set_camera_perspective(np.array(cd.sample_to_camera[:], np.float32),
                                 np.array(cd.camera_to_world[:], np.float32), cd.width, cd.height, cd.near_clip,
                                 cd.far_clip)
    elif cd is not None and cd.type == 2:
        s.set_camera_opencv(np.array(cd.camera_to_world[:], np.float32), np.array(cd.eye[:], np.float32),
                            np.array(cd.intrinsics[:], np.float32), cd.width, cd.height)
    elif cd is not None:
        s.L.dlo_scene_set_camera_basis(s.h, _f(list(cd.eye) + list(cd.U) + list(cd.V) + list(cd.W), 12),
                                       int(cd.width), int(cd.height))
        s.width, s.height = int(cd.width), int(cd.height)
    elif getattr(c, "type", "virtual") == "opencv":
        # camera_matrices: (inverse extrinsic, eye) as the product computed
        # them; otherwise a float64 inverse rounded to float32
        if camera_matrices is None:
            inv = np.linalg.inv(np.asarray(c.extrinsic, np.float64).reshape(4, 4)).astype(np.float32)
            camera_matrices = (inv.reshape(16), inv[:3, 3].copy())
        s.set_camera_opencv(camera_matrices[0], camera_matrices[1], (c.cx, c.cy, c.fx, c.fy), c.width, c.height)
    elif getattr(c, "type", "virtual") == "perspective":
        if camera_matrices is None:
            camera_matrices = perspective_matrices(c.fov, c.near, c.far, c.width, c.height, c.to_world)
        s.set_camera_perspective(camera_matrices[0], camera_matrices[1], c.width, c.height, c.near, c.far)
    else:
        s.set_camera(c.eye, c.lookat, c.up, c.fov, c.width, c.height)
    film = getattr(desc, "film", None)  # an exported filter table (XML scenes)
    if film is not None:
        s.L.dlo_scene_set_filter_table(s.h, float(film.radius), _f(film.table[:], 33))
    else:
        f = desc.rfilter
        s.set_filter(f.type, f.radius, f.p0, f.p1)
    s.set_background(desc.bg_color)
    if build:
        s.build()
    return s


def texture_sample(rgb, uv):
    """Bilinear, wrap-addressed lookups of an (h, w, 3) texture at (n, 2)
    normalized coordinates (the sampler of textured materials)."""
    rgb = np.ascontiguousarray(rgb, np.float32)
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((uv.shape[0], 3), np.float32)
    lib().dlo_texture_sample(rgb.reshape(-1), int(rgb.shape[1]), int(rgb.shape[0]), uv.reshape(-1), uv.shape[0],
                             out.reshape(-1))
    return out


def bvh4_replay(nodes, tris, rays, any_hit=False):
    """CPU replay of the HIP backend's traversal over its exported 4-wide
    nodes and triangles (oracle/replay.cpp).  Returns (prim, t, u, v,
    node_visits, tri_tests) for closest hits, (occ, node_visits, tri_tests)
    for any hits."""
    rays = np.ascontiguousarray(np.asarray(rays, np.float32).reshape(8, -1))
    n = rays.shape[1]
    nodes = np.ascontiguousarray(nodes, np.float32)
    tris = np.ascontiguousarray(tris, np.float32)
    cn = np.zeros(n, np.uint32)
    ct = np.zeros(n, np.uint32)
    if any_hit:
        occ = np.zeros(n, np.uint8)
        lib().dlo_bvh4_replay(nodes, tris, rays, n, 1, None, None, None, None, _ptr(occ), _ptr(cn), _ptr(ct))
        return occ, cn, ct
    prim = np.zeros(n, np.uint32)
    t, u, v = (np.zeros(n, np.float32) for _ in range(3))
    lib().dlo_bvh4_replay(nodes, tris, rays, n, 0, _ptr(prim), _ptr(t), _ptr(u), _ptr(v), None, _ptr(cn), _ptr(ct))
    return prim, t, u, v, cn, ct


def pcg32(initstate, initseq, n):
    out = np.zeros(n, np.uint32)
    lib().dlo_pcg32(initstate, initseq, n, out)
    return out


def pcg32_float(initstate, initseq, n):
    out = np.zeros(n, np.float32)
    lib().dlo_pcg32_float(initstate, initseq, n, out)
    return out


def warp(which, inp, a0=0.0, a1=0.0):
    inp = _f(inp)
    n = inp.size // 3
    out = np.zeros(3 * n, np.float32)
    lib().dlo_warp(int(which), inp, n, float(a0), float(a1), out)
    return out.reshape(n, 3)


def optics(which, inp):
    inp = _f(inp)
    n = inp.size // 6
    out = np.zeros(3 * n, np.float32)
    lib().dlo_optics(int(which), inp, n, out)
    return out.reshape(n, 3)


def spiral_blocks(W, H):
    n = lib().dlo_spiral_blocks(W, H, None)
    out = np.zeros(2 * n, np.int32)
    lib().dlo_spiral_blocks(W, H, out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(n, 2)
