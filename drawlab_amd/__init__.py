"""drawlab_amd -- the MI355X (gfx950) rendering backend for DrawLab.

The product is the native library ``lib/libdrawlab_hip.so`` (hand-written HIP
kernels + the C ABI declared in ``include/dl_hip.h``) and its C++ host layer.
This package is a thin ctypes binding used by tests, benches and Python
callers.  There is no CPU fallback: without the built library (or without a
GPU) every entry point raises.
"""
import ctypes
import dataclasses
import os

import numpy as np

from .scene import (Bsdf, Camera, EnvMap, Mesh, RFilter, SceneDesc, BSDF_DIFFUSE, BSDF_MIRROR, BSDF_DIELECTRIC,
                    BSDF_MICROFACET, BSDF_ANISO_GGX, BSDF_CONDUCTOR, FILTER_GAUSSIAN, FILTER_BOX, FILTER_TENT,
                    FILTER_MITCHELL)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdrawlab_hip.so")
REPO_ROOT = os.path.dirname(_HERE)

BUILD_GPU, BUILD_HOST_SAH, BUILD_GPU_LBVH = 0, 1, 2
RENDER_GUIDED, RENDER_NORMALS = 4, 16
LIGHT_AREA, LIGHT_POINT, LIGHT_ENV = 0, 1, 2

EXPORTED_SYMBOLS = (
    "dl_hip_create", "dl_hip_destroy", "dl_hip_last_error", "dl_hip_device_name", "dl_hip_upload_scene",
    "dl_hip_build_accel", "dl_hip_trace_closest", "dl_hip_trace_shadow", "dl_hip_trace_closest_counted",
    "dl_hip_trace_closest_noalloc", "dl_hip_trace_shadow_counted",
    "dl_hip_render", "dl_hip_render_paths", "dl_hip_malloc", "dl_hip_free", "dl_hip_memcpy_h2d",
    "dl_hip_memcpy_d2h", "dl_hip_memcpy_d2d", "dl_hip_sync", "dl_hip_last_kernel_ms", "dl_camera_virtual", "dl_camera_perspective",
    "dl_camera_opencv", "dl_film_filter", "dl_hip_set_view_options", "dl_hip_render_batch",
    "dl_mesh_tangents", "dl_hip_memory_usage", "dl_hip_trim", "dl_hip_set_camera",
    "dl_hip_scene_keep", "dl_hip_scene_select", "dl_hip_scene_drop", "dl_hip_scene_bytes", "dl_hip_accel_export",
    "dl_hip_service_start", "dl_hip_service_add_scene", "dl_hip_service_submit", "dl_hip_service_poll",
    "dl_hip_service_wait", "dl_hip_service_result", "dl_hip_service_cancel", "dl_hip_service_stats",
    "dl_hip_service_stop", "dl_hip_service_last_error",
    "dl_hip_guide_init", "dl_hip_guide_params", "dl_hip_guide_eval", "dl_hip_guide_train_step",
)


class DlError(RuntimeError):
    pass


class dl_mesh_desc(ctypes.Structure):
    _fields_ = [("positions", ctypes.c_void_p), ("indices", ctypes.c_void_p), ("normals", ctypes.c_void_p),
                ("texcoords", ctypes.c_void_p), ("tangents", ctypes.c_void_p), ("num_vertices", ctypes.c_uint32),
                ("num_triangles", ctypes.c_uint32), ("material", ctypes.c_int32), ("light", ctypes.c_int32)]


class dl_material_desc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("params", ctypes.c_float * 16), ("tex", ctypes.c_int32 * 5),
                ("maps_object_space", ctypes.c_int32)]


class dl_texture_desc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("rgb", ctypes.c_void_p)]


class dl_light_desc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("mesh", ctypes.c_int32), ("radiance", ctypes.c_float * 3),
                ("position", ctypes.c_float * 3), ("env_rgb", ctypes.c_void_p), ("env_width", ctypes.c_int32),
                ("env_height", ctypes.c_int32), ("env_scale", ctypes.c_float), ("env_visual", ctypes.c_int32)]


class dl_camera_desc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("eye", ctypes.c_float * 3), ("U", ctypes.c_float * 3), ("V", ctypes.c_float * 3),
                ("W", ctypes.c_float * 3), ("sample_to_camera", ctypes.c_float * 16),
                ("camera_to_world", ctypes.c_float * 16), ("inv_size", ctypes.c_float * 2),
                ("near_clip", ctypes.c_float), ("far_clip", ctypes.c_float), ("intrinsics", ctypes.c_float * 4)]


class dl_batch_item(ctypes.Structure):
    _fields_ = [("camera", dl_camera_desc), ("spp", ctypes.c_int32), ("first_sample", ctypes.c_int32),
                ("seed", ctypes.c_uint32), ("film", ctypes.c_int32), ("accumulate", ctypes.c_int32),
                ("rgb_out", ctypes.c_void_p), ("scene", ctypes.c_uint64)]


BATCH_MAX = 32


class dl_view_options(ctypes.Structure):
    _fields_ = [("bg_color", ctypes.c_float * 3), ("epsilon", ctypes.c_float), ("hide_lights", ctypes.c_int32)]


class dl_film_desc(ctypes.Structure):
    _fields_ = [("radius", ctypes.c_float), ("border", ctypes.c_int32), ("lookup", ctypes.c_float),
                ("table", ctypes.c_float * 33)]


class dl_scene_desc(ctypes.Structure):
    _fields_ = [("meshes", ctypes.POINTER(dl_mesh_desc)), ("num_meshes", ctypes.c_uint32),
                ("materials", ctypes.POINTER(dl_material_desc)), ("num_materials", ctypes.c_uint32),
                ("lights", ctypes.POINTER(dl_light_desc)), ("num_lights", ctypes.c_uint32),
                ("camera", dl_camera_desc), ("film", dl_film_desc), ("bg_color", ctypes.c_float * 3),
                ("textures", ctypes.POINTER(dl_texture_desc)), ("num_textures", ctypes.c_uint32)]


class dl_accel_stats(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint32), ("num_leaves", ctypes.c_uint32), ("max_depth", ctypes.c_uint32),
                ("num_prims", ctypes.c_uint32), ("build_ms", ctypes.c_double), ("sah_cost", ctypes.c_double),
                ("node_bytes", ctypes.c_uint64), ("tri_bytes", ctypes.c_uint64)]


class dl_render_params(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_int32), ("seed", ctypes.c_uint32), ("max_depth", ctypes.c_int32),
                ("rr_depth", ctypes.c_int32), ("spp_per_pass", ctypes.c_int32), ("first_sample", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("guide_steps", ctypes.c_int32)]


class dl_render_stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "paths", "closest_rays", "shadow_rays", "invalid_samples", "ms_total", "ms_closest", "ms_shadow", "ms_shade",
        "ms_film", "ms_other", "launches_closest", "launches_shadow", "nodes_closest", "tris_closest", "ms_guide",
        "ms_train", "train_steps", "nodes_shadow", "tris_shadow")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def lib():
    """Load libdrawlab_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DlError("libdrawlab_hip.so is not built: run __graft_entry__.build() or make -C drawlab_amd")
        L = ctypes.CDLL(LIB_PATH)
        P, i32, u32, c = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes
        sigs = {
            "dl_hip_create": (i32, [i32, c.POINTER(P)]),
            "dl_hip_destroy": (i32, [P]),
            "dl_hip_last_error": (c.c_char_p, [P]),
            "dl_hip_device_name": (c.c_char_p, [P]),
            "dl_hip_upload_scene": (i32, [P, c.POINTER(dl_scene_desc)]),
            "dl_hip_build_accel": (i32, [P, i32, c.POINTER(dl_accel_stats)]),
            "dl_hip_trace_closest": (i32, [P, P, u32, P, P, P, P]),
            "dl_hip_trace_closest_counted": (i32, [P, P, u32, P, P, P, P, P, P]),
            "dl_hip_trace_closest_noalloc": (i32, [P, P, u32, P, P, P, P]),
            "dl_hip_trace_shadow": (i32, [P, P, u32, P]),
            "dl_hip_trace_shadow_counted": (i32, [P, P, u32, P, P, P]),
            "dl_hip_render": (i32, [P, c.POINTER(dl_render_params), P, c.POINTER(dl_render_stats)]),
            "dl_hip_render_paths": (i32, [P, c.POINTER(dl_render_params), u32, P]),
            "dl_hip_malloc": (i32, [P, c.c_size_t, c.POINTER(P)]),
            "dl_hip_free": (i32, [P, P]),
            "dl_hip_memcpy_h2d": (i32, [P, P, P, c.c_size_t]),
            "dl_hip_memcpy_d2h": (i32, [P, P, P, c.c_size_t]),
            "dl_hip_memcpy_d2d": (i32, [P, P, P, c.c_size_t]),
            "dl_hip_sync": (i32, [P]),
            "dl_hip_last_kernel_ms": (i32, [P, c.POINTER(c.c_double)]),
            "dl_camera_virtual": (None, [P, P, P, c.c_float, i32, i32, c.POINTER(dl_camera_desc)]),
            "dl_camera_perspective": (i32, [P, c.c_float, c.c_float, c.c_float, i32, i32, c.POINTER(dl_camera_desc)]),
            "dl_camera_opencv": (i32, [P, c.c_float, c.c_float, c.c_float, c.c_float, i32, i32,
                                       c.POINTER(dl_camera_desc)]),
            "dl_hip_set_view_options": (i32, [P, c.POINTER(dl_view_options)]),
            "dl_hip_set_camera": (i32, [P, c.POINTER(dl_camera_desc)]),
            "dl_hip_render_batch": (i32, [P, c.POINTER(dl_render_params), c.POINTER(dl_batch_item), u32,
                                          c.POINTER(dl_render_stats)]),
            "dl_film_filter": (None, [i32, c.c_float, c.c_float, c.c_float, c.POINTER(dl_film_desc)]),
            "dl_mesh_tangents": (None, [P, u32, P, u32, P, P]),
            "dl_hip_guide_init": (i32, [P, c.c_uint64, c.c_float, P]),
            "dl_hip_guide_params": (i32, [P, P, P]),
            "dl_hip_guide_eval": (i32, [P, P, u32, P, P]),
            "dl_hip_guide_train_step": (i32, [P, P, P, P, u32, i32, c.POINTER(c.c_float)]),
            "dl_hip_memory_usage": (i32, [P, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64)]),
            "dl_hip_trim": (i32, [P]),
            "dl_hip_scene_keep": (i32, [P, c.c_uint64]),
            "dl_hip_scene_select": (i32, [P, c.c_uint64]),
            "dl_hip_scene_drop": (i32, [P, c.c_uint64]),
            "dl_hip_scene_bytes": (i32, [P, c.c_uint64, c.POINTER(c.c_uint64)]),
            "dl_hip_accel_export": (i32, [P, P, P]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def camera_desc(cam: Camera) -> dl_camera_desc:
    out = dl_camera_desc()
    if cam.type == "perspective":
        tw = np.ascontiguousarray(np.eye(4, dtype=np.float32) if cam.to_world is None else cam.to_world,
                                  dtype=np.float32).reshape(16)
        st = lib().dl_camera_perspective(_p(tw), float(cam.fov), float(cam.near), float(cam.far), int(cam.width),
                                         int(cam.height), ctypes.byref(out))
        if st != 0:
            raise ValueError("perspective camera: projection is not invertible")
        return out
    if cam.type == "opencv":
        ext = np.ascontiguousarray(cam.extrinsic, dtype=np.float32).reshape(16)
        st = lib().dl_camera_opencv(_p(ext), float(cam.cx), float(cam.cy), float(cam.fx), float(cam.fy),
                                    int(cam.width), int(cam.height), ctypes.byref(out))
        if st != 0:
            raise ValueError("opencv camera: extrinsic matrix not invertible or zero focal length")
        return out
    e = np.asarray(cam.eye, np.float32)
    la = np.asarray(cam.lookat, np.float32)
    up = np.asarray(cam.up, np.float32)
    lib().dl_camera_virtual(_p(e), _p(la), _p(up), float(cam.fov), int(cam.width), int(cam.height), ctypes.byref(out))
    return out


def film_desc(f: RFilter) -> dl_film_desc:
    out = dl_film_desc()
    lib().dl_film_filter(int(f.type), float(f.radius), float(f.p0), float(f.p1), ctypes.byref(out))
    return out


class DeviceBuffer:
    """A device allocation owned by a Backend context."""

    def __init__(self, be, nbytes):
        self.be = be
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        be._check(lib().dl_hip_malloc(be.h, max(1, self.nbytes), ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.be._check(lib().dl_hip_memcpy_h2d(self.be.h, self.ptr, _p(arr), arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.be._check(lib().dl_hip_memcpy_d2h(self.be.h, _p(out), self.ptr, out.nbytes))
        return out

    def copy_from(self, other, nbytes=None):
        """Device-to-device copy (asynchronous on the context stream)."""
        n = min(self.nbytes, other.nbytes) if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= other.nbytes
        self.be._check(lib().dl_hip_memcpy_d2d(self.be.h, self.ptr, other.ptr, n))

    def free(self):
        if self.ptr is not None and self.be.h:
            lib().dl_hip_free(self.be.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def scene_descriptor(desc: SceneDesc):
    """dl_scene_desc of a SceneDesc plus the arrays it points into (keep
    them alive while the descriptor is used)."""
    keep = []
    order = desc.emitter_order()
    light_of_mesh = {}
    lights = (dl_light_desc * max(1, len(order)))()
    for li, (kind, idx) in enumerate(order):
        ld = lights[li]
        if kind == "area":
            ld.type = LIGHT_AREA
            ld.mesh = idx
            ld.radiance[:] = desc.meshes[idx].radiance
            light_of_mesh[idx] = li
        elif kind == "point":
            ld.type = LIGHT_POINT
            ld.mesh = -1
            ld.position[:] = desc.point_lights[idx][0]
            ld.radiance[:] = desc.point_lights[idx][1]
        else:
            env = desc.envmap
            rgb = np.ascontiguousarray(env.rgb, dtype=np.float32)
            keep.append(rgb)
            ld.type = LIGHT_ENV
            ld.mesh = -1
            ld.env_rgb = _p(rgb)
            ld.env_height, ld.env_width = rgb.shape[0], rgb.shape[1]
            ld.env_scale = float(env.scale)
            ld.env_visual = int(bool(env.visual))
    meshes = (dl_mesh_desc * len(desc.meshes))()
    for i, m in enumerate(desc.meshes):
        V = np.ascontiguousarray(m.V, dtype=np.float32)
        F = np.ascontiguousarray(m.F, dtype=np.uint32)
        N = None if m.N is None else np.ascontiguousarray(m.N, dtype=np.float32)
        UV = None if m.UV is None else np.ascontiguousarray(m.UV, dtype=np.float32)
        keep += [V, F, N, UV]
        md = meshes[i]
        md.positions, md.indices, md.normals, md.texcoords, md.tangents = _p(V), _p(F), _p(N), _p(UV), None
        md.num_vertices, md.num_triangles = V.shape[0], F.shape[0]
        md.material = m.bsdf
        md.light = light_of_mesh.get(i, -1)
    mats = (dl_material_desc * max(1, len(desc.bsdfs)))()
    for i, b in enumerate(desc.bsdfs):
        mats[i].type = b.type
        for k, v in enumerate(b.params):
            mats[i].params[k] = v
        for slot, t in (b.textures or {}).items():
            mats[i].tex[slot] = int(t) + 1
        mats[i].maps_object_space = int(bool(b.maps_object_space))
    texs = (dl_texture_desc * max(1, len(desc.textures)))()
    for i, rgb in enumerate(desc.textures):
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        keep.append(rgb)
        texs[i].height, texs[i].width = rgb.shape[0], rgb.shape[1]
        texs[i].rgb = _p(rgb)
    sd = dl_scene_desc()
    sd.meshes = meshes
    sd.num_meshes = len(desc.meshes)
    sd.materials = mats
    sd.num_materials = len(desc.bsdfs)
    sd.lights = lights
    sd.num_lights = len(order)
    # exported descriptors (scenes loaded through the plugin layer) are
    # used as given
    cd = getattr(desc, "camera_desc", None)
    sd.camera = cd if cd is not None else camera_desc(desc.camera)
    fd = getattr(desc, "film", None)
    sd.film = fd if fd is not None else film_desc(desc.rfilter)
    sd.bg_color[:] = desc.bg_color
    sd.textures = texs
    sd.num_textures = len(desc.textures)
    return sd, (keep, meshes, mats, lights, texs)


class Backend:
    """One HIP context on one device (the reference's DeviceContext role)."""

    def __init__(self, device=0):
        self.L = lib()
        h = ctypes.c_void_p()
        rc = self.L.dl_hip_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise DlError("dl_hip_create failed with status %d (no usable HIP device?)" % rc)
        self.h = h.value
        self._keep = []
        self.desc = None
        self.accel_stats = None

    def close(self):
        if self.h:
            self.L.dl_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = self.L.dl_hip_last_error(self.h)
            raise DlError("dl_hip status %d: %s" % (rc, msg.decode() if msg else ""))

    @property
    def device_name(self):
        return self.L.dl_hip_device_name(self.h).decode()

    # -- scene -------------------------------------------------------------
    def upload_scene(self, desc: SceneDesc):
        sd, keep = scene_descriptor(desc)
        self._check(self.L.dl_hip_upload_scene(self.h, ctypes.byref(sd)))
        self.desc = desc
        return self

    def build_accel(self, builder=BUILD_HOST_SAH):
        st = dl_accel_stats()
        self._check(self.L.dl_hip_build_accel(self.h, int(builder), ctypes.byref(st)))
        self.accel_stats = {n: getattr(st, n) for n, _ in st._fields_}
        return self.accel_stats

    # -- traversal ---------------------------------------------------------
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def trace_closest_dev(self, d_rays, n, d_prim, d_t, d_u, d_v, d_nodes=None, d_tris=None):
        if d_nodes is not None:
            self._check(self.L.dl_hip_trace_closest_counted(self.h, d_rays.ptr, n, d_prim.ptr, d_t.ptr, d_u.ptr,
                                                            d_v.ptr, d_nodes.ptr, d_tris.ptr))
        else:
            self._check(self.L.dl_hip_trace_closest(self.h, d_rays.ptr, n, d_prim.ptr, d_t.ptr, d_u.ptr, d_v.ptr))

    def trace_shadow_dev(self, d_rays, n, d_occ, d_nodes=None, d_tris=None):
        if d_nodes is not None:
            self._check(self.L.dl_hip_trace_shadow_counted(self.h, d_rays.ptr, n, d_occ.ptr, d_nodes.ptr, d_tris.ptr))
        else:
            self._check(self.L.dl_hip_trace_shadow(self.h, d_rays.ptr, n, d_occ.ptr))

    def sync(self):
        self._check(self.L.dl_hip_sync(self.h))

    def accel_export(self):
        """(nodes, triangles) the traversal reads, as float32 arrays of 32
        and 12 words per record (layouts in csrc/bvh_layout.h)."""
        st = self.accel_stats
        nodes = np.zeros(st["node_bytes"] // 4, np.float32)
        tris = np.zeros(st["tri_bytes"] // 4, np.float32)
        self._check(self.L.dl_hip_accel_export(self.h, _p(nodes), _p(tris)))
        return nodes, tris

    # ---- device memory and resident scenes --------------------------------
    def memory_usage(self):
        """(current, peak since the previous query) device bytes held."""
        cur, peak = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.dl_hip_memory_usage(self.h, ctypes.byref(cur), ctypes.byref(peak)))
        return cur.value, peak.value

    def trim(self):
        self._check(self.L.dl_hip_trim(self.h))
        self.desc = None
        self._kept = {}

    def scene_keep(self, key):
        self._check(self.L.dl_hip_scene_keep(self.h, int(key)))
        self._kept = getattr(self, "_kept", {})
        self._kept[int(key)] = self.desc

    def scene_select(self, key):
        """Make the scene kept under `key` current; False if none is kept."""
        st = self.L.dl_hip_scene_select(self.h, int(key))
        if st == 4:  # DL_ERR_STATE: no such resident scene
            return False
        self._check(st)
        self.desc = self._kept[int(key)]
        return True

    def scene_drop(self, key):
        self._check(self.L.dl_hip_scene_drop(self.h, int(key)))

    def scene_bytes(self, key=0):
        """Device bytes held by the scene kept under `key` (0: the current one)."""
        b = ctypes.c_uint64()
        self._check(self.L.dl_hip_scene_bytes(self.h, int(key), ctypes.byref(b)))
        return b.value

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        self._check(self.L.dl_hip_last_kernel_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def trace_closest(self, rays, counted=False):
        """Host convenience wrapper: rays is (8, n) float32 planes."""
        rays = np.ascontiguousarray(np.asarray(rays, dtype=np.float32).reshape(8, -1))
        n = rays.shape[1]
        d_rays = self.alloc(rays.nbytes)
        d_rays.upload(rays)
        bufs = [self.alloc(4 * max(1, n)) for _ in range(4)]
        cnt = [self.alloc(4 * max(1, n)) for _ in range(2)] if counted else [None, None]
        self.trace_closest_dev(d_rays, n, *bufs, *cnt)
        prim = bufs[0].download(np.uint32, n)
        t = bufs[1].download(np.float32, n)
        u = bufs[2].download(np.float32, n)
        v = bufs[3].download(np.float32, n)
        if counted:
            return prim, t, u, v, cnt[0].download(np.uint32, n), cnt[1].download(np.uint32, n)
        return prim, t, u, v

    def trace_shadow(self, rays, counted=False):
        rays = np.ascontiguousarray(np.asarray(rays, dtype=np.float32).reshape(8, -1))
        n = rays.shape[1]
        d_rays = self.alloc(rays.nbytes)
        d_rays.upload(rays)
        d_occ = self.alloc(max(1, n))
        cnt = [self.alloc(4 * max(1, n)) for _ in range(2)] if counted else [None, None]
        self.trace_shadow_dev(d_rays, n, d_occ, *cnt)
        occ = d_occ.download(np.uint8, n)
        if counted:
            return occ, cnt[0].download(np.uint32, n), cnt[1].download(np.uint32, n)
        return occ

    # -- rendering ---------------------------------------------------------
    def _params(self, spp, seed, max_depth, rr_depth, spp_per_pass=0, first_sample=0, flags=0):
        p = dl_render_params()
        p.spp, p.seed, p.max_depth, p.rr_depth = int(spp), int(seed), int(max_depth), int(rr_depth)
        p.spp_per_pass, p.first_sample, p.flags = int(spp_per_pass), int(first_sample), int(flags)
        return p

    def render(self, spp=None, seed=0, max_depth=None, rr_depth=None, spp_per_pass=0, want_image=True,
               count_traversal=False, first_sample=0, accumulate=False, guided=False, guide_frozen=False,
               guide_steps=0, normals=False):
        d = self.desc
        spp = d.spp if spp is None else spp
        max_depth = d.max_depth if max_depth is None else max_depth
        rr_depth = d.rr_depth if rr_depth is None else rr_depth
        flags = (1 if count_traversal else 0) | (2 if accumulate else 0) | (4 if guided else 0) | \
            (8 if guide_frozen else 0) | (RENDER_NORMALS if normals else 0)
        p = self._params(spp, seed, max_depth, rr_depth, spp_per_pass, first_sample, flags)
        p.guide_steps = int(guide_steps)
        img = np.zeros((d.camera.height, d.camera.width, 3), np.float32) if want_image else None
        st = dl_render_stats()
        self._check(self.L.dl_hip_render(self.h, ctypes.byref(p), _p(img), ctypes.byref(st)))
        return img, st.as_dict()

    def set_camera(self, cam):
        """Replaces the uploaded scene's camera and output size
        (dl_hip_set_camera); cam: Camera or dl_camera_desc."""
        cd = cam if isinstance(cam, dl_camera_desc) else camera_desc(cam)
        self._check(self.L.dl_hip_set_camera(self.h, ctypes.byref(cd)))
        if isinstance(cam, dl_camera_desc):
            cam = dataclasses.replace(self.desc.camera, width=int(cd.width), height=int(cd.height))
        self.desc = dataclasses.replace(self.desc, camera=cam)

    def render_batch(self, items, max_depth=None, rr_depth=None, normals=False, count_traversal=False):
        """One pass shared by several images (dl_hip_render_batch).
        items: dicts with camera (Camera or dl_camera_desc), spp,
        first_sample (0), seed (0), film (slot), accumulate (False), image
        (True: return the resolved image), scene (0: the current scene, or
        the key of a resident scene, see scene_keep()).  Returns (images, stats);
        images[i] is None when not asked for."""
        d = self.desc
        shared = self._params(1, 0, d.max_depth if max_depth is None else max_depth,
                              d.rr_depth if rr_depth is None else rr_depth,
                              flags=(RENDER_NORMALS if normals else 0) | (1 if count_traversal else 0))
        arr = (dl_batch_item * len(items))()
        imgs = []
        for i, it in enumerate(items):
            cam = it["camera"]
            cd = cam if isinstance(cam, dl_camera_desc) else camera_desc(cam)
            a = arr[i]
            a.camera = cd
            a.spp, a.first_sample, a.seed = int(it["spp"]), int(it.get("first_sample", 0)), int(it.get("seed", 0))
            a.film, a.accumulate = int(it["film"]), int(bool(it.get("accumulate", False)))
            a.scene = int(it.get("scene", 0))
            img = np.zeros((cd.height, cd.width, 3), np.float32) if it.get("image", True) else None
            a.rgb_out = None if img is None else img.ctypes.data
            imgs.append(img)
        st = dl_render_stats()
        self._check(self.L.dl_hip_render_batch(self.h, ctypes.byref(shared), arr, len(items), ctypes.byref(st)))
        return imgs, st.as_dict()

    def set_view_options(self, bg_color=None, epsilon=0.0, hide_lights=False):
        """Viewer settings of the current scene (dl_hip_set_view_options):
        background colour (None: the scene's), ray offset (0: 1e-4) and
        hidden light surfaces; reset_view_options() restores the defaults."""
        o = dl_view_options()
        bg = self.desc.bg_color if bg_color is None else bg_color
        for k in range(3):
            o.bg_color[k] = float(bg[k])
        o.epsilon, o.hide_lights = float(epsilon), int(bool(hide_lights))
        self._check(self.L.dl_hip_set_view_options(self.h, ctypes.byref(o)))

    def reset_view_options(self):
        self._check(self.L.dl_hip_set_view_options(self.h, None))

    # ---- learned directional proposal (dl_hip_guide_*) ------------------
    GUIDE_PARAMS = 21808

    def guide_init(self, seed=0, lr=1e-2, params=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        if p is not None:
            assert p.size == self.GUIDE_PARAMS
        self._check(self.L.dl_hip_guide_init(self.h, int(seed), float(lr), _p(p)))

    def guide_params(self):
        th = np.zeros(self.GUIDE_PARAMS, np.float32)
        g = np.zeros(self.GUIDE_PARAMS, np.float32)
        self._check(self.L.dl_hip_guide_params(self.h, _p(th), _p(g)))
        return th, g

    def guide_eval(self, x):
        """Raw outputs (n, 48) and mixtures (n, 40) for features x (n, 96)."""
        x = np.ascontiguousarray(x, np.float32)
        n = x.shape[0]
        bx, bo, bm = DeviceBuffer(self, x.nbytes), DeviceBuffer(self, n * 48 * 4), DeviceBuffer(self, n * 40 * 4)
        bx.upload(x)
        self._check(self.L.dl_hip_guide_eval(self.h, bx.ptr, n, bo.ptr, bm.ptr))
        return bo.download(np.float32, n * 48).reshape(n, 48), bm.download(np.float32, n * 40).reshape(n, 40)

    def guide_train_step(self, x, dirs, weights, apply=True):
        x = np.ascontiguousarray(x, np.float32)
        d = np.ascontiguousarray(dirs, np.float32)
        w = np.ascontiguousarray(weights, np.float32)
        n = x.shape[0]
        bx, bd, bw = DeviceBuffer(self, x.nbytes), DeviceBuffer(self, d.nbytes), DeviceBuffer(self, w.nbytes)
        bx.upload(x)
        bd.upload(d)
        bw.upload(w)
        loss = ctypes.c_float()
        self._check(self.L.dl_hip_guide_train_step(self.h, bx.ptr, bd.ptr, bw.ptr, n, int(bool(apply)),
                                                   ctypes.byref(loss)))
        return float(loss.value)

    def render_paths(self, sample, seed=0, max_depth=None, rr_depth=None, normals=False):
        d = self.desc
        max_depth = d.max_depth if max_depth is None else max_depth
        rr_depth = d.rr_depth if rr_depth is None else rr_depth
        p = self._params(sample + 1, seed, max_depth, rr_depth, flags=RENDER_NORMALS if normals else 0)
        out = np.zeros((d.camera.height * d.camera.width, 5), np.float32)
        self._check(self.L.dl_hip_render_paths(self.h, ctypes.byref(p), int(sample), _p(out)))
        return out


# ---------------------------------------------------------------------------
# render service (dl_hip_service_*)
# ---------------------------------------------------------------------------
POLICY_FIFO, POLICY_SJF = 0, 1
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)
SERVICE_SYMBOLS = ("dl_hip_service_start", "dl_hip_service_add_scene", "dl_hip_service_submit",
                   "dl_hip_service_poll", "dl_hip_service_wait", "dl_hip_service_result", "dl_hip_service_cancel",
                   "dl_hip_service_stats", "dl_hip_service_stop", "dl_hip_service_last_error",
                   "dl_hip_memory_usage")


class dl_service_config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("lanes", ctypes.c_int32), ("policy", ctypes.c_int32),
                ("builder", ctypes.c_int32), ("budget_bytes", ctypes.c_uint64), ("max_batch", ctypes.c_int32),
                ("single_scene_passes", ctypes.c_int32)]


class dl_job_info(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("lane", ctypes.c_int32), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_finish", ctypes.c_double), ("paths", ctypes.c_double),
                ("ms_gpu", ctypes.c_double), ("reserved_bytes", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class dl_service_stats(ctypes.Structure):
    _fields_ = [("jobs_done", ctypes.c_uint64), ("jobs_failed", ctypes.c_uint64),
                ("jobs_cancelled", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("peak_reserved_bytes", ctypes.c_uint64), ("estimate_misses", ctypes.c_uint64),
                ("scene_uploads", ctypes.c_uint64), ("scene_selects", ctypes.c_uint64), ("paths", ctypes.c_double),
                ("batched_jobs", ctypes.c_uint64), ("shared_passes", ctypes.c_uint64),
                ("cross_scene_passes", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _service_lib():
    L = lib()
    if not getattr(L, "_service_sigs", False):
        P, i32, u64, c = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes
        sigs = {
            "dl_hip_service_start": (i32, [c.POINTER(dl_service_config), c.POINTER(P)]),
            "dl_hip_service_add_scene": (i32, [P, c.POINTER(dl_scene_desc), c.POINTER(u64)]),
            "dl_hip_service_submit": (i32, [P, u64, c.POINTER(dl_camera_desc), c.POINTER(dl_render_params),
                                            c.POINTER(u64)]),
            "dl_hip_service_poll": (i32, [P, u64, c.POINTER(dl_job_info)]),
            "dl_hip_service_wait": (i32, [P, u64, c.POINTER(dl_job_info)]),
            "dl_hip_service_result": (i32, [P, u64, P]),
            "dl_hip_service_cancel": (i32, [P, u64]),
            "dl_hip_service_stats": (i32, [P, c.POINTER(dl_service_stats)]),
            "dl_hip_service_stop": (i32, [P]),
            "dl_hip_service_last_error": (c.c_char_p, [P]),
            "dl_hip_memory_usage": (i32, [P, c.POINTER(u64), c.POINTER(u64)]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._service_sigs = True
    return L


class Service:
    """Many render jobs on one device: scenes registered once, jobs queued
    and run on `lanes` worker contexts under a device-memory budget."""

    def __init__(self, device=0, lanes=4, policy=POLICY_FIFO, budget_bytes=200 << 30, builder=BUILD_GPU, max_batch=0,
                 single_scene_passes=False):
        self.L = _service_lib()
        cfg = dl_service_config(int(device), int(lanes), int(policy), int(builder), int(budget_bytes),
                                int(max_batch), int(bool(single_scene_passes)))
        h = ctypes.c_void_p()
        st = self.L.dl_hip_service_start(ctypes.byref(cfg), ctypes.byref(h))
        if st != 0:
            raise DlError("dl_hip_service_start: %s (status %d)" % (self.L.dl_hip_service_last_error(None).decode(), st))
        self.h = h
        self.sizes = {}
        self.job_sizes = {}

    def _check(self, st):
        if st != 0:
            raise DlError(self.L.dl_hip_service_last_error(self.h).decode())

    def add_scene(self, desc):
        """Register a SceneDesc (its arrays are copied); returns the scene id."""
        sd, keep = scene_descriptor(desc)
        sid = ctypes.c_uint64()
        self._check(self.L.dl_hip_service_add_scene(self.h, ctypes.byref(sd), ctypes.byref(sid)))
        del keep
        self.sizes[sid.value] = (sd.camera.height, sd.camera.width)
        return sid.value

    def submit(self, scene_id, spp, seed=0, max_depth=-1, rr_depth=4, spp_per_pass=0, first_sample=0, camera=None,
               guided=False):
        """camera: None (the scene's camera) or a drawlab_amd.scene.Camera /
        dl_camera_desc for this job (e.g. another resolution); guided: render
        with the learned proposal (the job renders alone on its lane)."""
        p = dl_render_params(int(spp), int(seed), int(max_depth), int(rr_depth), int(spp_per_pass),
                             int(first_sample), RENDER_GUIDED if guided else 0)
        jid = ctypes.c_uint64()
        cd = None
        if camera is not None:
            cd = camera if isinstance(camera, dl_camera_desc) else camera_desc(camera)
        self._check(self.L.dl_hip_service_submit(self.h, int(scene_id), ctypes.byref(cd) if cd is not None else None,
                                                 ctypes.byref(p), ctypes.byref(jid)))
        self.job_sizes[jid.value] = (cd.height, cd.width) if cd is not None else self.sizes[int(scene_id)]
        return jid.value

    def poll(self, job_id):
        info = dl_job_info()
        self._check(self.L.dl_hip_service_poll(self.h, int(job_id), ctypes.byref(info)))
        return info.as_dict()

    def wait(self, job_id):
        info = dl_job_info()
        self._check(self.L.dl_hip_service_wait(self.h, int(job_id), ctypes.byref(info)))
        return info.as_dict()

    def result(self, job_id, scene_id=None):
        H, W = self.job_sizes[int(job_id)]
        img = np.zeros((H, W, 3), np.float32)
        self._check(self.L.dl_hip_service_result(self.h, int(job_id), _p(img)))
        return img

    def cancel(self, job_id):
        self._check(self.L.dl_hip_service_cancel(self.h, int(job_id)))

    def last_error(self):
        return self.L.dl_hip_service_last_error(self.h).decode()

    def stats(self):
        st = dl_service_stats()
        self._check(self.L.dl_hip_service_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def stop(self):
        if self.h:
            self.L.dl_hip_service_stop(self.h)
            self.h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
