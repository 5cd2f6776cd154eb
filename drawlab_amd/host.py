"""ctypes binding of the C++ plugin layer (lib/libdrawlab_host.so,
include/drawlab_host.h): DrawLab scene XML files loaded through the
reference's plugin model and rendered by the HIP backend.

``load_xml(path)`` returns an XmlScene whose ``desc`` is the exported scene
as a drawlab_amd.scene.SceneDesc (arrays copied out of the C descriptors),
so a scene read from XML can be uploaded through the Python binding or
rebuilt in the CPU oracle for parity checks.
"""
import ctypes
import os

import numpy as np

from . import (LIGHT_AREA, LIGHT_POINT, LIB_PATH, DlError, dl_render_params, dl_render_stats, dl_scene_desc, lib)
from .scene import (BSDF_ANISO_GGX, BSDF_CONDUCTOR, BSDF_DIELECTRIC, BSDF_DIFFUSE, BSDF_MICROFACET, BSDF_MIRROR,
                    Bsdf, Camera, EnvMap, RFilter, SceneDesc)

HOST_LIB_PATH = os.path.join(os.path.dirname(LIB_PATH), "libdrawlab_host.so")
CLI_PATH = os.path.join(os.path.dirname(LIB_PATH), "drawlab")
HOST_SYMBOLS = ("dlx_load_xml", "dlx_free", "dlx_last_error", "dlx_export", "dlx_describe", "dlx_render",
                "dlx_save_image", "dlx_load_image")

# number of meaningful parameters per BSDF type (dl_hip.h DL_BSDF_*)
_NPARAMS = {BSDF_DIFFUSE: 3, BSDF_MIRROR: 0, BSDF_DIELECTRIC: 2, BSDF_MICROFACET: 6, BSDF_ANISO_GGX: 8,
            BSDF_CONDUCTOR: 7}

_hl = None


def host_lib():
    global _hl
    if _hl is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise DlError("libdrawlab_host.so is not built: run __graft_entry__.build() or make -C drawlab_amd")
        lib()  # resolves libdrawlab_hip.so first
        L = ctypes.CDLL(HOST_LIB_PATH)
        P, i32, c = ctypes.c_void_p, ctypes.c_int, ctypes
        sigs = {
            "dlx_load_xml": (i32, [c.c_char_p, c.POINTER(P)]),
            "dlx_free": (None, [P]),
            "dlx_last_error": (c.c_char_p, []),
            "dlx_export": (i32, [P, c.POINTER(c.POINTER(dl_scene_desc)), c.POINTER(dl_render_params)]),
            "dlx_describe": (i32, [P, c.c_char_p, c.c_size_t]),
            "dlx_render": (i32, [P, i32, i32, i32, P, c.POINTER(dl_render_stats)]),
            "dlx_save_image": (i32, [c.c_char_p, P, i32, i32]),
            "dlx_load_image": (i32, [c.c_char_p, c.POINTER(i32), c.POINTER(i32), P]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _hl = L
    return _hl


class XmlError(DlError):
    """A drawlab::Exception raised by the plugin layer (message as the
    reference formats it)."""


def _check(rc):
    if rc != 0:
        raise XmlError(host_lib().dlx_last_error().decode())


def _arr(ptr, count, dtype):
    if not ptr or count == 0:
        return None
    ct = ctypes.c_float if dtype == np.float32 else ctypes.c_uint32
    buf = (ct * count).from_address(ptr)
    return np.ctypeslib.as_array(buf).astype(dtype, copy=True)


def _to_scene_desc(sd, params):
    d = SceneDesc()
    for i in range(sd.num_textures):
        t = sd.textures[i]
        d.add_texture(_arr(t.rgb, 3 * t.width * t.height, np.float32).reshape(t.height, t.width, 3).copy())
    for i in range(sd.num_materials):
        m = sd.materials[i]
        b = Bsdf(int(m.type), tuple(float(x) for x in m.params[:_NPARAMS[int(m.type)]]))
        slots = {s: int(m.tex[s]) - 1 for s in range(len(m.tex)) if m.tex[s] > 0}  # tex[] holds 1 + index
        if slots:
            b.textures = slots
            b.maps_object_space = bool(m.maps_object_space)
        d.bsdfs.append(b)
    lights = [sd.lights[i] for i in range(sd.num_lights)]
    for i in range(sd.num_meshes):
        md = sd.meshes[i]
        V = _arr(md.positions, 3 * md.num_vertices, np.float32).reshape(-1, 3)
        F = _arr(md.indices, 3 * md.num_triangles, np.uint32).reshape(-1, 3)
        N = _arr(md.normals, 3 * md.num_vertices, np.float32)
        UV = _arr(md.texcoords, 2 * md.num_vertices, np.float32)
        rad = tuple(float(x) for x in lights[md.light].radiance) if md.light >= 0 else None
        d.add_mesh(V, F, int(md.material), None if N is None else N.reshape(-1, 3),
                   None if UV is None else UV.reshape(-1, 2), rad)
    order = []
    for ld in lights:
        if ld.type == LIGHT_AREA:
            order.append(("area", int(ld.mesh)))
        elif ld.type == LIGHT_POINT:
            order.append(("point", len(d.point_lights)))
            d.point_lights.append((tuple(float(x) for x in ld.position), tuple(float(x) for x in ld.radiance)))
        else:
            rgb = _arr(ld.env_rgb, 3 * ld.env_width * ld.env_height, np.float32).reshape(ld.env_height,
                                                                                         ld.env_width, 3)
            d.envmap = EnvMap(rgb, float(ld.env_scale), bool(ld.env_visual))
            order.append(("env", 0))
    d.emitters = order
    c = sd.camera
    d.camera = Camera(width=int(c.width), height=int(c.height))
    d.camera_desc = c  # the exported camera (basis / matrices), used as given
    f = sd.film
    d.film = f
    d.rfilter = RFilter(radius=float(f.radius))
    d.bg_color = tuple(float(x) for x in sd.bg_color)
    d.spp, d.seed = int(params.spp), int(params.seed)
    d.max_depth, d.rr_depth = int(params.max_depth), int(params.rr_depth)
    return d


class XmlScene:
    """A scene loaded from XML by the plugin layer."""

    def __init__(self, path):
        self.path = str(path)
        h = ctypes.c_void_p()
        _check(host_lib().dlx_load_xml(self.path.encode(), ctypes.byref(h)))
        self.h = h
        sdp = ctypes.POINTER(dl_scene_desc)()
        self.params = dl_render_params()
        _check(host_lib().dlx_export(self.h, ctypes.byref(sdp), ctypes.byref(self.params)))
        self.raw = sdp.contents
        self.desc = _to_scene_desc(self.raw, self.params)

    def describe(self):
        buf = ctypes.create_string_buffer(1 << 16)
        _check(host_lib().dlx_describe(self.h, buf, len(buf)))
        return buf.value.decode()

    def render(self, device=0, builder=0, spp=0):
        """HipRenderer path: a fresh context renders the exported scene."""
        W, H = self.raw.camera.width, self.raw.camera.height
        img = np.zeros((H, W, 3), np.float32)
        st = dl_render_stats()
        _check(host_lib().dlx_render(self.h, int(device), int(builder), int(spp),
                                     img.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
        return img, st.as_dict()

    def close(self):
        if self.h:
            host_lib().dlx_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_xml(path):
    return XmlScene(path)


def save_image(path, rgb):
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    _check(host_lib().dlx_save_image(str(path).encode(), rgb.ctypes.data_as(ctypes.c_void_p), rgb.shape[1],
                                     rgb.shape[0]))


def load_image(path):
    w, h = ctypes.c_int(), ctypes.c_int()
    _check(host_lib().dlx_load_image(str(path).encode(), ctypes.byref(w), ctypes.byref(h), None))
    out = np.zeros((h.value, w.value, 3), np.float32)
    _check(host_lib().dlx_load_image(str(path).encode(), ctypes.byref(w), ctypes.byref(h),
                                     out.ctypes.data_as(ctypes.c_void_p)))
    return out
