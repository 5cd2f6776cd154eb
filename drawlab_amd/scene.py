"""Plain scene description shared by the XML loader, the procedural scene
generators and the HIP backend's C ABI (include/dl_hip.h).

It mirrors what the reference's plugin objects export to a backend: meshes
with a material and an optional area emitter (src/tracer/scene/scene.cpp:36-81),
point and environment emitters, the virtual camera, the reconstruction
filter and the scene background colour.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

BSDF_DIFFUSE, BSDF_MIRROR, BSDF_DIELECTRIC, BSDF_MICROFACET, BSDF_ANISO_GGX, BSDF_CONDUCTOR = range(6)
FILTER_GAUSSIAN, FILTER_BOX, FILTER_TENT, FILTER_MITCHELL = range(4)


TEX_ALBEDO, TEX_PD, TEX_PS, TEX_AXAY, TEX_NORMAL, TEX_TANGENT = 0, 0, 1, 2, 3, 4


@dataclass
class Bsdf:
    type: int
    params: Tuple[float, ...] = ()
    # texture slot -> index into SceneDesc.textures (diffuse: TEX_ALBEDO;
    # aniso_ggx: TEX_PD, TEX_PS, TEX_AXAY, TEX_NORMAL, TEX_TANGENT)
    textures: Optional[dict] = None
    maps_object_space: bool = False  # aniso_ggx normal / tangent maps (default: tangent space)

    @staticmethod
    def diffuse(albedo=(0.5, 0.5, 0.5)):
        return Bsdf(BSDF_DIFFUSE, tuple(float(a) for a in albedo))

    @staticmethod
    def mirror():
        return Bsdf(BSDF_MIRROR, ())

    @staticmethod
    def dielectric(int_ior=1.5046, ext_ior=1.000277):
        return Bsdf(BSDF_DIELECTRIC, (float(int_ior), float(ext_ior)))

    @staticmethod
    def microfacet(alpha=0.1, int_ior=1.5046, ext_ior=1.000277, kd=(0.5, 0.5, 0.5)):
        return Bsdf(BSDF_MICROFACET, (float(alpha), float(int_ior), float(ext_ior)) + tuple(float(k) for k in kd))

    @staticmethod
    def aniso_ggx(pd=(0.5, 0.5, 0.5), ps=(0.5, 0.5, 0.5), axay=(0.5, 0.5)):
        return Bsdf(BSDF_ANISO_GGX, tuple(float(x) for x in pd) + tuple(float(x) for x in ps) +
                    (float(axay[0]), float(axay[1])))

    @staticmethod
    def conductor(alpha=0.1, eta=(0.2, 0.92, 1.1), k=(3.912, 2.452, 2.142)):
        return Bsdf(BSDF_CONDUCTOR, (float(alpha),) + tuple(float(x) for x in eta) + tuple(float(x) for x in k))


@dataclass
class Mesh:
    V: np.ndarray                      # (nv, 3) float32
    F: np.ndarray                      # (nf, 3) uint32
    bsdf: int = 0
    N: Optional[np.ndarray] = None     # (nv, 3) float32
    UV: Optional[np.ndarray] = None    # (nv, 2) float32
    radiance: Optional[Tuple[float, float, float]] = None
    name: str = ""


@dataclass
class EnvMap:
    rgb: np.ndarray                    # (h, w, 3) float32, row 0 = +y
    scale: float = 1.0
    visual: bool = True


@dataclass
class Camera:
    """"virtual" (VirtualCamera: eye/lookat/up, vertical fov),
    "perspective" (PerspectiveCamera: toWorld 4x4, horizontal fov, clip
    distances) or "opencv" (OpencvCamera: extrinsic 4x4, cx, cy, fx, fy)."""
    eye: Tuple[float, float, float] = (0.0, 0.0, 5.0)
    lookat: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    up: Tuple[float, float, float] = (0.0, 1.0, 0.0)
    fov: float = 30.0
    width: int = 1280
    height: int = 720
    type: str = "virtual"
    to_world: Optional[np.ndarray] = None  # (4, 4) float32, row-major
    near: float = 1e-4
    far: float = 1e4
    # "opencv" (OpencvCamera): world-to-camera extrinsic (4, 4) row-major,
    # principal point and focal lengths in pixels
    extrinsic: Optional[np.ndarray] = None
    cx: float = 0.0
    cy: float = 0.0
    fx: float = 1.0
    fy: float = 1.0


@dataclass
class RFilter:
    type: int = FILTER_GAUSSIAN
    radius: float = 2.0
    p0: float = 0.5                    # gaussian stddev / mitchell B
    p1: float = 0.0                    # mitchell C


@dataclass
class SceneDesc:
    meshes: List[Mesh] = field(default_factory=list)
    bsdfs: List[Bsdf] = field(default_factory=list)
    point_lights: List[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = field(default_factory=list)
    envmap: Optional[EnvMap] = None
    camera: Camera = field(default_factory=Camera)
    rfilter: RFilter = field(default_factory=RFilter)
    bg_color: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    spp: int = 1
    max_depth: int = -1
    rr_depth: int = 4
    name: str = ""
    seed: int = 0
    # Scene::m_emitters order as (kind, index) pairs when it is not the
    # default (area meshes, then point lights, then the environment map), e.g.
    # for an XML file that declares the environment map first.
    emitters: Optional[List[Tuple[str, int]]] = None
    # bitmap textures: (height, width, 3) float32, row 0 at v = 0
    textures: List[np.ndarray] = field(default_factory=list)

    def add_texture(self, rgb) -> int:
        self.textures.append(np.ascontiguousarray(rgb, dtype=np.float32).reshape(np.shape(rgb)[0], -1, 3))
        return len(self.textures) - 1

    def add_bsdf(self, b: Bsdf) -> int:
        self.bsdfs.append(b)
        return len(self.bsdfs) - 1

    def add_mesh(self, V, F, bsdf, N=None, UV=None, radiance=None, name="") -> int:
        V = np.ascontiguousarray(V, dtype=np.float32).reshape(-1, 3)
        F = np.ascontiguousarray(F, dtype=np.uint32).reshape(-1, 3)
        N = None if N is None else np.ascontiguousarray(N, dtype=np.float32).reshape(-1, 3)
        UV = None if UV is None else np.ascontiguousarray(UV, dtype=np.float32).reshape(-1, 2)
        self.meshes.append(Mesh(V, F, bsdf, N, UV, None if radiance is None else tuple(float(r) for r in radiance),
                                name))
        return len(self.meshes) - 1

    @property
    def num_triangles(self) -> int:
        return int(sum(m.F.shape[0] for m in self.meshes))

    def emitter_order(self):
        """Emitters in Scene::m_emitters order (scene.cpp:36-81): `emitters`
        when given, else mesh emitters in mesh order, then point lights, then
        the environment map, as the procedural generators add them."""
        if self.emitters is not None:
            return list(self.emitters)
        out = [("area", i) for i, m in enumerate(self.meshes) if m.radiance is not None]
        out += [("point", i) for i in range(len(self.point_lights))]
        if self.envmap is not None:
            out.append(("env", 0))
        return out
