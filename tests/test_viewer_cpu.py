"""The OptiX backend's extra features on the oracle and the host side, no
GPU: the "opencv" camera (include/optix/camera/opencv.h:24-47, constructor
src/tracer/camera/opencv.cpp:9-24), the "normals" integrator
(src/tracer/integrator/normal.cpp:9-19) and the viewer settings of the OptiX
renderer (hidden lights, hitgroup_pgs.cu:55-64; ray epsilon; background
colour; optix_renderer.cpp:84-102)."""
import os

import numpy as np
import pytest

import drawlab_amd
import oracle
from drawlab_amd import scenes
from drawlab_amd.scene import Bsdf, Camera, RFilter, SceneDesc, FILTER_BOX


def _cbox_opencv(w=48, h=36):
    d = scenes.cornell_box(w, h, spp=4, max_depth=5, rr_depth=3)
    d.camera = scenes.opencv_camera((0.278, 0.273, -0.8), (0.278, 0.273, 0.0), (0, 1, 0), 39.3, w, h)
    return d


# ---------------------------------------------------------------------------
# opencv camera
# ---------------------------------------------------------------------------
def test_opencv_descriptor_inverts_the_extrinsic():
    cam = _cbox_opencv().camera
    cd = drawlab_amd.camera_desc(cam)
    assert cd.type == 2 and (cd.width, cd.height) == (48, 36)
    inv = np.array(cd.camera_to_world[:], np.float64).reshape(4, 4)
    assert np.allclose(inv @ cam.extrinsic.astype(np.float64), np.eye(4), atol=1e-6)
    assert np.allclose(cd.eye[:], (0.278, 0.273, -0.8), atol=1e-6)
    assert list(cd.intrinsics[:]) == pytest.approx([cam.cx, cam.cy, cam.fx, cam.fy])
    bad = Camera(type="opencv", width=4, height=4, extrinsic=np.zeros((4, 4), np.float32), fx=1, fy=1)
    with pytest.raises(ValueError):
        drawlab_amd.camera_desc(bad)
    bad = Camera(type="opencv", width=4, height=4, extrinsic=np.eye(4, dtype=np.float32), fx=0, fy=1)
    with pytest.raises(ValueError):
        drawlab_amd.camera_desc(bad)


def test_opencv_rays_closed_form():
    """Pixel (cx, cy) looks along the optical axis; one focal length to the
    right (down) of it looks 45 degrees to the camera's right (down); x
    grows to the right and y downward in the image, as in OpenCV."""
    d = _cbox_opencv()
    cd = drawlab_amd.camera_desc(d.camera)
    d.camera_desc = cd
    osc = oracle.from_desc(d, build=False)
    c = d.camera
    pos = np.array([[c.cx, c.cy], [c.cx + c.fx, c.cy], [c.cx, c.cy + c.fy], [c.cx - c.fx, c.cy - c.fy]],
                   np.float32)
    r = osc.camera_rays(pos)
    o, dirs = r[0:3].T, r[3:6].T
    assert np.allclose(o, (0.278, 0.273, -0.8), atol=1e-6)
    s = 1 / np.sqrt(2.0)
    want = np.array([[0, 0, 1], [-s, 0, s], [0, -s, s], [1 / np.sqrt(3), 1 / np.sqrt(3), 1 / np.sqrt(3)]])
    assert np.allclose(dirs, want, atol=2e-6)
    assert np.all(r[6] == np.float32(1e-4)) and np.all(np.isinf(r[7]))
    # the oracle's own float64 inverse gives the same rays to rounding
    o2 = oracle.from_desc(_cbox_opencv(), build=False).camera_rays(pos)
    assert np.allclose(o2, r, atol=2e-6)


def test_opencv_camera_reproduces_the_virtual_camera():
    """With the virtual camera's eye, target and vertical field of view and
    the principal point one pixel above the centre (VirtualCamera maps film
    row y to H - y - 1, virtual.cpp:52-65), the opencv camera casts the
    virtual camera's rays up to rounding: renders agree to relMSE 1e-12."""
    d = scenes.cornell_box(32, 32, spp=4, max_depth=5, rr_depth=3)
    a, _ = oracle.from_desc(d).render(4, seed=3, max_depth=5, rr_depth=3)
    d.camera = scenes.opencv_camera((0.278, 0.273, -0.8), (0.278, 0.273, 0.0), (0, 1, 0), 39.3, 32, 32)
    d.camera.cy -= 1
    d.camera_desc = drawlab_amd.camera_desc(d.camera)
    b, _ = oracle.from_desc(d).render(4, seed=3, max_depth=5, rr_depth=3)
    a, b = a.astype(np.float64), b.astype(np.float64)
    assert np.mean((a - b) ** 2 / (a * a + 1e-2)) < 1e-12 and a.mean() > 0.01


# ---------------------------------------------------------------------------
# normals integrator
# ---------------------------------------------------------------------------
def _plane_scene(N=None):
    """A quad filling the view of a virtual camera, a point light (a
    renderer needs an emitter), box filter."""
    d = SceneDesc()
    b = d.add_bsdf(Bsdf.diffuse())
    V, F, _ = scenes.quad((-4, -4, 0), (8, 0, 0), (0, 8, 0))
    d.add_mesh(V, F, b, N=N)
    d.point_lights.append(((0, 0, 3), (1, 1, 1)))
    d.camera = Camera(eye=(0, 0, 5), lookat=(0, 0, 0), up=(0, 1, 0), fov=30, width=12, height=10)
    d.rfilter = RFilter(FILTER_BOX)
    return d


def test_normals_integrator_closed_form():
    osc = oracle.from_desc(_plane_scene())
    osc.set_options(normals=True)
    img, st = osc.render(2, seed=1)
    assert np.array_equal(img, np.broadcast_to(np.float32([0, 0, 1]), img.shape))
    assert st["closest"] == st["paths"] and st["shadow"] == 0     # one camera ray, nothing else
    # vertex normals: |b0 n0 + b1 n1 + b2 normalize(n2)| of the hit (the
    # record's unnormalised shading normal, octree.cpp:202-205)
    n = np.float32([[0.6, -0.8, 0.0], [0.6, -0.8, 0.0], [0.6, -0.8, 0.0], [0.6, -0.8, 0.0]])
    osc = oracle.from_desc(_plane_scene(N=n))
    osc.set_options(normals=True)
    xys = np.array([(x, y, 0) for y in range(10) for x in range(12)], np.int32)
    out = osc.render_samples(xys, seed=2)
    assert np.allclose(out[:, 2:], [0.6, 0.8, 0.0], atol=1e-6)
    # a miss is black: look away from the plane
    d = _plane_scene()
    d.camera = Camera(eye=(0, 0, 5), lookat=(0, 0, 10), up=(0, 1, 0), fov=30, width=12, height=10)
    osc = oracle.from_desc(d)
    osc.set_options(normals=True)
    img, _ = osc.render(1, seed=1)
    assert not img.any()


# ---------------------------------------------------------------------------
# hidden lights
# ---------------------------------------------------------------------------
def _floor_and_lights(l2_offset):
    """A diffuse floor lit by a low light L1 off to the left (facing down)
    and a second light L2 facing up, between the camera and the floor,
    `l2_offset` along x from there.  Shadow rays from the visible floor to
    L1 stay below L2 (hidden lights still block shadow rays); L2's back
    faces the floor, so it adds no light there."""
    d = SceneDesc()
    wb = d.add_bsdf(Bsdf.diffuse((0.7, 0.7, 0.7)))
    V, F, _ = scenes.quad((-2, 0, -2), (0, 0, 4), (4, 0, 0))                  # floor, +y
    d.add_mesh(V, F, wb)
    V, F, _ = scenes.quad((-1.8, 1.0, -0.3), (0.6, 0, 0), (0, 0, 0.6))       # L1, -y
    d.add_mesh(V, F, wb, radiance=(6.0, 6.0, 6.0))
    x0 = 0.9 + l2_offset
    V, F, _ = scenes.quad((x0, 2.0, -0.3), (0, 0, 0.6), (0.6, 0, 0))         # L2, +y (toward the camera)
    d.add_mesh(V, F, wb, radiance=(9.0, 1.0, 1.0))
    d.camera = Camera(eye=(1.2, 3.0, 0.0), lookat=(1.2, 0, 0), up=(0, 0, 1), fov=40, width=24, height=24)
    d.rfilter = RFilter(FILTER_BOX)
    return d


def test_hidden_lights_are_invisible_and_transparent():
    """With hidden lights, L2 between the camera and the floor neither shows
    nor stops path rays: the image equals the one with L2 moved far out of
    view, up to the rounding of hit points found from the pass-through
    point, at every path length.  Without hiding, L2 shows."""
    near = oracle.from_desc(_floor_and_lights(0.0))
    far = oracle.from_desc(_floor_and_lights(50.0))
    for md in (2, 4, 6):
        imgs = []
        for osc in (near, far):
            osc.set_options(hide_lights=True)
            img, _ = osc.render(16, seed=3, max_depth=md, rr_depth=3)
            imgs.append(img.astype(np.float64))
        assert np.abs(imgs[0] - imgs[1]).max() <= 1e-4 * max(1.0, imgs[1].max()), md
    near.set_options(hide_lights=False)
    shown, _ = near.render(16, seed=3, max_depth=4, rr_depth=3)
    assert shown.max() >= 8.0 and imgs[0].max() < 8.0            # L2's radiance seen only without hiding


def test_hidden_light_filling_the_view_shows_the_background():
    d = SceneDesc()
    b = d.add_bsdf(Bsdf.diffuse())
    V, F, _ = scenes.quad((-4, -4, 0), (8, 0, 0), (0, 8, 0))                 # emitter facing the camera
    d.add_mesh(V, F, b, radiance=(3.0, 2.0, 1.0))
    d.camera = Camera(eye=(0, 0, 5), lookat=(0, 0, 0), up=(0, 1, 0), fov=30, width=8, height=8)
    d.rfilter = RFilter(FILTER_BOX)
    osc = oracle.from_desc(d)
    img, _ = osc.render(2, seed=1)
    assert np.array_equal(img, np.broadcast_to(np.float32([3, 2, 1]), img.shape))
    osc.set_options(hide_lights=True)
    img, _ = osc.render(2, seed=1)
    assert not img.any()


# ---------------------------------------------------------------------------
# epsilon
# ---------------------------------------------------------------------------
def test_epsilon_is_the_ray_offset():
    d = _plane_scene()
    osc = oracle.from_desc(d)
    pos = np.float32([[6.0, 5.0]])
    assert osc.camera_rays(pos)[6, 0] == np.float32(1e-4)
    osc.set_options(eps=1e-3)
    assert osc.camera_rays(pos)[6, 0] == np.float32(1e-3)
    osc.set_options(eps=0.0)
    assert osc.camera_rays(pos)[6, 0] == np.float32(1e-4)


# ---------------------------------------------------------------------------
# XML: the opencv camera and the normals integrator through the plugin layer
# ---------------------------------------------------------------------------
XML = """<scene>
  <integrator type="{integ}"/>
  <camera type="opencv">
    <float name="cx" value="24"/><float name="cy" value="18"/>
    <float name="fx" value="40"/><float name="fy" value="42"/>
    <integer name="width" value="48"/><integer name="height" value="36"/>
    <transform name="extrinsic">
      <matrix value="-1 0 0 0.5  0 -1 0 0.25  0 0 1 4  0 0 0 1"/>
    </transform>
  </camera>
  <mesh type="rectangle"><vector name="pos" value="-1, -1, 0"/><vector name="v0" value="2, 0, 0"/>
    <vector name="v1" value="0, 2, 0"/><emitter type="area"><color name="radiance" value="1, 1, 1"/></emitter></mesh>
</scene>"""


def test_xml_opencv_camera_and_normals_integrator(tmp_path):
    from drawlab_amd import host
    if not os.path.exists(host.HOST_LIB_PATH):
        pytest.skip("libdrawlab_host.so not built")
    p = tmp_path / "o.xml"
    p.write_text(XML.format(integ="normals"))
    xs = host.load_xml(str(p))
    cd = xs.raw.camera
    assert cd.type == 2 and (cd.width, cd.height) == (48, 36)
    assert list(cd.intrinsics[:]) == [24, 18, 40, 42]
    inv = np.array(cd.camera_to_world[:], np.float64).reshape(4, 4)
    ext = np.array([[-1, 0, 0, 0.5], [0, -1, 0, 0.25], [0, 0, 1, 4], [0, 0, 0, 1]], np.float64)
    assert np.allclose(inv @ ext, np.eye(4), atol=1e-6)
    assert np.allclose(cd.eye[:], (0.5, 0.25, -4.0), atol=1e-6)     # R^T (-t)
    assert xs.params.flags == drawlab_amd.RENDER_NORMALS
    assert "OpencvCamera" in xs.describe() and "NormalIntegrator" in xs.describe()
    p.write_text(XML.format(integ="path"))
    assert host.load_xml(str(p)).params.flags == 0
    p.write_text(XML.format(integ="whitted"))
    with pytest.raises(host.XmlError, match="only 'path' and 'normals'"):
        host.load_xml(str(p))
    # cx, cy, fx, fy and the extrinsic transform are required (opencv.cpp:11-16)
    p.write_text(XML.format(integ="path").replace('<float name="fx" value="40"/>', ""))
    with pytest.raises(host.XmlError, match="fx"):
        host.load_xml(str(p))
