"""The C++ plugin layer (libdrawlab_host.so): XML loading with the
reference's tags, properties, defaults and error messages
(src/core/parser/parser.cpp, proplist.cpp, the plugins under src/tracer/),
export to the C ABI descriptors, OBJ / rectangle meshes, transforms, image
I/O and the command-line front end.  No GPU is used."""
import math
import os
import struct
import subprocess
import textwrap

import numpy as np
import pytest

import drawlab_amd
from drawlab_amd import host
from drawlab_amd.scene import Camera, RFilter

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(HERE, "golden", "scenes")

pytestmark = pytest.mark.skipif(not os.path.exists(host.HOST_LIB_PATH), reason="libdrawlab_host.so not built")

MINIMAL = """<scene>
  <integrator type="path"/>
  <camera type="virtual">
    <vector name="eye" value="0, 0, 5"/><vector name="lookat" value="0, 0, 0"/><vector name="up" value="0, 1, 0"/>
    <integer name="width" value="8"/><integer name="height" value="6"/>
  </camera>
  {body}
</scene>"""

LIGHT = """<mesh type="rectangle"><vector name="pos" value="-1, -1, 0"/><vector name="v0" value="2, 0, 0"/>
  <vector name="v1" value="0, 2, 0"/><emitter type="area"><color name="radiance" value="1, 1, 1"/></emitter></mesh>"""


def write(tmp_path, body, name="s.xml", template=MINIMAL):
    p = tmp_path / name
    p.write_text(template.format(body=body) if template else body)
    return str(p)


def load(tmp_path, body, **kw):
    return host.load_xml(write(tmp_path, body, **kw))


# ---------------------------------------------------------------------------
# fixtures from the documentation
# ---------------------------------------------------------------------------
def test_docs_examples_load():
    s = host.load_xml(os.path.join(SCENES, "docs_examples.xml"))
    d = s.desc
    # camera of the example: VirtualCamera basis as dl_camera_virtual computes it
    ref = drawlab_amd.camera_desc(Camera(eye=(0, 0.919769, 5.41159), lookat=(0, 0.893051, 4.41198), up=(0, 1, 0),
                                         fov=27.7856, width=600, height=600))
    for f in ("eye", "U", "V", "W"):
        assert list(getattr(s.raw.camera, f)) == list(getattr(ref, f))
    # obj mesh: scale (0.06, 0.06, -1) then translate (10, 0, 25)
    cube = d.meshes[0].V
    assert np.allclose(cube.min(0), [10.0, 0.0, 24.0], atol=1e-6)
    assert np.allclose(cube.max(0), [10.06, 0.06, 25.0], atol=1e-6)
    # emitters in document order: the rectangle's area light, the point light, the envmap
    assert d.emitters == [("area", 1), ("point", 0), ("env", 0)]
    assert d.point_lights == [((0.0, 0.0, 0.0), (3.0, 3.0, 2.5))]
    assert d.envmap.scale == 0.5 and d.envmap.visual is False
    # materials: defaults of the reference constructors
    types = [b.type for b in d.bsdfs]
    assert types == [0, 0, 2, 1, 3, 4, 0]
    assert d.bsdfs[1].params == (0.5, 0.5, 0.5)                        # Mesh::activate default diffuse
    assert d.bsdfs[2].params == pytest.approx((1.5046, 1.000277))       # dielectric defaults
    assert d.bsdfs[4].params == pytest.approx((0.28, 1.7, 1.000277, 0.2, 0.2, 0.4))
    assert d.bsdfs[5].params == pytest.approx((0, 0, 0, 0.5, 0.5, 0.5, 0.2, 0.2))
    assert d.spp == 1 and d.max_depth == -1 and d.rr_depth == 4
    assert "PathIntegrator" in s.describe()


def test_cbox_fixture_geometry():
    s = host.load_xml(os.path.join(SCENES, "cbox.xml"))
    d = s.desc
    assert [m.F.shape[0] for m in d.meshes] == [2, 2, 2, 2, 2, 12, 12, 2]
    # the shared <ref id="white"/> material is one material; inline ones follow
    assert len(d.bsdfs) == 3
    assert [m.bsdf for m in d.meshes] == [0, 0, 0, 1, 2, 0, 0, 0]
    # boxes: unit cube scaled then translated (toWorld)
    assert np.allclose(d.meshes[5].V.min(0), [0.130, 0, 0.065], atol=1e-6)
    assert np.allclose(d.meshes[5].V.max(0), [0.295, 0.165, 0.230], atol=1e-6)
    assert np.allclose(d.meshes[6].V.max(0), [0.430, 0.330, 0.460], atol=1e-6)
    # the light faces down
    light = d.meshes[7]
    assert light.radiance == (17.0, 12.0, 4.0)
    assert np.allclose(light.N, [[0, -1, 0]] * 4)
    assert (d.spp, d.max_depth, d.rr_depth) == (16, 8, 3)


def test_rectangle_matches_reference_layout(tmp_path):
    s = load(tmp_path, """<mesh type="rectangle"><vector name="pos" value="1, 2, 3"/>
      <vector name="v0" value="0, 0, 2"/><vector name="v1" value="4, 0, 0"/>
      <emitter type="area"><color name="radiance" value="1, 2, 3"/></emitter></mesh>""")
    m = s.desc.meshes[0]
    assert m.V.tolist() == [[1, 2, 3], [1, 2, 5], [5, 2, 3], [5, 2, 5]]
    assert m.F.tolist() == [[0, 1, 2], [3, 2, 1]]
    assert m.UV.tolist() == [[1, 0], [1, 1], [0, 0], [0, 1]]
    assert np.allclose(m.N, [[0, 1, 0]] * 4)        # normalize(v0 x v1)


def test_obj_loader_deindexes_and_fans(tmp_path):
    (tmp_path / "q.obj").write_text(textwrap.dedent("""\
        v 0 0 0
        v 1 0 0
        v 1 1 0
        v 0 1 0
        vt 0 0
        vt 1 0
        vt 1 1
        vt 0 1
        vn 0 0 1
        f 1/1/1 2/2/1 3/3/1 4/4/1
        f -4/-4/-1 -2/-2/-1 -1/-1/-1
        """))
    s = load(tmp_path, '<mesh type="obj"><string name="filename" value="q.obj"/></mesh>' + LIGHT)
    m = s.desc.meshes[0]
    assert m.F.tolist() == [[0, 1, 2], [0, 2, 3], [0, 2, 3]]   # fan + negative indices resolve to the same corners
    assert m.V.shape == (4, 3) and m.UV.tolist() == [[0, 0], [1, 0], [1, 1], [0, 1]]
    assert np.allclose(m.N, [[0, 0, 1]] * 4)


def test_transform_operations_compose_on_the_left(tmp_path):
    (tmp_path / "p.obj").write_text("v 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\n")
    body = """<mesh type="obj"><string name="filename" value="p.obj"/>
      <transform name="toWorld">
        <scale value="2"/>
        <rotate axis="0, 0, 1" angle="90"/>
        <translate value="1, 2, 3"/>
      </transform></mesh>""" + LIGHT
    V = load(tmp_path, body).desc.meshes[0].V
    # p' = T(1,2,3) R_z(90) S(2) p
    want = np.array([[0, 2, 0], [-2, 0, 0], [0, 0, 2]], np.float64) + [1, 2, 3]
    assert np.allclose(V, want, atol=1e-6)
    body = """<mesh type="obj"><string name="filename" value="p.obj"/>
      <transform name="toWorld"><matrix value="1 0 0 5  0 1 0 6  0 0 1 7  0 0 0 1"/></transform></mesh>""" + LIGHT
    V = load(tmp_path, body).desc.meshes[0].V
    assert np.allclose(V, np.eye(3) + [5, 6, 7])


def test_lookat_transform_and_perspective_camera(tmp_path):
    body = """<camera type="perspective"><float name="fov" value="40"/>
      <integer name="width" value="64"/><integer name="height" value="48"/>
      <transform name="toWorld"><lookat origin="1, 2, 3" target="1, 2, 4" up="0, 1, 0"/></transform>
    </camera>""" + LIGHT
    tpl = "<scene><integrator type=\"path\"/>{body}</scene>"
    s = load(tmp_path, body, template=tpl)
    c = s.raw.camera
    assert c.type == 1 and c.width == 64 and c.height == 48
    C = np.array(c.camera_to_world[:]).reshape(4, 4)
    # lookat along +z from (1, 2, 3): columns left, up, dir, origin
    assert np.allclose(C, [[1, 0, 0, 1], [0, 1, 0, 2], [0, 0, 1, 3], [0, 0, 0, 1]], atol=1e-7)
    ref = drawlab_amd.camera_desc(Camera(type="perspective", fov=40, width=64, height=48, to_world=C))
    assert list(c.sample_to_camera) == list(ref.sample_to_camera)


def test_perspective_matrices_closed_form():
    """sampleToCamera maps the image centre to the optical axis and the
    corners to +-tan(fov/2) horizontally, +-tan(fov/2)/aspect vertically;
    the product's set-up agrees with the oracle's restatement."""
    import oracle
    W, H, fov = 64, 48, 40.0
    c = drawlab_amd.camera_desc(Camera(type="perspective", fov=fov, width=W, height=H))
    S = np.array(c.sample_to_camera[:], np.float64).reshape(4, 4)

    def ray(px, py):
        q = S @ np.array([px / W, py / H, 0.0, 1.0])
        p = q[:3] / q[3]
        return p / p[2]
    t = math.tan(math.radians(fov / 2))
    assert np.allclose(ray(W / 2, H / 2), [0, 0, 1], atol=1e-6)
    assert np.allclose(ray(0, 0), [t, t * H / W, 1], rtol=1e-5)
    assert np.allclose(ray(W, H), [-t, -t * H / W, 1], rtol=1e-5)
    So, Co = oracle.perspective_matrices(fov, 1e-4, 1e4, W, H)
    assert np.allclose(So, S.reshape(16), rtol=1e-6, atol=1e-9)
    assert (c.near_clip, c.far_clip) == (np.float32(1e-4), np.float32(1e4))


def test_filters_export(tmp_path):
    cases = [("<rfilter type=\"gaussian\"/>", RFilter(0, 2.0, 0.5)),
             ("<rfilter type=\"gaussian\"><float name=\"radius\" value=\"3\"/><float name=\"stddev\" value=\"0.7\"/></rfilter>",
              RFilter(0, 3.0, 0.7)),
             ("<rfilter type=\"box\"/>", RFilter(1, 0.5)),
             ("<rfilter type=\"tent\"/>", RFilter(2, 1.0)),
             ("<rfilter type=\"mitchell\"/>", RFilter(3, 2.0, 1 / 3, 1 / 3)),
             ("", RFilter(0, 2.0, 0.5))]   # no filter: the default Gaussian
    for xml, rf in cases:
        tpl = MINIMAL.replace("</camera>", xml + "</camera>")
        s = load(tmp_path, LIGHT, template=tpl)
        ref = drawlab_amd.film_desc(rf)
        assert s.raw.film.radius == ref.radius and s.raw.film.border == ref.border
        assert list(s.raw.film.table) == list(ref.table), xml


def test_shared_material_and_emitter_order(tmp_path):
    body = """<emitter type="envmap"><string name="filename" value="%s"/></emitter>
      <bsdf type="mirror" id="m"/>
      <mesh type="rectangle"><vector name="pos" value="0,0,0"/><vector name="v0" value="1,0,0"/>
        <vector name="v1" value="0,1,0"/><ref id="m"/></mesh>
      %s
      <emitter type="point"><color name="intensity" value="1,1,1"/><point name="position" value="0,1,0"/></emitter>
      <mesh type="rectangle"><vector name="pos" value="0,0,1"/><vector name="v0" value="1,0,0"/>
        <vector name="v1" value="0,1,0"/><ref id="m"/></mesh>""" % (os.path.join(SCENES, "envmap.pfm"), LIGHT)
    d = load(tmp_path, body).desc
    assert d.emitters == [("env", 0), ("area", 1), ("point", 0)]
    assert [m.bsdf for m in d.meshes] == [0, 1, 0] and [b.type for b in d.bsdfs] == [1, 0]


def test_render_params_and_extensions(tmp_path):
    tpl = MINIMAL.replace('<integrator type="path"/>', """<integrator type="path">
      <integer name="maxDepth" value="5"/><integer name="rrDepth" value="2"/></integrator>
      <sampler type="independent"><integer name="sampleCount" value="7"/><integer name="seed" value="9"/></sampler>""")
    s = load(tmp_path, LIGHT + """<mesh type="rectangle"><vector name="pos" value="0,0,0"/><vector name="v0" value="1,0,0"/>
      <vector name="v1" value="0,1,0"/><bsdf type="conductor"><float name="alpha" value="0.2"/></bsdf></mesh>""",
             template=tpl)
    p = s.params
    assert (p.spp, p.seed, p.max_depth, p.rr_depth) == (7, 9, 5, 2)
    assert s.desc.bsdfs[1].params == pytest.approx((0.2, 0.2, 0.92, 1.1, 3.912, 2.452, 2.142))


ERRORS = [
    # (body or full document, expected message fragment, full document?)
    ("<foo/>", 'unexpected tag "foo"', False),
    ("<mesh type=\"nosuch\"/>", 'A constructor for class "nosuch" could not be found!', False),
    ("<mesh type=\"rectangle\"><vector name=\"pos\" value=\"0,0,0\"/><vector name=\"v0\" value=\"1,0,0\"/></mesh>",
     "Property 'v1' is missing!", False),
    ("<mesh type=\"rectangle\"><float name=\"pos\" value=\"1\"/><vector name=\"v0\" value=\"1,0,0\"/>"
     "<vector name=\"v1\" value=\"0,1,0\"/></mesh>", "Property 'pos' has the wrong type!", False),
    ("<mesh type=\"rectangle\"><ref id=\"missing\"/></mesh>", "Can't find the referred node missing", False),
    ("<bsdf type=\"diffuse\" id=\"a\"/><bsdf type=\"diffuse\" id=\"a\"/>", "the node id has been used", False),
    ("<camera type=\"virtual\"><vector name=\"eye\" value=\"0,0,1\"/><vector name=\"lookat\" value=\"0,0,0\"/>"
     "<vector name=\"up\" value=\"0,1,0\"/></camera>", "There can only be one camera per scene!", False),
    ("<integer name=\"x\" value=\"1.5\"/>", 'Could not parse integer value "1.5"', False),
    ("<boolean name=\"x\" value=\"yes\"/>", 'Could not parse boolean value "yes"', False),
    ("<vector name=\"x\" value=\"1, 2\"/>", "Expected 3 values", False),
    ("<mesh type=\"rectangle\" bad=\"1\"/>", 'unexpected attribute "bad" in "mesh"', False),
    ("<translate value=\"1,2,3\"/>", "transform nodes can only contain transform operations", False),
    # where each kind of element may appear
    ("<mesh type=\"rectangle\"><transform name=\"toWorld\"><float name=\"x\" value=\"1\"/></transform></mesh>",
     "transform nodes can only contain transform operations", False),
    ("<mesh type=\"rectangle\"><float name=\"x\" value=\"1\"><bsdf type=\"diffuse\"/></float></mesh>",
     'node "bsdf" requires a object as parent', False),
    ("<mesh type=\"rectangle\"><float name=\"x\" value=\"1\"><translate value=\"1,2,3\"/></float></mesh>",
     "transform nodes can only contain transform operations", False),
    ("<mesh type=\"rectangle\"><transform name=\"toWorld\"><translate value=\"1,2,3\"><float name=\"x\" "
     "value=\"1\"/></translate></transform></mesh>", 'node "float" requires a object as parent', False),
    ("<mesh type=\"rectangle\"><float name=\"x\"/></mesh>", 'missing attribute "value" in "float"', False),
    ("<transform name=\"t\"><translate value=\"1,2,3\"/></transform>", "must be a object", True),
    ("<mesh type=\"rectangle\"><transform name=\"toWorld\"><matrix value=\"1 0 0\"/></transform></mesh>",
     "Expected 16 values", False),
    ("<float name=\"x\" value=\"1\"/>", None, True),   # property as root
    ("<scene><camera type=\"virtual\"><vector name=\"eye\" value=\"0,0,1\"/><vector name=\"lookat\" value=\"0,0,0\"/>"
     "<vector name=\"up\" value=\"0,1,0\"/></camera></scene>", "No integrator was specified!", True),
    ("<scene><integrator type=\"path\"/></scene>", "No camera was specified!", True),
    ("<scene><integrator type=\"path\"/>", "unterminated", True),
]


@pytest.mark.parametrize("body,msg,full", ERRORS)
def test_parse_errors(tmp_path, body, msg, full):
    path = write(tmp_path, body, template=None if full else MINIMAL)
    with pytest.raises(host.XmlError) as e:
        host.load_xml(path)
    text = str(e.value)
    assert text.startswith('Error while parsing "%s"' % path)
    if msg:
        assert msg in text
    if body == "<float name=\"x\" value=\"1\"/>":
        assert "must be a object" in text
    assert "line" in text


def test_two_envmaps_rejected(tmp_path):
    env = '<emitter type="envmap"><string name="filename" value="%s"/></emitter>' % os.path.join(SCENES, "envmap.pfm")
    with pytest.raises(host.XmlError, match="There is more than one envmap in the scene!"):
        load(tmp_path, env + env + LIGHT)


def test_export_errors(tmp_path):
    with pytest.raises(host.XmlError, match="There is not emitter in the scene."):
        load(tmp_path, "")
    tpl = MINIMAL.replace('<integrator type="path"/>', '<integrator type="whitted"/>')
    with pytest.raises(host.XmlError, match="not supported by the hip backend"):
        load(tmp_path, LIGHT, template=tpl)
    with pytest.raises(host.XmlError, match="Unsupported texture format"):
        load(tmp_path, LIGHT + """<mesh type="rectangle"><vector name="pos" value="0,0,0"/><vector name="v0" value="1,0,0"/>
          <vector name="v1" value="0,1,0"/><bsdf type="diffuse"><texture type="bitmap" name="albedo">
          <string name="filename" value="wood.jpg"/></texture></bsdf></mesh>""")


# ---------------------------------------------------------------------------
# images
# ---------------------------------------------------------------------------
def test_image_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.random((5, 7, 3), dtype=np.float32) * 100
    for ext in ("exr", "pfm"):
        p = tmp_path / ("a." + ext)
        host.save_image(p, img)
        assert np.array_equal(host.load_image(p), img), ext


def _rgbe(r, g, b):
    m = max(r, g, b)
    if m < 1e-32:
        return (0, 0, 0, 0)
    e = math.frexp(m)[1]
    f = 256.0 / 2 ** e
    return (int(r * f), int(g * f), int(b * f), e + 128)


def test_hdr_reader_flat_and_rle(tmp_path):
    """Radiance RGBE: value = mantissa * 2^(e - 136); rows top-down."""
    W, H = 9, 2
    px = [[(1.0, 0.5, 0.25), (2.0, 0.0, 0.0), (0.0, 0.0, 0.0)] * 3, [(0.125, 4.0, 8.0)] * 9]
    flat = b"".join(bytes(_rgbe(*c)) for row in px for c in row)
    hdr = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (H, W)
    (tmp_path / "flat.hdr").write_bytes(hdr + flat)
    # adaptive RLE: each row starts with 2 2 hi lo; per channel runs (128+n, v) or literals (n, v...)
    rle = b""
    for row in px:
        q = [_rgbe(*c) for c in row]
        rle += bytes([2, 2, W >> 8, W & 255])
        for ch in range(4):
            vals = [v[ch] for v in q]
            rle += bytes([W]) + bytes(vals)          # one literal run of W values
    (tmp_path / "rle.hdr").write_bytes(hdr + rle)
    want = np.array(px, np.float32)
    for name in ("flat.hdr", "rle.hdr"):
        got = host.load_image(tmp_path / name)
        assert got.shape == (H, W, 3)
        assert np.array_equal(got, want), name


def test_envmap_orientation(tmp_path):
    """Row 0 of the lat-long table is +y: the top row of .hdr / .pfm files,
    the bottom row of OpenEXR files (the reference flips float bitmaps on
    texture upload, src/tracer/texture/bitmap.cpp:86, and only .hdr was
    already flipped on load, src/core/bitmap/bitmap.cpp:109)."""
    img = host.load_image(os.path.join(SCENES, "envmap.pfm"))
    assert img[:, 0, 0].tolist() == [1, 2, 3, 4]
    host.save_image(tmp_path / "env.exr", img)
    body = '<emitter type="envmap"><string name="filename" value="%s"/></emitter>' + LIGHT
    pfm = load(tmp_path, body % os.path.join(SCENES, "envmap.pfm")).desc.envmap.rgb
    exr = load(tmp_path, body % str(tmp_path / "env.exr")).desc.envmap.rgb
    assert pfm[:, 0, 0].tolist() == [1, 2, 3, 4]
    assert exr[:, 0, 0].tolist() == [4, 3, 2, 1]


# ---------------------------------------------------------------------------
# command line
# ---------------------------------------------------------------------------
def run_cli(*args):
    return subprocess.run([host.CLI_PATH, *args], capture_output=True, text=True, timeout=60,
                          env=dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", "")))


def test_cli_usage_and_errors(tmp_path):
    r = run_cli("--help")
    assert r.returncode == 0 and "--scene" in r.stdout and "--backend" in r.stdout
    r = run_cli("-s", os.path.join(SCENES, "cbox.xml"), "-b", "cpu")
    assert r.returncode != 0 and "unknown backend" in r.stderr
    r = run_cli("-s", str(tmp_path / "scene.json"))
    assert r.returncode != 0 and "expected an extension of type .xml" in r.stderr
    bad = write(tmp_path, "<foo/>")
    r = run_cli("-s", bad)
    assert r.returncode != 0 and 'unexpected tag "foo"' in r.stderr


def test_png_output_is_srgb_8bit(tmp_path):
    """Bitmap::savePNG (src/core/bitmap/bitmap.cpp:190-214): toSRGB
    (color.h:76-86), x255, clamped, truncated; decoded by an independent
    reader.  powf may differ from numpy's float32 power in the last bit, so
    a value may land one step off where 255 * srgb is within rounding of an
    integer."""
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(4)
    img = (rng.random((9, 13, 3)) * 1.3).astype(np.float32)
    img[0, 0] = (-1.0, 0.001, 0.0031308)
    img[0, 1] = (np.inf, 0.0, 5.0)
    host.save_image(tmp_path / "a.png", img)
    got = np.asarray(Image.open(tmp_path / "a.png").convert("RGB")).astype(np.int64)
    v = img
    with np.errstate(invalid="ignore", over="ignore"):
        srgb = np.where(v <= np.float32(0.0031308), np.float32(12.92) * v,
                        np.float32(1.055) * np.power(v, np.float32(1.0 / 2.4), dtype=np.float32) - np.float32(0.055))
        want = np.clip(np.float32(255.0) * srgb.astype(np.float32), 0, 255).astype(np.int64)
    assert got.shape == (9, 13, 3)
    diff = np.abs(got - want)
    assert diff.max() <= 1 and (diff == 0).mean() > 0.99
    assert got[0, 0].tolist() == [0, 3, 10] and got[0, 1].tolist() == [255, 0, 255]


# ---------------------------------------------------------------------------
# OBJ parsing: rounding of decimals and the binary cache
# ---------------------------------------------------------------------------
def _nearest_float32(text):
    """Correctly rounded (ties to even) binary32 of a decimal string, by
    exact rational comparison."""
    from fractions import Fraction
    x = Fraction(text)
    f = np.float32(float(x))          # within one step of the answer
    cands = {f, np.nextafter(f, np.float32(np.inf)), np.nextafter(f, np.float32(-np.inf))}
    best = None
    for c in cands:
        if not np.isfinite(c):
            continue
        d = abs(Fraction(float(c)) - x)
        key = (d, int(np.float32(c).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def test_obj_float_parsing_is_correctly_rounded(tmp_path):
    """Coordinates are the nearest binary32 of their decimal text (ties to
    even): short and long decimals, exponents across the float range,
    exact midpoints between neighbouring floats, and decimals a hair off a
    midpoint (where rounding through binary64 first goes wrong)."""
    from fractions import Fraction
    rng = np.random.default_rng(12)
    vals = []
    for _ in range(2000):
        kind = rng.integers(0, 6)
        if kind == 0:
            vals.append("%.*g" % (int(rng.integers(1, 12)), rng.normal() * 10.0 ** rng.integers(-8, 9)))
        elif kind == 1:
            vals.append("%.*f" % (int(rng.integers(0, 10)), rng.normal() * 100))
        elif kind == 2:
            vals.append("%.*e" % (int(rng.integers(0, 20)), rng.normal() * 10.0 ** rng.integers(-40, 38)))
        elif kind == 3:
            digits = "".join(str(d) for d in rng.integers(0, 10, size=int(rng.integers(1, 25))))
            vals.append(("-" if rng.integers(0, 2) else "+") + digits[:3] + "." + digits[3:])
        else:
            f = np.float32(rng.normal() * 10.0 ** rng.integers(-5, 6))
            g = np.nextafter(f, np.float32(np.inf))
            mid = (Fraction(float(f)) + Fraction(float(g))) / 2
            if kind == 5:   # off the midpoint by far less than a binary64 step
                mid += Fraction(int(rng.choice([-1, 1])), 10 ** 40) * abs(mid)
                mid = Fraction(round(mid * 10 ** 60), 10 ** 60)
            vals.append(_frac_text(mid))
    while len(vals) % 3:
        vals.append("0.5")
    lines = ["v %s %s %s" % tuple(vals[i:i + 3]) for i in range(0, len(vals), 3)]
    lines.append("f " + " ".join(str(i + 1) for i in range(len(vals) // 3)))     # one fan over every position
    (tmp_path / "f.obj").write_text("\n".join(lines) + "\n")
    body = '<mesh type="obj"><string name="filename" value="%s"/></mesh>' % (tmp_path / "f.obj")
    got = load(tmp_path, body + LIGHT).desc.meshes[0].V.reshape(-1)
    want = np.array([_nearest_float32(v) for v in vals], np.float32)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert bad.size == 0, [(vals[i], got[i], want[i]) for i in bad[:5]]


def _frac_text(fr):
    """Exact decimal expansion of a dyadic fraction."""
    from fractions import Fraction
    neg = fr < 0
    fr = abs(fr)
    k = 0
    while (fr * 10 ** k).denominator != 1:
        k += 1
    digits = str((fr * 10 ** k).numerator).rjust(k + 1, "0")
    s = digits[:-k] + "." + digits[-k:] if k else digits
    return ("-" if neg else "") + s


def test_obj_mesh_cache(tmp_path, monkeypatch):
    """With DRAWLAB_MESH_CACHE set, the parsed OBJ is stored once and read
    back while the file's path, size and modification time match."""
    cache = tmp_path / "cache"
    cache.mkdir()
    monkeypatch.setenv("DRAWLAB_MESH_CACHE", str(cache))
    obj = tmp_path / "m.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvn 0 0 1\nvt 0 0\nvt 1 0\nvt 0 1\nvt 1 1\n"
                   "f 1/1/1 2/2/1 4/4/1 3/3/1\n")
    body = ('<mesh type="obj"><string name="filename" value="%s"/><transform name="toWorld">'
            '<translate value="0, 0, 1"/></transform></mesh>' % obj)
    first = load(tmp_path, body + LIGHT).desc.meshes[0]
    files = list(cache.iterdir())
    assert len(files) == 1 and files[0].suffix == ".dlmesh"
    # answered from the cache: change a cached position, the load shows it
    raw = bytearray(files[0].read_bytes())
    header = 8 + 8 + 8 + 4 + len(str(obj.resolve()).encode()) + 8 + 8 + 4 + 4
    struct.pack_into("<f", raw, header, 7.5)
    files[0].write_bytes(bytes(raw))
    again = load(tmp_path, body + LIGHT).desc.meshes[0]
    assert again.V[0].tolist() == [7.5, 0.0, 1.0]                      # transform applied after the cache
    assert np.array_equal(again.F, first.F) and np.array_equal(again.UV, first.UV)
    # a changed source file is parsed again (and re-cached)
    obj.write_text(obj.read_text() + "# touched\n")
    os.utime(obj, ns=(1, 10 ** 18))
    fresh = load(tmp_path, body + LIGHT).desc.meshes[0]
    assert np.array_equal(fresh.V, first.V)
    # an unreadable cache file is ignored
    for f in cache.iterdir():
        f.write_bytes(b"junk")
    assert np.array_equal(load(tmp_path, body + LIGHT).desc.meshes[0].V, first.V)
    # without the variable nothing is cached
    monkeypatch.delenv("DRAWLAB_MESH_CACHE")
    for f in cache.iterdir():
        f.unlink()
    load(tmp_path, body + LIGHT)
    assert not list(cache.iterdir())
