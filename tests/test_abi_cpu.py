"""The C ABI library loads and exports every entry point include/*.h declares;
host-side helpers (camera basis, filter table, tangents) agree bit-for-bit
with the oracle.  No GPU work."""
import ctypes
import os
import re

import numpy as np

import drawlab_amd
import oracle
from drawlab_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*(?:dl_status|void|const char\*)\s+(dl_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = drawlab_amd.lib()
    decl = declared_functions()
    assert decl, "no declarations found"
    assert decl == set(drawlab_amd.EXPORTED_SYMBOLS)
    for name in decl:
        assert hasattr(L, name), name


def test_create_without_device_fails_cleanly():
    # On a machine without a GPU the context cannot be created; the call
    # must report it, not crash.
    try:
        import torch  # noqa: F401  (only to know whether a GPU exists)
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        return
    h = ctypes.c_void_p()
    rc = drawlab_amd.lib().dl_hip_create(0, ctypes.byref(h))
    assert rc != 0


def test_camera_basis_matches_oracle():
    for cam in (scenes.cornell_box().camera, scenes.Camera(eye=(0.4, 0.9, 5.0), lookat=(0.4, -0.1, 0.0),
                                                          fov=38.0, width=1920, height=1080)):
        d = drawlab_amd.camera_desc(cam)
        s = oracle.Scene()
        s.set_camera(cam.eye, cam.lookat, cam.up, cam.fov, cam.width, cam.height)
        ref = s.camera()
        got = np.array([list(d.eye), list(d.U), list(d.V), list(d.W)], np.float32)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_filter_tables_match_oracle():
    for ftype, r, p0, p1 in [(0, 2.0, 0.5, 0.0), (1, 0.5, 0, 0), (2, 1.0, 0, 0), (3, 2.0, 1 / 3, 1 / 3),
                             (0, 3.0, 1.0, 0.0)]:
        f = drawlab_amd.film_desc(scenes.RFilter(ftype, r, p0, p1))
        s = oracle.Scene()
        s.set_filter(ftype, r, p0, p1)
        assert np.array_equal(np.array(list(f.table), np.float32).view(np.uint32), s.filter_table().view(np.uint32))


def test_lib_has_no_oracle_dependency():
    """The product library must not link the oracle."""
    data = open(drawlab_amd.LIB_PATH, "rb").read()
    assert b"liboracle" not in data and b"dlo_" not in data


def test_service_start_without_device_or_bad_config():
    L = drawlab_amd._service_lib()
    h = ctypes.c_void_p()
    cfg = drawlab_amd.dl_service_config(0, 0, 0, 0, 1 << 30)   # lanes = 0
    assert L.dl_hip_service_start(ctypes.byref(cfg), ctypes.byref(h)) == 1
    assert b"lanes" in L.dl_hip_service_last_error(None)
    cfg = drawlab_amd.dl_service_config(0, 1, 7, 0, 1 << 30)   # unknown policy
    assert L.dl_hip_service_start(ctypes.byref(cfg), ctypes.byref(h)) == 1
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        cfg = drawlab_amd.dl_service_config(0, 2, 0, 0, 1 << 30)
        st = L.dl_hip_service_start(ctypes.byref(cfg), ctypes.byref(h))
        assert st != 0 and not h.value


def test_host_library_exports_every_declared_symbol():
    from drawlab_amd import host
    text = open(os.path.join(os.path.dirname(drawlab_amd.LIB_PATH), "..", "..", "include", "drawlab_host.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decl = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(dlx_\w+)\s*\(", text, flags=re.M))
    assert decl == set(host.HOST_SYMBOLS)
    L = host.host_lib()
    for name in decl:
        assert hasattr(L, name), name


def test_ctypes_structures_match_the_c_layout(tmp_path):
    """Every ctypes mirror of an include/dl_hip.h structure has the C
    structure's size and each field's offset: a C program built from the
    header (its fields named after the mirror's) prints them."""
    import subprocess
    structs = [getattr(drawlab_amd, n) for n in dir(drawlab_amd)
               if n.startswith("dl_") and isinstance(getattr(drawlab_amd, n), type)
               and issubclass(getattr(drawlab_amd, n), ctypes.Structure)]
    assert len(structs) >= 15
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "dl_hip.h"', 'int main(void) {']
    for S in structs:
        lines.append('printf("%%s %%zu\\n", "%s", sizeof(%s));' % (S.__name__, S.__name__))
        for f in S._fields_:
            lines.append('printf("%%s %%zu\\n", "%s.%s", offsetof(%s, %s));' % (S.__name__, f[0], S.__name__, f[0]))
    lines += ["return 0;", "}"]
    src, exe = tmp_path / "layout.c", tmp_path / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for S in structs:
        assert int(out[S.__name__]) == ctypes.sizeof(S), S.__name__
        for f in S._fields_:
            assert int(out["%s.%s" % (S.__name__, f[0])]) == getattr(S, f[0]).offset, (S.__name__, f[0])
