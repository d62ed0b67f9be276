"""CPU tests of the library boundary: libgrt.so loads, exports every function
the public headers declare, refuses bad arguments with status codes instead of
aborting, and fails cleanly where there is no GPU.  Host-only parts (scene
generators, the PinholeCamera plugin's camera set-up) are checked here too."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import ROOT, SEED_C1

INCLUDE = os.path.join(ROOT, "include", "grt")
C_HEADERS = ("grt.h", "grt_scenes.h", "rt_cabi.h")


def _declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b((?:grt|rt|grt_)\w*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", grt.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_function():
    exported = _exported()
    total = 0
    for h in C_HEADERS:
        names = _declared_functions(os.path.join(INCLUDE, h))
        assert len(names) >= 5, (h, names)
        missing = [n for n in names if n not in exported]
        assert not missing, f"{h}: declared but not exported: {missing}"
        total += len(names)
    assert total >= 69
    # every bound signature in the Python package is a real export
    assert not [n for n in grt._SIGS if n not in exported]


def test_cpp_plugin_classes_are_exported():
    out = subprocess.run(["nm", "-DC", "--defined-only", grt.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("grt::HipGridAccStruct::traceNearestGeometry", "grt::HipGridAccStruct::traceNearest",
                "grt::HipGridAccStructBuilder::buildGeometry", "grt::HipRenderer::render",
                "grt::HipGridRenderer::render", "grt::HipGridRenderer::rebuildGrid",
                "grt::PinholeCamera::getRayDirection", "grt::IAccStruct::traceAnyInstance",
                "grt::IPlugin::receiveParameter", "grt::ICamera::setLookAt"):
        assert sym in out, sym


_NO_GPU_SCRIPT = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
import grt
L = grt.lib()
h = ctypes.c_void_p()
rc = L.grt_init(0, ctypes.byref(h))
print("init", rc, L.grt_last_error().decode())
print("rtInitContext", bool(L.rtInitContext()))
print("trace", L.grtRayTrace(), L.grtLastError().decode())
"""


def test_init_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    p = subprocess.run([sys.executable, "-c", _NO_GPU_SCRIPT, os.path.join(ROOT, "grid-rt_amd")],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = dict(l.split(" ", 1) for l in p.stdout.strip().splitlines())
    rc = int(lines["init"].split()[0])
    assert rc == grt.GRT_DEVICE_ERROR, p.stdout
    assert lines["rtInitContext"] == "False"
    assert int(lines["trace"].split()[0]) == grt.GRT_NOT_READY


def test_null_arguments_are_refused():
    L = grt.lib()
    P = ctypes.c_void_p
    out = P()
    assert L.grt_scene_create(None, ctypes.byref(out)) == grt.GRT_INVALID_PARAM
    assert L.grt_shutdown(None) == grt.GRT_INVALID_PARAM
    assert L.grt_grid_build(None, 5.0) == grt.GRT_INVALID_PARAM
    assert L.grt_trace_closest(None, None, None, 0, 0, None) == grt.GRT_INVALID_PARAM
    assert L.grt_trace_any(None, None, None, 0, 0, None) == grt.GRT_INVALID_PARAM
    assert L.grt_batch_set(None, None, 0) == grt.GRT_INVALID_PARAM
    assert L.grt_serve_run(None, None, 0, None, None, None) == grt.GRT_INVALID_PARAM
    assert L.grt_last_error()            # a message is left behind
    # the flat ABI checks its arguments before touching the device
    assert L.grtTransferTriangleVertices(None, 12) == grt.GRT_INVALID_PARAM
    assert L.grtTransferCellPointers(None, 4) == grt.GRT_INVALID_PARAM
    assert L.grtSetScreenSize(0, 10) == grt.GRT_INVALID_PARAM
    assert L.grtRegisterBuffer(None) == grt.GRT_INVALID_PARAM
    assert L.grtGetHitRecords(None, None, None, None) == grt.GRT_NOT_READY


def test_generators_are_deterministic_and_seeded():
    a = grt.gen_spheres(SEED_C1, 8)
    b = grt.gen_spheres(SEED_C1, 8)
    c = grt.gen_spheres(SEED_C1 + 1, 8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])
    s1 = grt.gen_stadium(0x6772740000000002, 4, 256, 1022)
    s2 = grt.gen_stadium(0x6772740000000002, 4, 256, 1022)
    assert all(np.array_equal(x, y) for x, y in zip(s1, s2))
    assert s1[0].shape == (1024, 9)


def test_sphere_mesh_is_closed_and_nondegenerate():
    v, n = grt.gen_spheres(SEED_C1, 3)
    assert v.shape == (3 * 256, 9)
    for s in range(3):
        tri = v[s * 256:(s + 1) * 256].reshape(256, 3, 3).astype(np.float64)
        # closed surface: every edge is shared by exactly two triangles
        edges = {}
        for t in tri:
            pts = [tuple(np.round(p, 6)) for p in t]
            for i in range(3):
                e = tuple(sorted((pts[i], pts[(i + 1) % 3])))
                edges[e] = edges.get(e, 0) + 1
        assert set(edges.values()) == {2}
        area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1)
        assert area.min() > 0
        nn = np.linalg.norm(n[s * 256:(s + 1) * 256].reshape(-1, 3), axis=1)
        assert np.allclose(nn, 1.0, atol=1e-5)


_CAMERA_PROG = r"""
#include <cstdio>
#include "grt/rt_compat.h"
int main() {
    grt::ref_ptr<grt::PinholeCamera> cam(new grt::PinholeCamera);
    cam->setLookAt(0.5f, 0.5f, -1.5f, 0.5f, 0.5f, 0.0f, 0.0f, 1.0f, 0.0f);
    cam->setPerspective(60.0f, 1.0f, 1000.0f);
    cam->setViewport(64, 48);
    cam->newFrame();
    grt::vec3f d, o;
    const float px[3][2] = {{0, 0}, {31, 17}, {63, 47}};
    for (auto& p : px) {
        cam->getRayOrigin(o, p[0], p[1]);
        cam->getRayDirection(d, p[0], p[1]);
        std::printf("%a %a %a %a %a %a\n", o.x, o.y, o.z, d.x, d.y, d.z);
    }
    // a full turn about Y and a translation along the camera's own axes
    for (int k = 0; k < 4; ++k) cam->rotate(1.5707963267948966f, grt::RT_AXIS_Y);
    cam->translate(0.0f, 0.0f, 1.0f);
    cam->newFrame();
    cam->getRayOrigin(o, 0, 0);
    cam->getRayDirection(d, 31, 17);
    std::printf("%a %a %a %a %a %a\n", o.x, o.y, o.z, d.x, d.y, d.z);
    return cam->refCount() == 1 ? 0 : 1;
}
"""


def test_pinhole_camera_plugin_matches_oracle(tmp_path):
    src = tmp_path / "cam.cpp"
    src.write_text(_CAMERA_PROG)
    exe = tmp_path / "cam"
    pkg = os.path.join(ROOT, "grid-rt_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", str(src), "-o", str(exe),
                    f"-L{pkg}", "-lgrt", f"-Wl,-rpath,{pkg}"], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    rows = [[float.fromhex(x) for x in line.split()] for line in p.stdout.strip().splitlines()]
    cam = orc.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0, 1, 0), 60.0, 64, 48)
    rays = orc.primary_rays(cam)
    for row, (x, y) in zip(rows[:3], ((0, 0), (31, 17), (63, 47))):
        i = y * 64 + x
        want = [rays[k][i] for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
        assert np.array_equal(np.float32(row), np.float32(want)), (x, y, row, want)
    # a full turn leaves the basis (nearly) unchanged; +1 along the camera's Z
    # (which points away from the view direction) moves the eye to z = -2.5
    o, d = np.float32(rows[3][:3]), np.float32(rows[3][3:])
    assert np.allclose(o, (0.5, 0.5, -2.5), atol=1e-5)
    i = 17 * 64 + 31
    assert np.allclose(d, [rays[k][i] for k in ("dx", "dy", "dz")], atol=1e-5)


def test_phong_defaults_are_the_shader_constants():
    """grt_phong_defaults (host only, no GPU) equals the oracle's restatement of
    shading_FRAG.glsl:8-36."""
    p = grt.phong()
    for k, v in orc.PHONG_DEFAULTS.items():
        got = getattr(p, k)
        if isinstance(v, tuple):
            assert list(got) == [float(np.float32(x)) for x in v], k
        else:
            assert got == np.float32(v), k
