"""GPU tests of the two reference-facing boundaries.

1. The flat C ABI of the reference's GPU tracer (include/grt/rt_cabi.h, the
   calls of src/rtcuda/RayTracing.h:6-44), fed exactly as
   rtc::CudaRenderer::initialize fed it (CudaRenderer.cpp:65-209): a grid built
   on the host, three float4 per triangle, {start, count} cell pointers and
   3*id triangle references.  The rendered RGBA8 frame and the hit records must
   equal the oracle's bit for bit.
2. The C++ plugin classes (include/grt/rt_compat.h), driven by a compiled
   program (grid-rt_amd/examples/render_compat.cpp) the way the reference's
   Canvas drove its renderer plugin: builder plugin, PinholeCamera, renderer
   newFrame + render, plus per-ray traceNearestInstance calls.
"""
import os
import subprocess

import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import SEED_C1
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

EYE, AT = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0)


def _cuda_renderer_buffers(v, nrm, lam):
    """The host-side buffers CudaRenderer::initialize builds (CudaRenderer.cpp:86-199)."""
    fr = orc.grid_frame(v, lam)
    _, offs, refs = orc.grid_build(v, fr)
    nt = v.shape[0]
    v4 = np.ones((nt, 3, 4), np.float32)
    v4[:, :, :3] = v.reshape(nt, 3, 3)
    n4 = np.ones((nt, 3, 4), np.float32)
    n4[:, :, :3] = nrm.reshape(nt, 3, 3)
    counts = np.diff(offs.astype(np.int64))
    cell_ptrs = np.stack([offs[:-1].astype(np.int64), counts], axis=1).astype(np.int32)
    cell_ids = (refs.astype(np.int64) * 3).astype(np.int32)
    return fr, offs, refs, v4.reshape(-1), n4.reshape(-1), cell_ptrs.reshape(-1), cell_ids


def _oracle_frame(v, nrm, fr, offs, refs, cam):
    rays = orc.primary_rays(grt.camera_dict(cam))
    o = orc.trace_closest(fr, offs, refs, v, rays)
    return o, orc.headlight_image(rays, o, nrm)


class _Cabi:
    def __init__(self):
        self.L = grt.lib()

    def __call__(self, name, *args):
        rc = getattr(self.L, name)(*args)
        return rc

    def ok(self, name, *args):
        rc = self(name, *args)
        assert rc == grt.GRT_OK, f"{name}: {rc} {self.L.grtLastError().decode()}"

    def upload(self, fr, v4, n4, cell_ptrs, cell_ids):
        self.ok("grtTransferTriangleVertices", v4.ctypes.data, v4.size)
        self.ok("grtTransferTriangleNormals", n4.ctypes.data, n4.size)
        self.ok("grtTransferCellPointers", cell_ptrs.ctypes.data, cell_ptrs.size)
        self.ok("grtTransferCellTriangleIds", cell_ids.ctypes.data, cell_ids.size)
        self.ok("grtSetBoxMin", *map(float, fr["bmin"]))
        self.ok("grtSetBoxMax", *map(float, fr["bmax"]))
        self.ok("grtSetGridParameters", *map(float, fr["cell"]), *map(int, fr["n"]))

    def camera(self, cam):
        self.ok("grtSetScreenSize", cam.width, cam.height)
        self.ok("grtSetCameraParameters", *cam.pos, *cam.base, *cam.nu, *cam.nv)

    def hits(self, n):
        out = {"id": np.empty(n, np.int32)}
        for k in ("t", "u", "v"):
            out[k] = np.empty(n, np.float32)
        self.ok("grtGetHitRecords", *(out[k].ctypes.data for k in ("id", "t", "u", "v")))
        return out


@pytest.fixture(scope="module")
def cabi(gpu_ctx):
    c = _Cabi()
    assert c.L.rtInitContext()
    yield c
    c.ok("rtShutdownContext")


def test_cabi_frame_matches_oracle(cabi, c1_scene):
    import torch
    v, nrm = c1_scene
    fr, offs, refs, v4, n4, cp, ci = _cuda_renderer_buffers(v, nrm, 5.0)
    cam = grt.camera_lookat(EYE, AT, (0, 1, 0), 60.0, 320, 240)
    n = cam.width * cam.height
    frame = torch.full((n,), -5, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    assert cabi("grtRayTrace") == grt.GRT_NOT_READY            # nothing set yet
    cabi.ok("grtRegisterBuffer", frame.data_ptr())
    cabi.upload(fr, v4, n4, cp, ci)
    cabi.camera(cam)
    cabi.ok("grtSetBlockSize", 8, 8)
    cabi.ok("grtRayTrace")
    o, img = _oracle_frame(v, nrm, fr, offs, refs, cam)
    g = cabi.hits(n)
    for k in ("id", "t", "u", "v"):
        assert_same(g[k], o[k], f"hit.{k}")
    assert_same(frame.cpu().numpy().view(np.uint32), img, "RGBA8 frame")
    assert np.count_nonzero(o["id"] >= 0) > n // 20
    # a second frame with another camera reuses the uploaded scene
    cam2 = grt.camera_lookat((1.6, 0.9, 0.2), (0.5, 0.4, 0.5), (0, 1, 0), 45.0, 320, 240)
    cabi.camera(cam2)
    cabi.ok("grtRayTrace")
    o2, img2 = _oracle_frame(v, nrm, fr, offs, refs, cam2)
    assert_same(cabi.hits(n)["id"], o2["id"], "second frame hit.id")
    assert_same(frame.cpu().numpy().view(np.uint32), img2, "second RGBA8 frame")
    cabi.ok("grtUnregisterBuffer", frame.data_ptr())
    assert cabi("grtRayTrace") == grt.GRT_NOT_READY


def test_cabi_cell_layout_and_validation(cabi):
    """Cells whose id ranges are stored out of order give the same frame; ids
    that are not 3*t and pointer counts that do not match the grid are refused
    on the host, before anything runs on the device."""
    import torch
    v, nrm = grt.gen_spheres(SEED_C1 ^ 0x55, 24, 0.03, 0.09)
    fr, offs, refs, v4, n4, cp, ci = _cuda_renderer_buffers(v, nrm, 3.0)
    # store the cells' id ranges back to front, leaving a gap of 7 ids
    pairs = cp.reshape(-1, 2).astype(np.int64)
    m = pairs.shape[0]
    ids2, ptr2 = [], np.zeros((m, 2), np.int64)
    pos = 7
    chunks = []
    for c in range(m - 1, -1, -1):
        s, k = pairs[c]
        ptr2[c] = (pos, k)
        chunks.append(ci[s:s + k])
        pos += k
    ids2 = np.concatenate([np.full(7, 3, np.int32)] + chunks).astype(np.int32)
    ptr2 = ptr2.astype(np.int32).reshape(-1)
    cam = grt.camera_lookat(EYE, AT, (0, 1, 0), 60.0, 128, 96)
    n = cam.width * cam.height
    frame = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    cabi.ok("grtRegisterBuffer", frame.data_ptr())
    cabi.camera(cam)
    cabi.upload(fr, v4, n4, ptr2, ids2)
    cabi.ok("grtRayTrace")
    o, img = _oracle_frame(v, nrm, fr, offs, refs, cam)
    g = cabi.hits(n)
    for k in ("id", "t", "u", "v"):
        assert_same(g[k], o[k], f"hit.{k}")
    assert_same(frame.cpu().numpy().view(np.uint32), img, "RGBA8 frame")
    bad = ci.copy()
    bad[len(bad) // 2] += 1                                      # not a multiple of 3
    cabi.ok("grtTransferCellTriangleIds", bad.ctypes.data, bad.size)
    assert cabi("grtRayTrace") == grt.GRT_INVALID_PARAM
    cabi.ok("grtTransferCellTriangleIds", ci.ctypes.data, ci.size)
    cabi.ok("grtTransferCellPointers", cp.ctypes.data, cp.size - 2)   # one cell short
    assert cabi("grtRayTrace") == grt.GRT_INVALID_PARAM
    assert cabi("grtTransferCellPointers", cp.ctypes.data, 3) == grt.GRT_INVALID_PARAM
    assert cabi("grtTransferTriangleVertices", v4.ctypes.data, 13) == grt.GRT_INVALID_PARAM
    assert cabi("grtSetGridParameters", 0.1, 0.0, 0.1, 4, 4, 4) == grt.GRT_INVALID_PARAM
    cabi.ok("grtUnregisterBuffer", frame.data_ptr())


RENDER = os.path.join(os.path.dirname(grt.__file__), "..", "_build", "render_compat")


@pytest.mark.parametrize("accel", ["grid", "kd"])
def test_cpp_plugin_classes_render(gpu_ctx, c1_scene, tmp_path, accel):
    """The CudaRenderer frame through the plugin classes, with the grid builder
    plugin or the kd-tree builder plugin (whose closest hits are the grid's)."""
    v, nrm = c1_scene
    w, h, lam = 256, 200, 5.0
    scene = tmp_path / "scene.bin"
    with open(scene, "wb") as f:
        np.asarray([v.shape[0]], np.uint32).tofile(f)
        np.ascontiguousarray(v, np.float32).tofile(f)
        np.ascontiguousarray(nrm, np.float32).tofile(f)
    out = tmp_path / "frame.bin"
    cmd = [RENDER, str(scene), str(out), str(w), str(h), str(lam), *map(str, EYE), *map(str, AT)]
    if accel == "kd":
        cmd.append("kd")
    print("running", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    n = w * h
    raw = np.fromfile(out, np.uint8)
    assert raw.size == 4 * (5 * n + 9)
    words = raw.view(np.uint32)
    g = {"id": words[:n].view(np.int32), "t": words[n:2 * n].view(np.float32),
         "u": words[2 * n:3 * n].view(np.float32), "v": words[3 * n:4 * n].view(np.float32)}
    rgba = words[4 * n:5 * n]
    gn = words[5 * n:5 * n + 3].view(np.int32)
    gb = words[5 * n + 3:].view(np.float32)
    fr = orc.grid_frame(v, lam)
    if accel == "kd":
        assert list(gn) == [-1, -1, -1]
        assert_same(gb[:3], v.reshape(-1, 3).min(axis=0), "kd-tree bmin")
        assert_same(gb[3:], v.reshape(-1, 3).max(axis=0), "kd-tree bmax")
    else:
        assert list(gn) == list(fr["n"])
        assert_same(gb[:3], np.asarray(fr["bmin"], np.float32), "grid bmin")
        assert_same(gb[3:], np.asarray(fr["bmax"], np.float32), "grid bmax")
    _, offs, refs = orc.grid_build(v, fr)
    cam = grt.camera_lookat(EYE, AT, (0, 1, 0), 60.0, w, h)
    o, img = _oracle_frame(v, nrm, fr, offs, refs, cam)
    for k in ("id", "t", "u", "v"):
        assert_same(g[k], o[k], f"hit.{k}")
    assert_same(rgba, img, "RGBA8 frame")


PHONG = os.path.join(os.path.dirname(grt.__file__), "..", "_build", "render_phong")


@pytest.mark.parametrize("colours", [False, True])
def test_cpp_phong_renderer(gpu_ctx, c1_scene, tmp_path, colours):
    """grt::HipGridRenderer (the rtgl::GpuRenderer counterpart): primary,
    shadow, Phong, one reflection with its shadow, against the oracle."""
    from test_gpu_parity import _phong_frame_oracle
    v, nrm = c1_scene
    w, h, lam = 160, 120, 5.0
    scene = tmp_path / "scene.bin"
    with open(scene, "wb") as f:
        np.asarray([v.shape[0]], np.uint32).tofile(f)
        np.ascontiguousarray(v, np.float32).tofile(f)
        np.ascontiguousarray(nrm, np.float32).tofile(f)
    out = tmp_path / "frame.bin"
    cmd = [PHONG, str(scene), str(out), str(w), str(h), str(lam), *map(str, EYE), *map(str, AT)]
    diffuse = None
    if colours:
        diffuse = np.random.default_rng(5).uniform(0.1, 1.0, (v.shape[0], 3)).astype(np.float32)
        cfile = tmp_path / "colours.bin"
        diffuse.tofile(cfile)
        cmd.append(str(cfile))
    print("running", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    n = w * h
    raw = np.fromfile(out, np.float32)
    assert raw.size == 4 * n
    rgb, rgba = raw[:3 * n].reshape(n, 3), raw[3 * n:].view(np.uint32)
    fr = orc.grid_frame(v, lam)
    _, offs, refs = orc.grid_build(v, fr)
    cam = grt.camera_lookat(EYE, AT, (0, 1, 0), 60.0, w, h)
    light = (1e10, 1e10, 1e10)
    _, o_rgb, o_img, o, _ = _phong_frame_oracle(fr, offs, refs, v, nrm, cam, light, {}, diffuse)
    assert_same(rgb, o_rgb, "rgb")
    assert_same(rgba, o_img, "RGBA8 frame")
    assert np.count_nonzero(o["id"] >= 0) > n // 20
