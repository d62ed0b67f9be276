"""Committed fixtures (tests/golden/c1_small.npz, written by make_golden.py).

CPU: the oracle and the host-side generators reproduce every array bit for
bit.  GPU: the HIP path (grid build, primary rays, closest / any hit, shadow,
reflection and ambient-occlusion ray generation) reproduces them through the C
ABI."""
import importlib.util
import os

import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "c1_small.npz")
_spec = importlib.util.spec_from_file_location(
    "make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
_mg = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_mg)
SPEC = vars(_mg)


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, (what, a.shape, b.shape, a.dtype, b.dtype)
    bad = np.nonzero(_bits(a).reshape(-1) != _bits(b).reshape(-1))[0]
    assert bad.size == 0, f"{what}: {bad.size} differ, first {bad[:5]}"


def _frame(g):
    f = g["frame_f"]
    n = [int(x) for x in g["frame_n"]]
    return {"bmin": f[0:3], "bmax": f[3:6], "cell": f[6:9], "inv_cell": f[9:12], "n": n,
            "m": n[0] * n[1] * n[2]}


def _rays(g, prefix):
    return {k: g[prefix + k] for k in orc.RAY_FIELDS}


def test_fixture_is_nontrivial(gold):
    assert (gold["hit_id"] >= 0).mean() > 0.2
    assert gold["shadow_occluded"].sum() > 10 and gold["ao_occluded"].sum() > 10
    assert (gold["refl_hit_id"] >= 0).sum() > 10
    assert gold["refs"].size == gold["counts"].sum() == gold["offsets"][-1]


def test_oracle_reproduces_golden(gold):
    want = SPEC["make"]()
    assert sorted(want) == sorted(gold)
    for k in gold:
        _same(want[k], gold[k], k)


@pytest.mark.gpu
def test_hip_path_reproduces_golden(gpu_ctx, gold):
    import torch
    v, nrm = gold["verts"], gold["normals"]
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, nrm)
    sc.build_grid(SPEC["LAMBDA"])
    info, offs, refs = sc.download_grid()
    fr = _frame(gold)
    for k in ("bmin", "bmax", "cell", "inv_cell"):
        _same(np.float32(info[k]), fr[k], "frame " + k)
    assert list(info["n"]) == fr["n"]
    _same(offs, gold["offsets"], "offsets")
    _same(refs, gold["refs"], "refs")
    c = gold["camera"]
    cam = grt.Camera()
    for i, k in enumerate(("pos", "base", "nu", "nv")):
        for j in range(3):
            getattr(cam, k)[j] = float(c[3 * i + j])
    cam.width, cam.height = SPEC["WIDTH"], SPEC["HEIGHT"]
    n = cam.width * cam.height
    rb = grt.RayBatch.empty(n)
    gpu_ctx.gen_primary(cam, rb)
    got = rb.numpy()
    for k in orc.RAY_FIELDS:
        _same(got[k], gold["ray_" + k], "primary " + k)
    hb = grt.HitBatch.empty(n)
    cnt = torch.zeros(2 * n, dtype=torch.int32, device="cuda:0")
    sc.trace_closest(rb, hb, image_width=cam.width, counters=cnt)
    h = hb.numpy()
    for k in ("id", "t", "u", "v"):
        _same(h[k], gold["hit_" + k], "hit " + k)
    _same(cnt.cpu().numpy().view(np.uint32).reshape(n, 2), gold["hit_counters"], "counters")
    # shadow
    sb = grt.RayBatch.empty(n)
    sc.gen_shadow(rb, hb, SPEC["LIGHT"], SPEC["OFFSET"], sb)
    for k, a in sb.numpy().items():
        _same(a, gold["shadow_" + k], "shadow " + k)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    sc.trace_any(sb, occ, image_width=cam.width)
    _same(occ.cpu().numpy(), gold["shadow_occluded"], "shadow occluded")
    # reflection
    fb = grt.RayBatch.empty(n)
    sc.gen_reflection(rb, hb, SPEC["OFFSET"], fb)
    for k, a in fb.numpy().items():
        _same(a, gold["refl_" + k], "reflection " + k)
    fh = grt.HitBatch.empty(n)
    sc.trace_closest(fb, fh)
    for k, a in fh.numpy().items():
        _same(a, gold["refl_hit_" + k], "reflection hit " + k)
    # ambient occlusion
    spp = SPEC["AO_SPP"]
    ab = grt.RayBatch.empty(n * spp)
    sc.gen_ao(rb, hb, spp, SPEC["AO_SEED"], SPEC["OFFSET"], SPEC["AO_TMAX"], ab)
    for k, a in ab.numpy().items():
        _same(a, gold["ao_" + k], "ao " + k)
    aocc = torch.empty(n * spp, dtype=torch.uint8, device="cuda:0")
    sc.trace_any(ab, aocc)
    _same(aocc.cpu().numpy(), gold["ao_occluded"], "ao occluded")
