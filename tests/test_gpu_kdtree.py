"""GPU parity of the kd-tree comparison accelerator.

The device-built tree equals the host-only build of the same vertices; the
HIP traversal (grt_kdtree_trace_closest / _any) equals the oracle's kd
traversal bit for bit -- hit id, t, u, v, occlusion flags and the per-ray
leaf and test counters -- and its closest hits equal the HIP grid
traversal's on the stadium frame (the teapot-in-stadium comparison).
All calls go through the C ABI.
"""
import numpy as np
import pytest
import torch

import grt
import pyoracle as orc
from conftest import SEED_C1, SEED_C2, c1_camera
from test_gpu_parity import _degenerate_rays, _edge_scenes, _random_rays, assert_same

pytestmark = pytest.mark.gpu


def _kd(ctx, v, nrm):
    sc = grt.Scene(ctx)
    sc.upload(v, nrm)
    kd = grt.KdTree(sc)
    info = kd.info()
    nodes, elements = kd.download()
    host = grt.KdTree.from_host(v)
    hn, he = host.download()
    host.close()
    assert_same(nodes.ravel(), hn.ravel(), "device-built nodes vs host build")
    assert_same(elements, he, "device-built elements vs host build")
    assert info["depth"] <= 48
    return sc, kd, ((info["bmin"], info["bmax"]), nodes, elements)


def check_kd(kd, tree, v, rays, image_width=0):
    n = rays["ox"].shape[0]
    rb = grt.RayBatch.from_numpy(rays)
    hb = grt.HitBatch.empty(n)
    cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
    kd.trace_closest(rb, hb, image_width=image_width, counters=cnt)
    o = orc.kd_trace_closest(*tree, v, rays, counters=True)
    g = hb.numpy()
    gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
    for k in ("id", "t", "u", "v"):
        assert_same(g[k], o[k], f"kd hit.{k}")
    assert_same(gc[:, 0], o["leaves"], "kd leaves visited")
    assert_same(gc[:, 1], o["tests"], "kd triangle tests")
    occ = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
    cnt.fill_(-7)
    kd.trace_any(rb, occ, image_width=image_width, counters=cnt)
    oa = orc.kd_trace_any(*tree, v, rays, counters=True)
    gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
    assert_same(occ.cpu().numpy(), oa["occluded"], "kd occluded")
    assert_same(gc[:, 0], oa["leaves"], "kd any leaves visited")
    assert_same(gc[:, 1], oa["tests"], "kd any triangle tests")
    return g, o


def test_kd_c1_primary_random_and_degenerate_rays(gpu_ctx, c1_scene):
    v, nrm = c1_scene
    sc, kd, tree = _kd(gpu_ctx, v, nrm)
    try:
        rays = orc.primary_rays(grt.camera_dict(c1_camera()))
        g, o = check_kd(kd, tree, v, rays, image_width=512)
        assert (o["id"] >= 0).sum() > 10000
        check_kd(kd, tree, v, _random_rays(40000, 5, inside=False))
        check_kd(kd, tree, v, _random_rays(40000, 6, inside=True))
        check_kd(kd, tree, v, _degenerate_rays(8192, 43))
        # the same closest hits as the grid traversal
        sc.build_grid(5.0)
        hb = grt.HitBatch.empty(len(rays["ox"]))
        sc.trace_closest(grt.RayBatch.from_numpy(rays), hb, image_width=512)
        gg = hb.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], gg[k], f"kd vs grid hit.{k}")
        # shadow and reflection rays from those hits
        rb, hbk = grt.RayBatch.from_numpy(rays), grt.HitBatch.from_numpy(g)
        for gen in ("shadow", "reflection"):
            out = grt.RayBatch.empty(rb.n)
            if gen == "shadow":
                sc.gen_shadow(rb, hbk, (1000.0, 1000.0, 1000.0), 1e-3, out)
            else:
                sc.gen_reflection(rb, hbk, 1e-3, out)
            check_kd(kd, tree, v, out.numpy())
    finally:
        kd.close()
        sc.close()


@pytest.mark.parametrize("name", ["degenerate", "duplicates", "big_triangle", "hot_cell"])
def test_kd_edge_scenes(gpu_ctx, name):
    v = _edge_scenes()[name]
    sc, kd, tree = _kd(gpu_ctx, v, np.zeros_like(v))
    try:
        rays = _random_rays(20000, 23, inside=False)
        c = v.reshape(-1, 3).mean(axis=0)
        for k, ax in enumerate(("dx", "dy", "dz")):
            rays[ax][::2] = (c[k] - rays["o" + ax[1]][::2]).astype(np.float32)
        g, o = check_kd(kd, tree, v, rays)
        assert (o["id"] >= 0).sum() > 100
    finally:
        kd.close()
        sc.close()


def test_kd_stadium_frame_vs_oracle_and_grid(gpu_ctx):
    """The full 2^20-triangle stadium: 960 x 540 primary and shadow rays through
    the kd-tree equal the oracle's kd traversal, and the closest hits equal the
    grid traversal's (lambda 256)."""
    v, nrm, _ = grt.gen_stadium(SEED_C2, 256, 4096, (1 << 20) - 2)
    sc, kd, tree = _kd(gpu_ctx, v, nrm)
    try:
        cam = grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, 960, 540)
        rays = orc.primary_rays(grt.camera_dict(cam))
        g, o = check_kd(kd, tree, v, rays, image_width=960)
        assert (o["id"] >= 0).sum() > 100000
        sc.build_grid(256.0)
        n = len(rays["ox"])
        hb = grt.HitBatch.empty(n)
        rb = grt.RayBatch.from_numpy(rays)
        sc.trace_closest(rb, hb, image_width=960)
        gg = hb.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], gg[k], f"kd vs grid hit.{k}")
        sh = grt.RayBatch.empty(n)
        sc.gen_shadow(rb, grt.HitBatch.from_numpy(g), (1000.0, 1000.0, 1000.0), 1e-4, sh)
        check_kd(kd, tree, v, sh.numpy(), image_width=960)
    finally:
        kd.close()
        sc.close()


def test_kd_rejects_host_only_trace(gpu_ctx, c1_scene):
    v, _ = c1_scene
    host = grt.KdTree.from_host(v)
    try:
        with pytest.raises(grt.GrtError) as e:
            host.trace_closest(grt.RayBatch.empty(4), grt.HitBatch.empty(4))
        assert e.value.code == grt.GRT_NOT_READY
    finally:
        host.close()
