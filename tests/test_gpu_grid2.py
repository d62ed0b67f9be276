"""GPU parity of the two-level grid (the second level: uniform sub-grids in the
dense top cells, rebuilt on the device).

Bar: bit-exact against the oracle's two-level restatement -- the sub-grid
plan of every top cell, sub-cell offsets and references (ascending ids per
sub-cell), and the traversal's hit id / t / u / v, occlusion flags and
per-ray counters (top cells + sub-cells, tests) -- plus the one-level
oracle's hits (the reference semantics) on every ray whose two nearest
float64 hits are not within the exit bound's slack of each other.  All calls
go through the C ABI.
"""
import numpy as np
import pytest
import torch

import grt
import pyoracle as orc
from conftest import SEED_C1, SEED_C2, c1_camera
from test_gpu_parity import _degenerate_rays, _edge_scenes, _random_rays, assert_same

pytestmark = pytest.mark.gpu


def _built(ctx, v, nrm, lam, lam2, min_refs):
    sc = grt.Scene(ctx)
    sc.upload(v, nrm)
    sc.set_subgrid(lam2, min_refs)
    sc.build_grid(lam)
    return sc


def check_levels(sc, v, lam, lam2, min_refs, tag=""):
    """Top grid and second level equal the oracle's; returns the oracle's arrays."""
    info, offs, refs = sc.download_grid()
    fr = orc.grid_frame(v, lam)
    _, ooffs, orefs = orc.grid_build(v, fr)
    assert list(info["n"]) == list(fr["n"]), tag
    assert_same(offs, ooffs, f"top offsets {tag}")
    assert_same(refs, orefs, f"top refs {tag}")
    si = sc.subgrid_info()
    assert si["built"] and si["status"] == 0, si
    cs, so, sr = sc.download_subgrid()
    ocs, oso, osr = orc.grid2_build(v, fr, ooffs, orefs, min_refs, lam2, 16)
    assert_same(cs.ravel(), ocs.ravel(), f"cell plans {tag}")
    assert_same(so, oso, f"sub-cell offsets {tag}")
    assert_same(sr, osr, f"sub-cell refs {tag}")
    return fr, ooffs, orefs, (ocs, oso, osr)


# Trace modes: one ray per lane (simple; pooled and persistent launch the same
# kernel on a two-level grid), and 2, 8 and 16 lanes per ray walking the ray's
# top items in turn (the split schedule).
MODES = [("simple", 16), ("pooled", 16), ("persistent", 16), ("split", 2), ("split", 8),
         ("split", 16)]


def check_trace2(sc, fr, offs, refs, g2, v, rays, image_width=0, one_level=True, modes=MODES):
    n = rays["ox"].shape[0]
    rb = grt.RayBatch.from_numpy(rays)
    o = orc.grid2_trace_closest(fr, offs, refs, g2, v, rays, counters=True)
    oa = orc.grid2_trace_any(fr, offs, refs, g2, v, rays, counters=True)
    ctx = sc.ctx
    g = None
    for mode, param in modes:
        ctx.set_trace_mode(mode, param)
        tag = f"[{mode} {param}]"
        hb = grt.HitBatch.empty(n)
        cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
        sc.trace_closest(rb, hb, image_width=image_width, counters=cnt)
        gm = hb.numpy()
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        for k in ("id", "t", "u", "v"):
            assert_same(gm[k], o[k], f"two-level hit.{k} {tag}")
        assert_same(gc[:, 0], o["cells"], f"two-level cells visited {tag}")
        assert_same(gc[:, 1], o["tests"], f"two-level triangle tests {tag}")
        occ = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
        cnt.fill_(-7)
        sc.trace_any(rb, occ, image_width=image_width, counters=cnt)
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        assert_same(occ.cpu().numpy(), oa["occluded"], f"two-level occluded {tag}")
        assert_same(gc[:, 0], oa["cells"], f"two-level any cells {tag}")
        assert_same(gc[:, 1], oa["tests"], f"two-level any tests {tag}")
        g = g if g is not None else gm
    ctx.set_trace_mode("simple", 16)
    if one_level:
        # the reference semantics: the one-level grid's hits and flags
        o1 = orc.trace_closest(fr, offs, refs, v, rays)
        same = (o1["id"] == o["id"]) & (o1["t"].view(np.uint32) == o["t"].view(np.uint32))
        assert same.mean() > 0.9999, (int((~same).sum()), n)
        a1 = orc.trace_any(fr, offs, refs, v, rays)["occluded"]
        assert_same(oa["occluded"], a1, "two-level vs one-level occlusion")
    return g, o


@pytest.mark.parametrize("lam,lam2,min_refs", [(1.0, 4.0, 8), (2.0, 2.0, 4), (0.5, 8.0, 16)])
def test_grid2_c1(gpu_ctx, c1_scene, lam, lam2, min_refs):
    v, nrm = c1_scene
    sc = _built(gpu_ctx, v, nrm, lam, lam2, min_refs)
    try:
        fr, offs, refs, g2 = check_levels(sc, v, lam, lam2, min_refs)
        p = g2[0][:, 1]
        assert int(((p != 0) & (p >> 31 == 0)).sum()) > 10, "no dense cells"
        rays = orc.primary_rays(grt.camera_dict(c1_camera()))
        g, o = check_trace2(sc, fr, offs, refs, g2, v, rays, image_width=512)
        assert (o["id"] >= 0).sum() > 10000
        check_trace2(sc, fr, offs, refs, g2, v, _random_rays(30000, 7, inside=True))
        check_trace2(sc, fr, offs, refs, g2, v, _random_rays(30000, 8, inside=False))
        check_trace2(sc, fr, offs, refs, g2, v, _degenerate_rays(8192, 47), one_level=False)
    finally:
        sc.close()


@pytest.mark.parametrize("name", ["degenerate", "duplicates", "big_triangle", "hot_cell"])
def test_grid2_edge_scenes(gpu_ctx, name):
    """Duplicates give exact t ties inside sub-cells; the hot cell holds
    40,000 references (a 16^3 sub-grid streamed in 1024-reference rounds); a
    triangle spanning the whole grid sits in every sub-cell."""
    v = _edge_scenes()[name]
    sc = _built(gpu_ctx, v, np.zeros_like(v), 2.0, 4.0, 8)
    try:
        fr, offs, refs, g2 = check_levels(sc, v, 2.0, 4.0, 8, name)
        rays = _random_rays(20000, 29, inside=False)
        c = v.reshape(-1, 3).mean(axis=0)
        for k, ax in enumerate(("dx", "dy", "dz")):
            rays[ax][::2] = (c[k] - rays["o" + ax[1]][::2]).astype(np.float32)
        check_trace2(sc, fr, offs, refs, g2, v, rays)
    finally:
        sc.close()


def test_grid2_stadium_frame_rebuilds(gpu_ctx):
    """The full 2^20-triangle stadium, two frames of transform + asynchronous
    rebuild of both levels (lambda 8, lambda2 4, 8 references): both levels
    equal the oracle's, and 960 x 540 primary, shadow and bounce rays through
    them equal the oracle's two-level and one-level traversals."""
    v0, n0, centers = grt.gen_stadium(SEED_C2, 256, 4096, (1 << 20) - 2)
    sc = grt.Scene(gpu_ctx)
    try:
        sc.upload(v0, n0)
        sc.set_subgrid(4.0, 8)
        xf = grt.gen_cluster_transforms(SEED_C2, 0, centers, 0.05)
        sc.transform_clusters(xf, 4096, (1 << 20) - 2)
        sc.build_grid(8.0)
        info = sc.grid_info()
        sc.reserve_grid(int(info["m"] * 1.25) + 4096, int(info["r"] * 1.5) + 4096)
        for frame in (1, 2):
            xf = grt.gen_cluster_transforms(SEED_C2, frame, centers, 0.05)
            sc.transform_clusters(xf, 4096, (1 << 20) - 2)
            sc.rebuild_grid_async(8.0)
            sc.check_grid()
            v, nrm = orc.transform(v0, n0, xf, 4096, (1 << 20) - 2)
            fr, offs, refs, g2 = check_levels(sc, v, 8.0, 4.0, 8, f"frame {frame}")
        cam = grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, 960, 540)
        rays = orc.primary_rays(grt.camera_dict(cam))
        g, o = check_trace2(sc, fr, offs, refs, g2, v, rays, image_width=960)
        assert (o["id"] >= 0).sum() > 100000
        sh = orc.gen_shadow(rays, o, nrm, (1000.0, 1000.0, 1000.0), 1e-4)
        check_trace2(sc, fr, offs, refs, g2, v, sh, image_width=960)
        rf = orc.gen_reflection(rays, o, nrm, 1e-4)
        check_trace2(sc, fr, offs, refs, g2, v, rf, image_width=960)
    finally:
        sc.close()


def test_grid2_lists_and_capacity(gpu_ctx, c1_scene):
    """List traces walk both levels too; a rebuild whose second level outgrows
    its buffers traces the top grid alone (the one-level results) and the check
    reports GRT_CAPACITY; turning the level off restores the one-level path."""
    v, nrm = c1_scene
    sc = _built(gpu_ctx, v, nrm, 1.0, 2.0, 8)
    try:
        fr, offs, refs, g2 = check_levels(sc, v, 1.0, 2.0, 8)
        rays = orc.primary_rays(grt.camera_dict(c1_camera(200, 150)))
        n = len(rays["ox"])
        rb = grt.RayBatch.from_numpy(rays)
        lst = torch.empty(n, dtype=torch.int32, device="cuda:0")
        count = torch.zeros(1, dtype=torch.int32, device="cuda:0")
        hb = grt.HitBatch.empty(n)
        gpu_ctx.compact_rays(rb, 200, lst, count, miss_hits=hb)
        sc.trace_closest_list(rb, hb, lst, count)
        o = orc.grid2_trace_closest(fr, offs, refs, g2, v, rays)
        g = hb.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], o[k], f"listed two-level hit.{k}")
        # a finer second level than the buffers sized at lambda2 2 hold
        sc.set_subgrid(64.0, 2)
        sc.rebuild_grid_async(1.0)
        hb2 = grt.HitBatch.empty(n)
        sc.trace_closest(rb, hb2, image_width=200)
        with pytest.raises(grt.GrtError) as e:
            sc.check_grid()
        assert e.value.code == grt.GRT_CAPACITY
        o1 = orc.trace_closest(fr, offs, refs, v, rays)
        g2h = hb2.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g2h[k], o1[k], f"top grid alone hit.{k}")
        sc.set_subgrid(0.0, 8)
        sc.build_grid(1.0)
        assert not sc.subgrid_info()["built"]
        hb3 = grt.HitBatch.empty(n)
        cnt = torch.empty(2 * n, dtype=torch.int32, device="cuda:0")
        sc.trace_closest(rb, hb3, image_width=200, counters=cnt)
        o1c = orc.trace_closest(fr, offs, refs, v, rays, counters=True)
        assert_same(cnt.cpu().numpy().view(np.uint32).reshape(n, 2)[:, 0], o1c["cells"],
                    "one-level cells after turning the level off")
    finally:
        sc.close()


def test_grid2_rebuild_and_trace_replay_in_a_graph(gpu_ctx, c1_scene):
    """A two-level rebuild and a trace captured in a launch graph (under
    capture the rebuild stays on the captured stream instead of forking onto
    the context's second stream) give the eager results on every replay, with
    eager work between the replays."""
    v, nrm = c1_scene
    stream = grt.stream_create()
    torch.cuda.synchronize()
    gpu_ctx.set_stream(stream)
    sc = _built(gpu_ctx, v, nrm, 2.0, 4.0, 8)
    try:
        rays = orc.primary_rays(grt.camera_dict(c1_camera(200, 150)))
        n = len(rays["ox"])
        rb = grt.RayBatch.from_numpy(rays)
        hb = grt.HitBatch.empty(n)
        torch.cuda.synchronize()

        def work():
            sc.rebuild_grid_async(2.0)
            sc.trace_closest(rb, hb, image_width=200)

        work()
        gpu_ctx.synchronize()
        eager = hb.numpy()
        g = grt.Graph(gpu_ctx)
        g.begin()
        work()
        g.end()
        try:
            for rep in range(3):
                for t in hb.f.values():
                    t.fill_(-3)
                torch.cuda.synchronize()
                g.launch()
                gpu_ctx.synchronize()
                got = hb.numpy()
                for k in ("id", "t", "u", "v"):
                    assert_same(got[k], eager[k], f"replay {rep} hit.{k}")
                sc.trace_closest(rb, grt.HitBatch.empty(n), image_width=200)   # eager work
                gpu_ctx.synchronize()
        finally:
            g.close()
        sc.check_grid()
        assert sc.subgrid_info()["built"]
        fr, offs, refs, g2 = check_levels(sc, v, 2.0, 4.0, 8, "after the replays")
        o = orc.grid2_trace_closest(fr, offs, refs, g2, v, rays)
        for k in ("id", "t", "u", "v"):
            assert_same(eager[k], o[k], f"eager hit.{k}")
    finally:
        sc.close()
        gpu_ctx.synchronize()
        gpu_ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        grt.stream_destroy(stream)
