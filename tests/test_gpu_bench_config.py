"""Parity at the benchmarked configuration (the headline frame).

Workload: the full 2^20-triangle stadium -- 256 Gaussian clusters of 4096
triangles plus a 2-triangle ground plane 100x the cluster extent -- after
per-frame rigid cluster transforms, with its grid rebuilt on the device at the
two resolutions the frame autotuner chooses between: lambda 256 (about
2658 x 39 x 2658 = 2.8e8 cells, 1.5e7 references) and lambda 512 (about
3348 x 48 x 3348 = 5.4e8 cells, 2.3e7 references).  Both grids hold more than
2^24 cells and ground triangles that span millions of cells each.

The four traversal stages of the 1920 x 1080 frame -- primary closest hit,
shadow any hit, one mirror bounce, bounce shadow -- run under every schedule
the autotuner may pick (one ray per lane with per-lane or pooled tests,
persistent refill, 2, 4, 8 and 16 lanes per ray), over the whole image and
over compacted ray lists, and the Phong frame is shaded from them.

Bar: bit-exact against the oracle (its OpenMP path) -- transformed vertices,
grid frame, offsets, references, generated rays, hit id / t / u / v,
occlusion flags, per-ray cell and test counters, RGB and RGBA8 pixels.

The two-level grids the autotuner also chooses among (a coarse top grid at
lambda 4..16 with uniform sub-grids in its dense cells) are checked the same
way against the oracle's two-level restatement, and against the one-level
oracle's hits and occlusion.
"""
import numpy as np
import pytest
import torch

import grt
import pyoracle as orc
from conftest import SEED_C2

pytestmark = pytest.mark.gpu

N_CLUSTERS, PER_CLUSTER, N_CLUSTER_TRIS = 256, 4096, (1 << 20) - 2
WIDTH, HEIGHT = 1920, 1080
EYE, CENTER, UP, FOVY = (0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0.0, 1.0, 0.0), 60.0
LIGHT, OFFSET = (1000.0, 1000.0, 1000.0), 1e-4

# (schedule, refill threshold or lanes per ray): every candidate of the frame
# autotuner ('s', 'w', 'p', '2', '4', '8', 'x'), all with cooperative dense
# cells at the bench's threshold of 16.
SCHEDULES = [("simple", 16), ("pooled", 16), ("persistent", 16), ("split", 2), ("split", 4),
             ("split", 8), ("split", 16)]


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    diff = np.nonzero(bits(a) != bits(b))[0]
    assert diff.size == 0, (f"{what}: {diff.size} of {a.size} differ, first at {diff[:5]}: "
                            f"{a[diff[:5]]} vs {b[diff[:5]]}")


@pytest.fixture(scope="module")
def stadium():
    return grt.gen_stadium(SEED_C2, N_CLUSTERS, PER_CLUSTER, N_CLUSTER_TRIS)


def _oracle_pose(stadium, frame):
    v, nrm, centers = stadium
    xf = grt.gen_cluster_transforms(SEED_C2, frame, centers, 0.05)
    ov, on = orc.transform(v, nrm, xf, PER_CLUSTER, N_CLUSTER_TRIS)
    return xf, ov, on


def _scene_like_bench(ctx, stadium, lam):
    """The bench's set-up: frame 0 transformed and built synchronously, the
    grid reserved with the bench's headroom (per-frame rebuilds never
    reallocate)."""
    v, nrm, _ = stadium
    sc = grt.Scene(ctx)
    sc.upload(v, nrm)
    xf0, _, _ = _oracle_pose(stadium, 0)
    sc.transform_clusters(xf0, PER_CLUSTER, N_CLUSTER_TRIS)
    sc.build_grid(lam)
    info = sc.grid_info()
    sc.reserve_grid(int(info["m"] * 1.25) + 4096, int(info["r"] * 1.5) + 4096)
    return sc


def _check_grid(sc, ov, lam, tag):
    info, offs, refs = sc.download_grid()
    fr = orc.grid_frame(ov, lam)
    _, ooffs, orefs = orc.grid_build(ov, fr)
    for k in ("bmin", "bmax", "cell", "inv_cell"):
        assert_same(np.asarray(info[k], np.float32), np.asarray(fr[k], np.float32), f"{k} {tag}")
    assert list(info["n"]) == list(fr["n"]), tag
    assert info["m"] == fr["m"] and info["m"] > (1 << 24), (info["m"], tag)
    assert info["r"] == len(orefs), tag
    assert_same(offs, ooffs, f"offsets {tag}")
    assert_same(refs, orefs, f"refs {tag}")
    return fr, ooffs, orefs


@pytest.mark.parametrize("lam", [256.0, 512.0])
def test_bench_grid_per_frame_rebuild(gpu_ctx, stadium, lam):
    """Two frames of transform + asynchronous rebuild, as the timed loop runs
    them: vertices, normals and the whole grid equal the oracle's."""
    sc = _scene_like_bench(gpu_ctx, stadium, lam)
    try:
        for frame in (1, 2):
            xf, ov, on = _oracle_pose(stadium, frame)
            sc.transform_clusters(xf, PER_CLUSTER, N_CLUSTER_TRIS)
            sc.rebuild_grid_async(lam)
            sc.check_grid()
            gv, gn = sc.download_geometry()
            assert_same(gv.ravel(), ov.ravel(), f"vertices frame {frame}")
            assert_same(gn.ravel(), on.ravel(), f"normals frame {frame}")
            _check_grid(sc, ov, lam, f"[lambda {lam:g} frame {frame}]")
    finally:
        sc.close()


def _trace_closest_all(ctx, sc, rays, o, tag, list_args=None):
    n = rays.n
    active = None
    if list_args is not None:
        f = rays.numpy()
        active = f["tmin"] <= f["tmax"]
    for mode, param in SCHEDULES:
        ctx.set_trace_mode(mode, param)
        hb = grt.HitBatch.empty(n)
        cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
        if list_args is None:
            sc.trace_closest(rays, hb, image_width=WIDTH, counters=cnt)
        else:
            lst, count = list_args
            ctx.compact_rays(rays, WIDTH, lst, count, miss_hits=hb)
            sc.trace_closest_list(rays, hb, lst, count, counters=cnt)
        g = hb.numpy()
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        t = f"{tag} [{mode} {param}{' listed' if list_args is not None else ''}]"
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], o[k], f"hit.{k} {t}")
        sel = slice(None) if active is None else active
        assert_same(gc[sel, 0], o["cells"][sel], f"cells {t}")
        assert_same(gc[sel, 1], o["tests"][sel], f"tests {t}")
    ctx.set_trace_mode("simple", 16)
    return g


def _trace_any_all(ctx, sc, rays, o, tag, list_args=None):
    n = rays.n
    active = None
    if list_args is not None:
        f = rays.numpy()
        active = f["tmin"] <= f["tmax"]
    for mode, param in SCHEDULES:
        ctx.set_trace_mode(mode, param)
        occ = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
        cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
        if list_args is None:
            sc.trace_any(rays, occ, image_width=WIDTH, counters=cnt)
        else:
            lst, count = list_args
            ctx.compact_rays(rays, WIDTH, lst, count, occluded=occ)
            sc.trace_any_list(rays, occ, lst, count, counters=cnt)
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        t = f"{tag} [{mode} {param}{' listed' if list_args is not None else ''}]"
        assert_same(occ.cpu().numpy(), o["occluded"], f"occluded {t}")
        sel = slice(None) if active is None else active
        assert_same(gc[sel, 0], o["cells"][sel], f"cells {t}")
        assert_same(gc[sel, 1], o["tests"][sel], f"tests {t}")
    ctx.set_trace_mode("simple", 16)
    return occ


def _assert_rays(rb, o, tag):
    g = rb.numpy()
    for k in g:
        assert_same(g[k], o[k], f"{tag}.{k}")


@pytest.mark.parametrize("lam", [256.0, 512.0])
def test_bench_frame_four_stages(gpu_ctx, stadium, lam):
    """The whole 1920 x 1080 frame after one per-frame rebuild: every stage
    under every schedule, unlisted and listed, equals the oracle."""
    ctx = gpu_ctx
    ctx.set_coop(16)
    sc = _scene_like_bench(ctx, stadium, lam)
    try:
        xf, ov, on = _oracle_pose(stadium, 1)
        sc.transform_clusters(xf, PER_CLUSTER, N_CLUSTER_TRIS)
        sc.rebuild_grid_async(lam)
        sc.check_grid()
        fr, offs, refs = _check_grid(sc, ov, lam, f"[lambda {lam:g}]")
        n = WIDTH * HEIGHT
        lst = torch.empty(n, dtype=torch.int32, device="cuda:0")
        count = torch.zeros(1, dtype=torch.int32, device="cuda:0")
        tag = f"[lambda {lam:g}]"

        cam = grt.camera_lookat(EYE, CENTER, UP, FOVY, WIDTH, HEIGHT)
        prim = grt.RayBatch.empty(n)
        ctx.gen_primary(cam, prim)
        orays = orc.primary_rays(grt.camera_dict(cam))
        _assert_rays(prim, orays, f"primary rays {tag}")
        oh = orc.trace_closest(fr, offs, refs, ov, orays, counters=True)
        assert (oh["id"] >= 0).sum() > n // 4
        g = _trace_closest_all(ctx, sc, prim, oh, f"primary {tag}")
        hits = grt.HitBatch.from_numpy(g)

        shadow = grt.RayBatch.empty(n)
        sc.gen_shadow(prim, hits, LIGHT, OFFSET, shadow)
        osh = orc.gen_shadow(orays, oh, on, LIGHT, OFFSET)
        _assert_rays(shadow, osh, f"shadow rays {tag}")
        oocc = orc.trace_any(fr, offs, refs, ov, osh, counters=True)
        assert oocc["occluded"].sum() > 1000
        occ = _trace_any_all(ctx, sc, shadow, oocc, f"shadow {tag}")
        _trace_any_all(ctx, sc, shadow, oocc, f"shadow {tag}", (lst, count))

        refl = grt.RayBatch.empty(n)
        sc.gen_reflection(prim, hits, OFFSET, refl)
        orf = orc.gen_reflection(orays, oh, on, OFFSET)
        _assert_rays(refl, orf, f"bounce rays {tag}")
        orh = orc.trace_closest(fr, offs, refs, ov, orf, counters=True)
        assert (orh["id"] >= 0).sum() > 10000
        rg = _trace_closest_all(ctx, sc, refl, orh, f"bounce {tag}")
        _trace_closest_all(ctx, sc, refl, orh, f"bounce {tag}", (lst, count))
        rhits = grt.HitBatch.from_numpy(rg)

        shadow2 = grt.RayBatch.empty(n)
        sc.gen_shadow(refl, rhits, LIGHT, OFFSET, shadow2)
        osh2 = orc.gen_shadow(orf, orh, on, LIGHT, OFFSET)
        _assert_rays(shadow2, osh2, f"bounce shadow rays {tag}")
        oocc2 = orc.trace_any(fr, offs, refs, ov, osh2, counters=True)
        occ2 = _trace_any_all(ctx, sc, shadow2, oocc2, f"bounce shadow {tag}")
        _trace_any_all(ctx, sc, shadow2, oocc2, f"bounce shadow {tag}", (lst, count))

        # Phong frame of both levels (the bench's shading pass).
        rgb = torch.empty((n, 3), dtype=torch.float32, device="cuda:0")
        rgba = torch.empty(n, dtype=torch.int32, device="cuda:0")
        params = {"light": LIGHT}
        sc.shade_phong(prim, hits, rgb, occ, 0, params)
        sc.shade_phong(refl, rhits, rgb, occ2, 1, params, rgba=rgba)
        o_rgb, o_rgba = orc.shade_phong(orays, oh, on, oocc["occluded"], params, depth=0)
        o_rgb, o_rgba = orc.shade_phong(orf, orh, on, oocc2["occluded"], params, depth=1,
                                        rgb=o_rgb)
        assert_same(rgb.cpu().numpy().ravel(), o_rgb.ravel(), f"Phong rgb {tag}")
        assert_same(rgba.cpu().numpy().view(np.uint32), o_rgba, f"Phong rgba8 {tag}")
    finally:
        sc.close()


# Two-level grids (top lambda, lambda2, minimum references of a subdivided
# cell): the two the frame autotuner chooses among and two neighbours.
TWO_LEVEL = [(8.0, 4.0, 8), (16.0, 4.0, 8), (16.0, 4.0, 4), (8.0, 8.0, 8), (4.0, 4.0, 8)]


def _two_level_stage(ctx, sc, rays, fr, offs, refs, g2, ov, closest, tag, lst, count):
    """One stage through the two-level grid: whole image one ray per lane
    (simple, pooled, persistent: the same kernel), 4 and 16 lanes per ray
    (split), and as compacted lists, all equal to the oracle's two-level
    traversal; returns the oracle's result."""
    n = rays.n
    o = rays.numpy()
    active = o["tmin"] <= o["tmax"]
    fn = orc.grid2_trace_closest if closest else orc.grid2_trace_any
    ores = fn(fr, offs, refs, g2, ov, o, counters=True)
    for mode, param, listed in (("simple", 16, False), ("persistent", 16, False),
                                ("pooled", 16, False), ("split", 4, False), ("split", 16, False),
                                ("simple", 16, True), ("pooled", 16, True), ("split", 8, True),
                                ("persistent", 32, True)):
        ctx.set_trace_mode(mode, param)
        cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
        t = f"{tag} [{mode} {param}{' listed' if listed else ''}]"
        if closest:
            hb = grt.HitBatch.empty(n)
            if listed:
                ctx.compact_rays(rays, WIDTH, lst, count, miss_hits=hb)
                sc.trace_closest_list(rays, hb, lst, count, counters=cnt)
            else:
                sc.trace_closest(rays, hb, image_width=WIDTH, counters=cnt)
            g = hb.numpy()
            for k in ("id", "t", "u", "v"):
                assert_same(g[k], ores[k], f"hit.{k} {t}")
        else:
            occ = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
            if listed:
                ctx.compact_rays(rays, WIDTH, lst, count, occluded=occ)
                sc.trace_any_list(rays, occ, lst, count, counters=cnt)
            else:
                sc.trace_any(rays, occ, image_width=WIDTH, counters=cnt)
            assert_same(occ.cpu().numpy(), ores["occluded"], f"occluded {t}")
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        sel = active if listed else slice(None)
        assert_same(gc[sel, 0], ores["cells"][sel], f"cells {t}")
        assert_same(gc[sel, 1], ores["tests"][sel], f"tests {t}")
    ctx.set_trace_mode("simple", 16)
    return ores


@pytest.mark.parametrize("lam,lam2,min_refs", TWO_LEVEL)
def test_bench_two_level_frame(gpu_ctx, stadium, lam, lam2, min_refs):
    """The frame on the two-level grids the autotuner picks from (and two
    neighbours), set up as the bench does (frame 0 built synchronously with
    headroom, then a per-frame transform + asynchronous rebuild of both
    levels): both levels equal the oracle's, the four stages equal the
    oracle's two-level traversal bit for bit (hits, flags, top cells +
    sub-cells, tests), and the closest hits and occlusion equal the one-level
    oracle's (the reference semantics) except on rays whose two nearest hits
    lie within the exit bound's 2^-20 slack."""
    ctx = gpu_ctx
    v, nrm, _ = stadium
    sc = grt.Scene(ctx)
    try:
        sc.upload(v, nrm)
        xf0, _, _ = _oracle_pose(stadium, 0)
        sc.transform_clusters(xf0, PER_CLUSTER, N_CLUSTER_TRIS)
        sc.set_subgrid(lam2, min_refs)
        sc.build_grid(lam)
        info = sc.grid_info()
        sc.reserve_grid(int(info["m"] * 1.25) + 4096, int(info["r"] * 1.5) + 4096)
        xf, ov, on = _oracle_pose(stadium, 1)
        sc.transform_clusters(xf, PER_CLUSTER, N_CLUSTER_TRIS)
        sc.rebuild_grid_async(lam)
        sc.check_grid()
        tag = f"[two-level {lam:g}/{lam2:g}/{min_refs}]"
        info, offs_g, refs_g = sc.download_grid()
        fr = orc.grid_frame(ov, lam)
        _, offs, refs = orc.grid_build(ov, fr)
        assert_same(offs_g, offs, f"top offsets {tag}")
        assert_same(refs_g, refs, f"top refs {tag}")
        si = sc.subgrid_info()
        assert si["built"] and si["status"] == 0, si
        g2 = orc.grid2_build(ov, fr, offs, refs, min_refs, lam2, 16)
        cs, so, sr = sc.download_subgrid()
        assert_same(cs.ravel(), g2[0].ravel(), f"cell plans {tag}")
        assert_same(so, g2[1], f"sub-cell offsets {tag}")
        assert_same(sr, g2[2], f"sub-cell refs {tag}")

        n = WIDTH * HEIGHT
        lst = torch.empty(n, dtype=torch.int32, device="cuda:0")
        count = torch.zeros(1, dtype=torch.int32, device="cuda:0")
        cam = grt.camera_lookat(EYE, CENTER, UP, FOVY, WIDTH, HEIGHT)
        prim = grt.RayBatch.empty(n)
        ctx.gen_primary(cam, prim)
        oh = _two_level_stage(ctx, sc, prim, fr, offs, refs, g2, ov, True, f"primary {tag}",
                              lst, count)
        assert (oh["id"] >= 0).sum() > n // 4
        hits = grt.HitBatch.from_numpy(oh)
        orays = prim.numpy()
        o1 = orc.trace_closest(fr, offs, refs, ov, orays)
        same = (o1["id"] == oh["id"]) & (o1["t"].view(np.uint32) == oh["t"].view(np.uint32))
        assert same.mean() > 0.9999, (int((~same).sum()), tag)

        shadow = grt.RayBatch.empty(n)
        sc.gen_shadow(prim, hits, LIGHT, OFFSET, shadow)
        oocc = _two_level_stage(ctx, sc, shadow, fr, offs, refs, g2, ov, False,
                                f"shadow {tag}", lst, count)
        assert oocc["occluded"].sum() > 1000
        a1 = orc.trace_any(fr, offs, refs, ov, shadow.numpy())["occluded"]
        assert_same(oocc["occluded"], a1, f"two-level vs one-level occlusion {tag}")

        refl = grt.RayBatch.empty(n)
        sc.gen_reflection(prim, hits, OFFSET, refl)
        orh = _two_level_stage(ctx, sc, refl, fr, offs, refs, g2, ov, True, f"bounce {tag}",
                               lst, count)
        shadow2 = grt.RayBatch.empty(n)
        sc.gen_shadow(refl, grt.HitBatch.from_numpy(orh), LIGHT, OFFSET, shadow2)
        _two_level_stage(ctx, sc, shadow2, fr, offs, refs, g2, ov, False,
                         f"bounce shadow {tag}", lst, count)
    finally:
        sc.close()


@pytest.mark.parametrize("grid,sched,lists", [((16.0, (4.0, 8)), "ss24", "--c"),
                                              ((8.0, (4.0, 8)), "4448", "ccc"),
                                              ((64.0, None), "pp2p", "-c-")])
def test_bench_frame_branches_concurrent_same_as_sequential(gpu_ctx, grid, sched, lists):
    """The benchmarked frame with its shadow and bounce branches on two streams
    (the bounce rays, their shadow rays, list compaction and both traces on a
    second stream) writes exactly what the one-stream frame writes: hits,
    occlusion flags of both levels and the RGBA8 image, for two frames; on the
    one-level grid both branches run the persistent schedule at once (each
    stream has its own slot counter)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    lam, sub = grid
    fr = bench.Frame(grt, torch, lam=lam, sched=sched, sub=sub)
    fr.lists_kind = lists
    assert fr.concurrent_ok()
    try:
        for f in (1, 2):
            fr.run(f)
            torch.cuda.synchronize()
            want = {"hits": fr.hits.numpy(), "rhits": fr.rhits.numpy(),
                    "occ": fr.occ.cpu().numpy(), "occ2": fr.occ2.cpu().numpy(),
                    "rgba": fr.rgba.cpu().numpy()}
            for t in (fr.occ, fr.occ2, fr.rgba):
                t.fill_(0x5a)
            fr.run(f, concurrent=True)
            torch.cuda.synchronize()
            got = {"hits": fr.hits.numpy(), "rhits": fr.rhits.numpy(),
                   "occ": fr.occ.cpu().numpy(), "occ2": fr.occ2.cpu().numpy(),
                   "rgba": fr.rgba.cpu().numpy()}
            tag = f"[{lam}/{sub} {sched} {lists} frame {f}]"
            for k in ("hits", "rhits"):
                for fld in ("id", "t", "u", "v"):
                    assert_same(got[k][fld], want[k][fld], f"{k}.{fld} {tag}")
            for k in ("occ", "occ2", "rgba"):
                assert_same(got[k], want[k], f"{k} {tag}")
            assert (want["occ2"] != 0).any() and (want["hits"]["id"] >= 0).mean() > 0.5
    finally:
        fr.scene.close()
        fr.ctx.close()
