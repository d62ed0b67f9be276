"""GPU parity: the HIP path against the CPU oracle on identical seeded inputs.

Bar: bit-exact.  Grid frames, offsets and references are integers / exactly
specified binary32 values; traversal results (id, t, u, v) and the per-ray
counters (cells visited, triangle tests) are compared bit for bit because both
sides evaluate the same operation sequence (explicit FMAs, IEEE division and
sqrt, no contraction).  All calls go through the C ABI (libgrt.so).
"""
import math

import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import SEED_C1, SEED_C2, c1_camera, rng

pytestmark = pytest.mark.gpu


def bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    diff = np.nonzero(bits(a) != bits(b))[0]
    assert diff.size == 0, (f"{what}: {diff.size} of {a.size} differ, first at {diff[:5]}: "
                            f"{a[diff[:5]]} vs {b[diff[:5]]}")


def assert_frame(info, fr):
    for k in ("bmin", "bmax", "cell", "inv_cell"):
        assert_same(np.asarray(info[k], np.float32), np.asarray(fr[k], np.float32), k)
    assert list(info["n"]) == list(fr["n"])
    assert info["m"] == fr["m"]


def build_both(ctx, v, n, lam):
    sc = grt.Scene(ctx)
    sc.upload(v, n)
    sc.build_grid(lam)
    info, offs, refs = sc.download_grid()
    fr = orc.grid_frame(v, lam)
    _, ooffs, orefs = orc.grid_build(v, fr)
    return sc, fr, info, offs, refs, ooffs, orefs


def check_grid(ctx, v, n, lam):
    sc, fr, info, offs, refs, ooffs, orefs = build_both(ctx, v, n, lam)
    assert_frame(info, fr)
    assert_same(offs, ooffs, "offsets")
    assert info["r"] == len(orefs)
    assert_same(refs, orefs, "refs")
    return sc, fr, ooffs, orefs


# Every traversal schedule must give identical per-ray results and counters:
# (schedule, refill threshold or lanes per ray, cooperative threshold).  The
# per-lane, split and persistent schedules finish dense cells cooperatively;
# thresholds 1 and 2 make nearly every such cell cooperative, 0 turns the pass
# off.  The pooled schedule spreads every wave's tests over its lanes; the
# split schedule gives each ray 2..16 lanes that take its cells in turn.
MODES = [("simple", 16, 16), ("simple", 16, 0), ("pooled", 16, 16), ("persistent", 16, 0),
         ("persistent", 16, 16), ("persistent", 1, 1), ("persistent", 64, 2),
         ("split", 2, 16), ("split", 4, 0), ("split", 8, 1), ("split", 16, 16)]


def check_closest(sc, fr, offs, refs, v, rays, image_width=0):
    import torch
    n = rays["ox"].shape[0]
    o = orc.trace_closest(fr, offs, refs, v, rays, counters=True)
    rb = grt.RayBatch.from_numpy(rays)
    for mode, refill, coop in MODES:
        sc.ctx.set_trace_mode(mode, refill)
        sc.ctx.set_coop(coop)
        hb = grt.HitBatch.empty(n)
        cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
        sc.trace_closest(rb, hb, image_width=image_width, counters=cnt)
        g = hb.numpy()
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        tag = f"[{mode} param={refill} coop={coop}]"
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], o[k], f"hit.{k} {tag}")
        assert_same(gc[:, 0], o["cells"], f"cells visited {tag}")
        assert_same(gc[:, 1], o["tests"], f"triangle tests {tag}")
    sc.ctx.set_trace_mode("simple", 16)
    sc.ctx.set_coop(16)
    return g, o


def check_any(sc, fr, offs, refs, v, rays):
    import torch
    n = rays["ox"].shape[0]
    o = orc.trace_any(fr, offs, refs, v, rays, counters=True)
    rb = grt.RayBatch.from_numpy(rays)
    for mode, refill, coop in MODES:
        sc.ctx.set_trace_mode(mode, refill)
        sc.ctx.set_coop(coop)
        occ = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
        cnt = torch.full((2 * n,), -7, dtype=torch.int32, device="cuda:0")
        sc.trace_any(rb, occ, counters=cnt)
        gc = cnt.cpu().numpy().view(np.uint32).reshape(n, 2)
        tag = f"[{mode} param={refill} coop={coop}]"
        assert_same(occ.cpu().numpy(), o["occluded"], f"occluded {tag}")
        assert_same(gc[:, 0], o["cells"], f"cells visited {tag}")
        assert_same(gc[:, 1], o["tests"], f"triangle tests {tag}")
    sc.ctx.set_trace_mode("simple", 16)
    sc.ctx.set_coop(16)
    return o


def test_trace_mode_validation(gpu_ctx):
    for bad in (0, 1, 3, 32):
        with pytest.raises(grt.GrtError):
            gpu_ctx.set_trace_mode("split", bad)
    for bad in (0, 65):
        with pytest.raises(grt.GrtError):
            gpu_ctx.set_trace_mode("persistent", bad)
    with pytest.raises(grt.GrtError):
        grt.check(grt.lib().grt_set_trace_mode(gpu_ctx.h, 4, 16))
    gpu_ctx.set_trace_mode("simple", 16)


# ---------------------------------------------------------------------------
# Scan
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 5, 63, 64, 4095, 4096, 4097, 100003, 4096 * 1024 + 7])
def test_scan_exclusive(gpu_ctx, n):
    import torch
    x = rng(n).integers(0, 50, size=n, dtype=np.uint32)
    xd = torch.from_numpy(x.view(np.int32)).cuda()
    out = torch.empty(n + 1, dtype=torch.int32, device="cuda:0")
    gpu_ctx.debug_scan(xd, out)
    got = out.cpu().numpy().view(np.uint32)
    ref = np.zeros(n + 1, np.uint64)
    ref[1:] = np.cumsum(x, dtype=np.uint64)
    assert np.array_equal(got, ref.astype(np.uint32))


# ---------------------------------------------------------------------------
# Grid build
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("lam", [2.0, 5.0, 8.0])
def test_grid_build_c1(gpu_ctx, c1_scene, lam):
    v, n = c1_scene
    check_grid(gpu_ctx, v, n, lam)


def test_grid_build_rebuild_twice_same(gpu_ctx, c1_scene):
    v, n = c1_scene
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, n)
    sc.build_grid(5.0)
    a = sc.download_grid()
    sc.rebuild_grid_async(5.0)
    sc.check_grid()
    b = sc.download_grid()
    assert_same(a[1], b[1], "offsets")
    assert_same(a[2], b[2], "refs")


def _edge_scenes():
    r = rng(7)
    out = {}
    out["single"] = np.array([[0.1, 0.2, 0.3, 0.9, 0.25, 0.35, 0.4, 0.8, 0.6]], np.float32)
    planar = r.random((200, 9), dtype=np.float32)
    planar[:, 1::3] = 0.5
    out["planar_y"] = planar
    degen = r.random((300, 9), dtype=np.float32)
    degen[::7, 3:6] = degen[::7, 0:3]
    degen[::7, 6:9] = degen[::7, 0:3]
    out["degenerate"] = degen
    dup = np.repeat(r.random((50, 9), dtype=np.float32), 4, axis=0)
    out["duplicates"] = dup
    big = r.random((500, 9), dtype=np.float32) * 0.05 + 0.4
    big = np.vstack([big, np.array([[-1, -1, -1, 2, -1, -1, -1, 2, 2]], np.float32)])
    out["big_triangle"] = big
    # One hot cell holding > 16384 references (exercises the tiled global sort).
    hot = r.random((40000, 9), dtype=np.float32) * 1e-3 + 0.5
    spread = r.random((64, 9), dtype=np.float32)
    out["hot_cell"] = np.vstack([hot, spread]).astype(np.float32)
    return out


@pytest.mark.parametrize("name", ["single", "planar_y", "degenerate", "duplicates",
                                  "big_triangle", "hot_cell"])
def test_grid_build_edge(gpu_ctx, name):
    v = _edge_scenes()[name]
    n = np.zeros_like(v)
    check_grid(gpu_ctx, v, n, 5.0)


@pytest.mark.parametrize("name", ["degenerate", "duplicates", "big_triangle", "hot_cell"])
def test_trace_edge(gpu_ctx, name):
    """Traversal on the edge scenes.  Duplicates put exact t ties in one cell
    (the earliest reference must win, also in cooperative passes); the hot
    cell holds 40,000 references, so drained waves test it cooperatively."""
    v = _edge_scenes()[name]
    n = np.zeros_like(v)
    sc, fr, offs, refs = check_grid(gpu_ctx, v, n, 5.0)
    rays = _random_rays(20000, 21, inside=False)
    # aim half of the rays at the scene's centre so they cross its dense cells
    c = v.reshape(-1, 3).mean(axis=0)
    for k, ax in enumerate(("dx", "dy", "dz")):
        o = rays["o" + ax[1]]
        rays[ax][::2] = (c[k] - o[::2]).astype(np.float32)
    g, o = check_closest(sc, fr, offs, refs, v, rays)
    assert (o["id"] >= 0).sum() > 100
    check_any(sc, fr, offs, refs, v, rays)


@pytest.mark.parametrize("over", ["cells", "refs"])
def test_rebuild_over_capacity_is_safe(gpu_ctx, c1_scene, over):
    """An asynchronous rebuild that outgrows the reserved grid (more cells, or
    more references with the cells reserved) leaves a grid every trace treats
    as empty: traces already queued return misses, the check reports
    GRT_CAPACITY, later traces are refused, and a synchronous build recovers
    with the oracle's grid and hits."""
    import torch
    v, n = c1_scene
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, n)
    sc.build_grid(2.0)
    if over == "refs":
        sc.reserve_grid(int(orc.grid_frame(v, 40.0)["m"]) + 4096, 1)
    cam = c1_camera(96, 64)
    rb, hb = grt.RayBatch.empty(cam.width * cam.height), grt.HitBatch.empty(cam.width * cam.height)
    gpu_ctx.gen_primary(cam, rb)
    sc.rebuild_grid_async(40.0)
    hb.f["id"].fill_(7)
    sc.trace_closest(rb, hb, image_width=cam.width)
    torch.cuda.synchronize()
    assert (hb.numpy()["id"] == -1).all()
    with pytest.raises(grt.GrtError) as e:
        sc.check_grid()
    assert e.value.code == grt.GRT_CAPACITY
    with pytest.raises(grt.GrtError):
        sc.trace_closest(rb, hb, image_width=cam.width)
    sc.build_grid(5.0)
    fr = orc.grid_frame(v, 5.0)
    _, ooffs, orefs = orc.grid_build(v, fr)
    _, offs, refs = sc.download_grid()
    assert_same(offs, ooffs, "offsets after recovery")
    assert_same(refs, orefs, "refs after recovery")
    sc.trace_closest(rb, hb, image_width=cam.width)
    o = orc.trace_closest(fr, ooffs, orefs, v, rb.numpy())
    g = hb.numpy()
    for k in ("id", "t", "u", "v"):
        assert_same(g[k], o[k], f"hit.{k} after recovery")


@pytest.mark.parametrize("then", ["build", "reserve_rebuild"])
def test_rebuild_over_refs_then_build_without_check(gpu_ctx, c1_scene, then):
    """A rebuild that outgrows the reference capacity leaves the per-cell
    counts clean on the device: a synchronous build, or a reserve plus
    asynchronous rebuild, issued without grt_grid_check in between still
    returns the oracle's grid."""
    v, n = c1_scene
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, n)
    sc.build_grid(2.0)
    fr40 = orc.grid_frame(v, 40.0)
    sc.reserve_grid(int(fr40["m"]) + 4096, 1)
    sc.rebuild_grid_async(40.0)     # cells fit, references do not
    if then == "build":
        lam = 5.0
        sc.build_grid(lam)
    else:
        lam = 40.0
        _, o40, _ = orc.grid_build(v, fr40)
        sc.reserve_grid(int(fr40["m"]) + 4096, int(o40[-1]) + 4096)
        sc.rebuild_grid_async(lam)
        sc.check_grid()
    fr = orc.grid_frame(v, lam)
    _, ooffs, orefs = orc.grid_build(v, fr)
    info, offs, refs = sc.download_grid()
    assert_frame(info, fr)
    assert_same(offs, ooffs, f"offsets after the overflow ({then})")
    assert_same(refs, orefs, f"refs after the overflow ({then})")


@pytest.mark.parametrize("lam2", [0.0, 4.0])
def test_growing_reserve_drops_the_grid_until_rebuilt(gpu_ctx, c1_scene, lam2):
    """grt_grid_reserve that grows the buffers reallocates them without their
    contents: a trace before the next build is refused with GRT_NOT_READY
    (not run over uninitialised offsets), a reserve that fits leaves the grid
    usable, and the rebuild into the grown buffers traces like the oracle --
    one level, and two levels (lambda2 4)."""
    v, n = c1_scene
    sc = grt.Scene(gpu_ctx)
    try:
        sc.upload(v, n)
        sc.set_subgrid(lam2, 8)
        sc.build_grid(5.0)
        cam = c1_camera(96, 64)
        rays = orc.primary_rays(grt.camera_dict(cam))
        rb, hb = grt.RayBatch.from_numpy(rays), grt.HitBatch.empty(cam.width * cam.height)
        info = sc.grid_info()
        sc.reserve_grid(info["m"], info["r"])        # fits: the grid stays
        sc.trace_closest(rb, hb, image_width=cam.width)
        sc.reserve_grid(4 * info["m"] + 65536, 4 * info["r"] + 65536)   # grows
        for trace in (lambda: sc.trace_closest(rb, hb, image_width=cam.width),
                      lambda: sc.trace_any(rb, hb.f["id"], image_width=cam.width)):
            with pytest.raises(grt.GrtError) as e:
                trace()
            assert e.value.code == grt.GRT_NOT_READY
        sc.rebuild_grid_async(5.0)
        sc.check_grid()
        sc.trace_closest(rb, hb, image_width=cam.width)
        fr = orc.grid_frame(v, 5.0)
        _, offs, refs = orc.grid_build(v, fr)
        if lam2 > 0:
            assert sc.subgrid_info()["built"]
            g2 = orc.grid2_build(v, fr, offs, refs, 8, lam2, 16)
            o = orc.grid2_trace_closest(fr, offs, refs, g2, v, rays)
        else:
            o = orc.trace_closest(fr, offs, refs, v, rays)
        g = hb.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], o[k], f"hit.{k} after the grown reserve and rebuild")
    finally:
        sc.close()


def test_grid_build_stadium_small(gpu_ctx):
    v, n, _ = grt.gen_stadium(SEED_C2, 16, 2048, 16 * 2048 - 2)
    check_grid(gpu_ctx, v, n, 5.0)


def _huge_scene(kind):
    """A flat slab of small triangles plus triangles spanning more than 65,536
    cells at lambda 64: a ground pair (boxes overlap: one merged cell walk),
    planes at two heights (their union box is mostly empty) and five planes
    (more than are merged); the last two take the per-triangle path."""
    r = rng(11)
    p = r.random((20000, 3), dtype=np.float32)
    p[:, 1] *= 0.05
    small = np.tile(p, 3) + (r.random((20000, 9), dtype=np.float32) - 0.5) * 0.01
    def plane(y, flip=False):
        a = [[0, y, 0, 1, y, 0, 1, y, 1], [0, y, 0, 1, y, 1, 0, y, 1]]
        return np.array(a[1:] + a[:1] if flip else a, np.float32)
    if kind == "ground_pair":
        big = plane(0.0)
    elif kind == "two_levels":
        big = np.vstack([plane(0.0), plane(0.05)])
    else:
        big = np.vstack([plane(0.0), plane(0.0, True), plane(0.0)[:1]])
    return np.vstack([small, big]).astype(np.float32)


@pytest.mark.parametrize("kind", ["ground_pair", "two_levels", "five_planes"])
def test_grid_build_huge_triangles(gpu_ctx, kind):
    v = _huge_scene(kind)
    fr = orc.grid_frame(v, 64.0)
    assert fr["n"][0] * fr["n"][2] > 65536, fr["n"]
    check_grid(gpu_ctx, v, np.zeros_like(v), 64.0)


def test_grid_build_stadium_full(gpu_ctx):
    """Config 2 size: 2^20 triangles, 256 clusters + ground plane."""
    v, n, _ = grt.gen_stadium(SEED_C2, 256, 4096, (1 << 20) - 2)
    check_grid(gpu_ctx, v, n, 5.0)


# ---------------------------------------------------------------------------
# Rays
# ---------------------------------------------------------------------------
def test_gen_primary_parity(gpu_ctx):
    cam = c1_camera(640, 360)
    rb = grt.RayBatch.empty(640 * 360)
    gpu_ctx.gen_primary(cam, rb)
    g = rb.numpy()
    o = orc.primary_rays(grt.camera_dict(cam))
    for k in g:
        assert_same(g[k], o[k], f"primary.{k}")


def test_camera_host_parity():
    for eye, ctr, up, fov, w, h in [((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0, 1, 0), 60.0, 512, 512),
                                    ((0.3, 2.0, -3.0), (0.5, 0.2, 0.5), (0, 1, 0), 45.0, 1920, 1080),
                                    ((5, 5, 5), (0, 0, 0), (0, 0, 1), 30.0, 64, 48)]:
        a = grt.camera_dict(grt.camera_lookat(eye, ctr, up, fov, w, h))
        b = orc.camera_lookat(eye, ctr, up, fov, w, h)
        for k in ("pos", "base", "nu", "nv"):
            assert_same(np.float32(a[k]), np.float32(b[k]), k)


@pytest.mark.parametrize("image_width", [0, 512])
def test_closest_c1_primary(gpu_ctx, c1_scene, image_width):
    v, n = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, n, 5.0)
    rays = orc.primary_rays(grt.camera_dict(c1_camera()))
    g, o = check_closest(sc, fr, offs, refs, v, rays, image_width)
    assert (o["id"] >= 0).sum() > 10000


def _random_rays(n, seed, inside=True):
    r = rng(seed)
    o = r.random((n, 3), dtype=np.float32) if inside else \
        (r.random((n, 3), dtype=np.float32) * 4 - 1.5).astype(np.float32)
    d = r.normal(size=(n, 3)).astype(np.float32)
    # axis-aligned and zero components, including negative zero
    d[::11, 1:] = 0.0
    d[1::13, 0] = -0.0
    d[2::17, :2] = 0.0
    return {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
            "dz": d[:, 2], "tmin": np.full(n, 1e-4, np.float32),
            "tmax": np.where(np.arange(n) % 5 == 0, np.float32(0.3), np.float32(np.inf))}


def _degenerate_rays(n, seed):
    """Rays that break the usual assumptions, interleaved with ordinary ones
    so that they share waves: zero, NaN, infinite and tiny directions, NaN
    and far-away origins, empty and one-point intervals, negative tmin."""
    rays = _random_rays(n, seed, inside=True)
    f = np.float32
    cases = [((0.5, 0.5, 0.5), (0, 0, 0), 0, np.inf),
             ((0.5, 0.5, 0.5), (np.nan, 1, 0), 0, np.inf),
             ((np.nan, 0.5, 0.5), (1, 0, 0), 0, np.inf),
             ((0.2, 0.5, 0.5), (1, 0.1, 0), 0.5, 0.2),
             ((0.2, 0.5, 0.5), (1, 0.1, 0), 0.3, 0.3),
             ((1e30, 0.5, 0.5), (-1, 0, 0), 0, np.inf),
             ((0.5, 0.5, 0.5), (np.inf, 0, 0), 0, np.inf),
             ((0.5, 0.5, 0.5), (1, 1, 1), -1.0, np.inf),
             ((0.5, 0.5, 0.5), (1e-30, 1e-30, 1e-30), 0, np.inf),
             ((-0.5, 0.5, 0.5), (1, 0, 0), 0, np.nan)]
    for k, (o, d, t0, t1) in enumerate(cases):
        sl = slice(k, n, 7 * len(cases))
        for a, key in enumerate(("ox", "oy", "oz")):
            rays[key][sl] = f(o[a])
        for a, key in enumerate(("dx", "dy", "dz")):
            rays[key][sl] = f(d[a])
        rays["tmin"][sl] = f(t0)
        rays["tmax"][sl] = f(t1)
    return rays


def test_degenerate_rays(gpu_ctx, c1_scene):
    """Every schedule returns the oracle's results and counters on degenerate
    rays, and zero-length launches are accepted."""
    import torch
    v, n = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, n, 5.0)
    rays = _degenerate_rays(8192, 41)
    check_closest(sc, fr, offs, refs, v, rays)
    check_any(sc, fr, offs, refs, v, rays)
    empty = grt.RayBatch.empty(0)
    sc.trace_closest(empty, grt.HitBatch.empty(0))
    sc.trace_any(empty, torch.empty(0, dtype=torch.uint8, device="cuda:0"))
    torch.cuda.synchronize()


@pytest.mark.parametrize("inside", [True, False])
def test_closest_random_rays(gpu_ctx, c1_scene, inside):
    v, n = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, n, 5.0)
    check_closest(sc, fr, offs, refs, v, _random_rays(50000, 3 + inside, inside))


def test_any_random_rays(gpu_ctx, c1_scene):
    v, n = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, n, 5.0)
    check_any(sc, fr, offs, refs, v, _random_rays(50000, 11))


def _secondary_parity(gpu_ctx, sc, fr, offs, refs, v, nrm, rays, light):
    """Primary hits -> shadow, reflection and AO rays generated on both sides,
    then traced on both sides."""
    g, o = check_closest(sc, fr, offs, refs, v, rays)
    rb = grt.RayBatch.from_numpy(rays)
    hb = grt.HitBatch.from_numpy(g)
    n = rays["ox"].shape[0]
    # shadow
    sb = grt.RayBatch.empty(n)
    sc.gen_shadow(rb, hb, light, 1e-3, sb)
    so = orc.gen_shadow(rays, o, nrm, light, 1e-3)
    sg = sb.numpy()
    for k in sg:
        assert_same(sg[k], so[k], f"shadow.{k}")
    check_any(sc, fr, offs, refs, v, so)
    # reflection
    fb = grt.RayBatch.empty(n)
    sc.gen_reflection(rb, hb, 1e-3, fb)
    fo = orc.gen_reflection(rays, o, nrm, 1e-3)
    fg = fb.numpy()
    for k in fg:
        assert_same(fg[k], fo[k], f"reflection.{k}")
    check_closest(sc, fr, offs, refs, v, fo)
    # ambient occlusion
    spp = 4
    ab = grt.RayBatch.empty(n * spp)
    sc.gen_ao(rb, hb, spp, 1234, 1e-3, 0.05, ab)
    ao = orc.gen_ao(rays, o, nrm, spp, 1234, 1e-3, 0.05)
    ag = ab.numpy()
    for k in ag:
        assert_same(ag[k], ao[k], f"ao.{k}")
    check_any(sc, fr, offs, refs, v, ao)
    # headlight image
    import torch
    rgba = torch.empty(n, dtype=torch.int32, device="cuda:0")
    sc.shade_headlight(rb, hb, rgba)
    img = rgba.cpu().numpy().view(np.uint32)
    idx = np.nonzero(o["id"] >= 0)[0][:200]
    for i in idx:
        d = (rays["dx"][i], rays["dy"][i], rays["dz"][i])
        assert img[i] == orc.headlight(d, nrm[o["id"][i]], o["id"][i], o["u"][i], o["v"][i])
    assert np.all(img[o["id"] < 0] == 0xFFFFFFFF)


def test_secondary_c1(gpu_ctx, c1_scene):
    v, nrm = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, nrm, 5.0)
    rays = orc.primary_rays(grt.camera_dict(c1_camera(256, 256)))
    _secondary_parity(gpu_ctx, sc, fr, offs, refs, v, nrm, rays, (1000.0, 1000.0, 1000.0))


def test_secondary_stadium(gpu_ctx):
    v, nrm, _ = grt.gen_stadium(SEED_C2, 16, 2048, 16 * 2048 - 2)
    sc, fr, offs, refs = check_grid(gpu_ctx, v, nrm, 5.0)
    cam = grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, 192, 108)
    rays = orc.primary_rays(grt.camera_dict(cam))
    _secondary_parity(gpu_ctx, sc, fr, offs, refs, v, nrm, rays, (1000.0, 1000.0, 1000.0))


def test_uploaded_grid_matches_built(gpu_ctx, c1_scene):
    v, n = c1_scene
    fr = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, fr)
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, n)
    sc.upload_grid(fr, offs, refs)
    rays = orc.primary_rays(grt.camera_dict(c1_camera(128, 128)))
    check_closest(sc, fr, offs, refs, v, rays)


def test_transform_parity(gpu_ctx):
    v, nrm, centers = grt.gen_stadium(SEED_C2, 16, 2048, 16 * 2048 - 2)
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, nrm)
    for frame in (1, 2):
        xf = grt.gen_cluster_transforms(SEED_C2, frame, centers, 0.05)
        sc.transform_clusters(xf, 2048, 16 * 2048 - 2)
        ov, on = orc.transform(v, nrm, xf, 2048, 16 * 2048 - 2)
        gv, gn = sc.download_geometry()
        assert_same(gv.ravel(), ov.ravel(), "transformed vertices")
        assert_same(gn.ravel(), on.ravel(), "transformed normals")
        sc.build_grid(5.0)
        info, offs, refs = sc.download_grid()
        fr = orc.grid_frame(ov, 5.0)
        _, ooffs, orefs = orc.grid_build(ov, fr)
        assert_frame(info, fr)
        assert_same(offs, ooffs, "offsets after transform")
        assert_same(refs, orefs, "refs after transform")
        cam = grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, 96, 54)
        rays = orc.primary_rays(grt.camera_dict(cam))
        check_closest(sc, fr, ooffs, orefs, ov, rays)


# ---------------------------------------------------------------------------
# Phong frame: primary + shadow shaded, one mirror bounce + its shadow added
# ---------------------------------------------------------------------------
def _phong_frame_gpu(sc, cam, light, params, diffuse):
    import torch
    n = cam.width * cam.height
    rb, hb = grt.RayBatch.empty(n), grt.HitBatch.empty(n)
    sb, fb, fhb = grt.RayBatch.empty(n), grt.RayBatch.empty(n), grt.HitBatch.empty(n)
    s2 = grt.RayBatch.empty(n)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    occ2 = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    rgb = torch.full((n, 3), -1.0, dtype=torch.float32, device="cuda:0")
    rgba = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    kd = torch.from_numpy(diffuse).cuda() if diffuse is not None else None
    sc.ctx.gen_primary(cam, rb)
    sc.trace_closest(rb, hb, image_width=cam.width)
    sc.gen_shadow(rb, hb, light, 1e-3, sb)
    sc.trace_any(sb, occ)
    sc.shade_phong(rb, hb, rgb, occ, 0, params, kd)
    first = rgb.cpu().numpy().copy()
    sc.gen_reflection(rb, hb, 1e-3, fb)
    sc.trace_closest(fb, fhb)
    sc.gen_shadow(fb, fhb, light, 1e-3, s2)
    sc.trace_any(s2, occ2)
    sc.shade_phong(fb, fhb, rgb, occ2, 1, params, kd, rgba)
    return first, rgb.cpu().numpy(), rgba.cpu().numpy().view(np.uint32)


def _phong_frame_oracle(fr, offs, refs, v, nrm, cam, light, params, diffuse):
    rays = orc.primary_rays(grt.camera_dict(cam))
    o = orc.trace_closest(fr, offs, refs, v, rays)
    so = orc.gen_shadow(rays, o, nrm, light, 1e-3)
    occ = orc.trace_any(fr, offs, refs, v, so)["occluded"]
    first, _ = orc.shade_phong(rays, o, nrm, occ, params, 0, None, diffuse)
    fo = orc.gen_reflection(rays, o, nrm, 1e-3)
    fho = orc.trace_closest(fr, offs, refs, v, fo)
    s2 = orc.gen_shadow(fo, fho, nrm, light, 1e-3)
    occ2 = orc.trace_any(fr, offs, refs, v, s2)["occluded"]
    rgb, img = orc.shade_phong(fo, fho, nrm, occ2, params, 1, first, diffuse)
    return first, rgb, img, o, occ


@pytest.mark.parametrize("colours", [False, True])
def test_phong_frame_c1(gpu_ctx, c1_scene, colours):
    v, nrm = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, nrm, 5.0)
    light = (3.0, 4.0, -2.0)
    params = {"light": light}
    diffuse = None
    if colours:
        diffuse = rng(77).uniform(0.05, 1.0, size=(v.shape[0], 3)).astype(np.float32)
        params.update({"spec_exp": 9, "ambient": (0.05, 0.02, 0.0), "shadow_coeff": 0.3})
    cam = c1_camera(192, 160)
    g_first, g_rgb, g_img = _phong_frame_gpu(sc, cam, light, grt.phong(params), diffuse)
    o_first, o_rgb, o_img, o, occ = _phong_frame_oracle(fr, offs, refs, v, nrm, cam, light,
                                                        params, diffuse)
    assert_same(g_first, o_first, "primary Phong rgb")
    assert_same(g_rgb, o_rgb, "rgb with the reflection added")
    assert_same(g_img, o_img, "RGBA8 frame")
    # the frame exercises every branch: misses, lit and shadowed hits, and
    # reflections that add colour
    hit = o["id"] >= 0
    assert np.any(~hit) and np.any(hit & (occ == 1)) and np.any(hit & (occ == 0))
    assert np.any(o_rgb[hit] != o_first[hit])
    assert np.all(o_rgb[~hit] == 1.0)


def test_mesh_scene_parity(gpu_ctx, tmp_path):
    """A scene assembled on the host (procedural spheres and cubes through
    transforms, plus an OBJ quad with a degenerate face) goes through
    grt_scene_upload_mesh, the device grid build and the traversal, against
    the oracle on the mesh's own arrays."""
    m = grt.Mesh()
    for k in range(5):
        s = 0.05 + 0.02 * k
        m.set_transform([s, 0, 0, 0.2 * k, 0, s, 0, 0.5, 0, 0, s, 0.3 * (k % 2)])
        m.add_sphere(24, 20)
        m.set_transform([s, 0, 0, 0.2 * k, 0, 2 * s, 0, 0.2, 0, 0, s, 0.6])
        m.add_cube()
    (tmp_path / "q.obj").write_text("v -1 -0.2 -1\nv 2 -0.2 -1\nv 2 -0.2 2\nv -1 -0.2 2\n"
                                    "vn 0 1 0\nf 1//1 2//1 3//1 4//1\nf 1 1 2\n")
    m.load_obj(tmp_path / "q.obj", drop_degenerate=True)
    v, nrm, _, _ = m.arrays()
    sc = grt.Scene(gpu_ctx)
    sc.upload_mesh(m)
    sc.build_grid(5.0)
    fr = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, fr)
    assert_frame(sc.grid_info(), fr)
    _, go, gr = sc.download_grid()
    assert_same(go, offs, "mesh grid offsets")
    assert_same(gr, refs, "mesh grid refs")
    cam = grt.camera_lookat((0.4, 0.8, -2.0), (0.4, 0.4, 0.3), (0, 1, 0), 50.0, 160, 120)
    rays = orc.primary_rays(grt.camera_dict(cam))
    g, o = check_closest(sc, fr, offs, refs, v, rays, image_width=cam.width)
    assert np.count_nonzero(o["id"] >= 0) > rays["ox"].shape[0] // 4
    sh = orc.gen_shadow(rays, o, nrm, (5.0, 9.0, -4.0), 1e-3)
    check_any(sc, fr, offs, refs, v, sh)


@pytest.mark.parametrize("layout", ["even_spaced", "mixed", "teapots125"])
def test_rtdb_layout_scene_parity(gpu_ctx, tmp_path, layout):
    """rtdb's scene layouts (rtdb.cpp:215-287) assembled on the host -- with a
    model read from an OBJ in place of the teapot -- through
    grt_scene_upload_mesh, the device grid build and the traversal, against
    the oracle on the mesh's own arrays."""
    src = grt.Mesh()
    src.set_transform([0.5, 0, 0, 0.25, 0, 0.25, 0, 0.5, 0, 0, 0.75, 0.125])
    src.add_cube()
    src.set_transform([0.3, 0, 0, 0.5, 0, 0.3, 0, 0.2, 0, 0, 0.3, 0.5])
    src.add_sphere(12, 10)
    v0 = src.arrays()[0].reshape(-1, 3, 3)
    lines = ["v %.9g %.9g %.9g" % tuple(p) for tri in v0 for p in tri]
    lines += ["f %d %d %d" % (3 * k + 1, 3 * k + 2, 3 * k + 3) for k in range(len(v0))]
    (tmp_path / "model.obj").write_text("\n".join(lines) + "\n")
    model = grt.Mesh()
    model.load_obj(tmp_path / "model.obj", drop_degenerate=True)
    m = grt.Mesh()
    if layout == "even_spaced":
        m.add_even_spaced()
        eye, at = (1.5, 1.5, -6.0), (1.5, 1.5, 1.5)
    elif layout == "mixed":
        m.add_mixed(model)
        eye, at = (14.0, 12.0, -25.0), (12.0, 3.0, 12.0)
    else:
        m.add_teapots125(model)
        eye, at = (-30.0, 25.0, -30.0), (25.0, 25.0, 25.0)
    v, nrm, _, _ = m.arrays()
    sc = grt.Scene(gpu_ctx)
    sc.upload_mesh(m)
    sc.build_grid(5.0)
    fr = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, fr)
    assert_frame(sc.grid_info(), fr)
    _, go, gr = sc.download_grid()
    assert_same(go, offs, f"{layout} grid offsets")
    assert_same(gr, refs, f"{layout} grid refs")
    cam = grt.camera_lookat(eye, at, (0, 1, 0), 50.0, 200, 150)
    rays = orc.primary_rays(grt.camera_dict(cam))
    g, o = check_closest(sc, fr, offs, refs, v, rays, image_width=cam.width)
    assert np.count_nonzero(o["id"] >= 0) > 500
    sh = orc.gen_shadow(rays, o, nrm, (1000.0, 1000.0, 1000.0), 1e-3)
    check_any(sc, fr, offs, refs, v, sh)


# ---------------------------------------------------------------------------
# Instances: rays through each instance's inverse matrix, nearest hit kept
# ---------------------------------------------------------------------------
def _oracle_instances(items, rays):
    """The oracle restatement of IAccStruct::traceNearestInstance on the
    oracle's own traversal: items are (frame, offsets, refs, verts, inverse)."""
    n = rays["ox"].shape[0]
    best = {"id": np.full(n, -1, np.int32), "t": np.full(n, np.inf, np.float32),
            "u": np.zeros(n, np.float32), "v": np.zeros(n, np.float32)}
    inst = np.full(n, -1, np.int32)
    o = [rays[k] for k in ("ox", "oy", "oz")]
    d = [rays[k] for k in ("dx", "dy", "dz")]
    for k, (fr, offs, refs, v, a) in enumerate(items):
        a = np.asarray(a, np.float32)
        loc = {}
        for r, name in enumerate("xyz"):
            loc["o" + name] = ((a[4 * r] * o[0] + a[4 * r + 1] * o[1]) + a[4 * r + 2] * o[2]) + a[4 * r + 3]
            loc["d" + name] = (a[4 * r] * d[0] + a[4 * r + 1] * d[1]) + a[4 * r + 2] * d[2]
        loc["tmin"] = rays["tmin"].copy()
        tmax = rays["tmax"].copy()
        fin = best["t"] < np.inf
        tmax[fin] = np.minimum(np.nextafter(best["t"][fin], np.float32(-np.inf)), tmax[fin])
        loc["tmax"] = tmax
        h = orc.trace_closest(fr, offs, refs, v, loc)
        m = h["id"] >= 0
        for key in best:
            best[key][m] = h[key][m]
        inst[m] = k
    return best, inst


def test_instances_parity(gpu_ctx, c1_scene):
    import torch
    v, nrm = c1_scene
    sc, fr, offs, refs = check_grid(gpu_ctx, v, nrm, 5.0)
    cube = grt.Mesh()
    cube.add_cube()
    cv, cn, _, _ = cube.arrays()
    csc, cfr, coffs, crefs = check_grid(gpu_ctx, cv, cn, 5.0)
    c, s = math.cos(0.5), math.sin(0.5)
    placed = [
        (sc, fr, offs, refs, v, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]),
        (sc, fr, offs, refs, v, [0.8 * c, 0, 0.8 * s, 0.3, 0, 0.8, 0, 0.05, -0.8 * s, 0, 0.8 * c, 0.5]),
        (csc, cfr, coffs, crefs, cv, [2.0, 0, 0, 0.5, 0, 0.1, 0, -0.1, 0, 0, 2.0, 0.5]),
        (sc, fr, offs, refs, v, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]),   # ties: instance 0 wins
    ]
    cam = grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, 128, 96)
    rays = orc.primary_rays(grt.camera_dict(cam))
    items = [(f, o, r, vv, grt.affine_inverse(xf)) for (_, f, o, r, vv, xf) in placed]
    want, want_inst = _oracle_instances(items, rays)
    assert set(np.unique(want_inst)) >= {-1, 0, 1, 2} and not np.any(want_inst == 3)
    rb = grt.RayBatch.from_numpy(rays)
    for mode, param in (("simple", 16), ("pooled", 16), ("split", 4)):
        gpu_ctx.set_trace_mode(mode, param)
        hb = grt.HitBatch.empty(rb.n)
        inst = torch.full((rb.n,), 77, dtype=torch.int32, device="cuda:0")
        gpu_ctx.trace_instances([(p[0], p[5]) for p in placed], rb, hb, inst, image_width=cam.width)
        g = hb.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], want[k], f"instances hit.{k} [{mode}]")
        assert_same(inst.cpu().numpy(), want_inst, f"instance index [{mode}]")
    gpu_ctx.set_trace_mode("simple", 16)
    with pytest.raises(grt.GrtError):
        gpu_ctx.trace_instances([(sc, [0] * 12)], rb, grt.HitBatch.empty(rb.n),
                                torch.empty(rb.n, dtype=torch.int32, device="cuda:0"))


# ---------------------------------------------------------------------------
# Ray lists: the listed (active) rays alone give the full stage's results
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scene", ["c1", "stadium"])
def test_ray_lists_match_full_stages(gpu_ctx, c1_scene, scene):
    import torch
    if scene == "c1":
        v, nrm = c1_scene
        cam = c1_camera(200, 136)
    else:
        v, nrm, _ = grt.gen_stadium(SEED_C2, 16, 2048, 16 * 2048 - 2)
        cam = grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, 200, 136)
    sc = grt.Scene(gpu_ctx)
    sc.upload(v, nrm)
    sc.build_grid(5.0)
    n, W = cam.width * cam.height, cam.width
    light = (3.0, 4.0, -2.0)
    prim, hits = grt.RayBatch.empty(n), grt.HitBatch.empty(n)
    gpu_ctx.gen_primary(cam, prim)
    sc.trace_closest(prim, hits, image_width=W)
    # full stages
    sh, refl = grt.RayBatch.empty(n), grt.RayBatch.empty(n)
    occ, rhits = torch.empty(n, dtype=torch.uint8, device="cuda:0"), grt.HitBatch.empty(n)
    sc.gen_shadow(prim, hits, light, 1e-3, sh)
    sc.gen_reflection(prim, hits, 1e-3, refl)
    cnt_full = torch.zeros(2 * n, dtype=torch.int32, device="cuda:0")
    sc.trace_any(sh, occ, image_width=W, counters=cnt_full)
    rcnt_full = torch.zeros(2 * n, dtype=torch.int32, device="cuda:0")
    sc.trace_closest(refl, rhits, image_width=W, counters=rcnt_full)
    want_occ, want_r = occ.cpu().numpy(), rhits.numpy()
    sh_active = (sh.f["tmin"] <= sh.f["tmax"]).cpu().numpy()
    rf_active = (refl.f["tmin"] <= refl.f["tmax"]).cpu().numpy()
    assert 0 < sh_active.sum() < n and 0 < rf_active.sum() < n
    for mode, param in (("simple", 16), ("pooled", 16), ("persistent", 16), ("split", 4)):
        gpu_ctx.set_trace_mode(mode, param)
        lst = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
        count = torch.full((1,), 12345, dtype=torch.int32, device="cuda:0")
        sh2 = grt.RayBatch.empty(n)
        occ2 = torch.full((n,), 9, dtype=torch.uint8, device="cuda:0")
        sc.gen_shadow(prim, hits, light, 1e-3, sh2)
        gpu_ctx.compact_rays(sh2, W, lst, count, occluded=occ2)
        c = int(count.item())
        assert c == int(sh_active.sum())
        # the list is the active rays in 8x8-tile order
        slots = np.arange(((W + 7) // 8) * ((cam.height + 7) // 8) * 64)
        tile, lane = slots // 64, slots % 64
        px = (tile % ((W + 7) // 8)) * 8 + lane % 8
        py = (tile // ((W + 7) // 8)) * 8 + lane // 8
        keep = (px < W) & (py < cam.height)
        order = (py * W + px)[keep]
        assert np.array_equal(lst[:c].cpu().numpy(), order[sh_active[order]])
        for k in sh.f:
            assert_same(sh2.f[k].cpu().numpy(), sh.f[k].cpu().numpy(), f"shadow list ray {k}")
        cnt = torch.zeros(2 * n, dtype=torch.int32, device="cuda:0")
        sc.trace_any_list(sh2, occ2, lst, count, counters=cnt)
        assert_same(occ2.cpu().numpy(), want_occ, f"listed shadow rays [{mode}]")
        assert_same(cnt.cpu().numpy(), cnt_full.cpu().numpy(), f"listed shadow counters [{mode}]")
        rf2, rh2 = grt.RayBatch.empty(n), grt.HitBatch.empty(n)
        rh2.f["id"].fill_(-9)
        sc.gen_reflection(prim, hits, 1e-3, rf2)
        gpu_ctx.compact_rays(rf2, W, lst, count, miss_hits=rh2)
        assert int(count.item()) == int(rf_active.sum())
        cnt = torch.zeros(2 * n, dtype=torch.int32, device="cuda:0")
        sc.trace_closest_list(rf2, rh2, lst, count, counters=cnt)
        g = rh2.numpy()
        for k in ("id", "t", "u", "v"):
            assert_same(g[k], want_r[k], f"listed bounce hit.{k} [{mode}]")
        assert_same(cnt.cpu().numpy(), rcnt_full.cpu().numpy(), f"listed bounce counters [{mode}]")
    gpu_ctx.set_trace_mode("simple", 16)
