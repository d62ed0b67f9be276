"""GPU parity on the batching, serving and ambient-occlusion workloads at their
own sizes.

  batch-64   64 seeded sphere scenes of 2k..500k triangles (5.8 M in all),
             square images of 64..1024 pixels a side, depth 1..3, every
             level of every job traced in one batched launch per level
  serve-16   16 resident sphere scenes of 10k..4M triangles (14.9 M in all),
             requests with Poisson arrivals so that new requests arrive while
             others are on the GPU
  ao-16M     2^24 triangles, its grid rebuilt on the device, 8 cosine-
             weighted ambient-occlusion rays per pixel at 1024 x 512

Bar: bit-exact.  Every grid equals the oracle's; every batched item equals the
same item traced alone and the oracle; every served request equals the same
request served alone; the AO frame's primary hits, AO rays and occlusion flags
equal the oracle's.
"""
import numpy as np
import pytest
import torch

import grt
import pyoracle as orc
from grt import perf
from grt import workloads as W

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


def _oracle_grid_equal(sc, v, lam, tag):
    info, offs, refs = sc.download_grid()
    fr = orc.grid_frame(v, lam)
    _, ooffs, orefs = orc.grid_build(v, fr)
    assert list(info["n"]) == list(fr["n"]), tag
    assert_same(offs, ooffs, f"offsets {tag}")
    assert_same(refs, orefs, f"refs {tag}")
    return fr, ooffs, orefs


# ---------------------------------------------------------------------------
# batch-64
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode,param", [("pooled", 16), ("split", 4)])
def test_batch64_levels_match_alone_and_oracle(gpu_ctx, mode, param):
    """The 64 jobs of the batching workload, one batched launch per depth
    level (the autotuner's lambda 2; pooled one ray per lane, and 4 lanes per
    ray).  Every item of every level equals the same item traced alone (one
    ray per lane), and 512 seeded rays per item equal the oracle."""
    ctx = gpu_ctx
    lam = 2.0
    jobs = W.batch_jobs()
    assert len(jobs) == 64
    bj = perf.BatchedJobs(ctx, jobs, lam)
    grids = []
    try:
        for i, j in enumerate(bj.jobs):
            v, _ = W.sphere_scene(W.SEED_BATCH + 1000 + i, jobs[i][0])
            grids.append((v,) + _oracle_grid_equal(j["scene"], v, lam, f"job {i}"))
        assert sum(g[0].shape[0] for g in grids) > 5_000_000
        rng = np.random.default_rng(0x6772)
        checked = 0
        for lvl in range(bj.levels):
            for j in bj.jobs:
                if lvl <= j["depth"]:
                    bj._gen(j, lvl)
            ctx.set_trace_mode(mode, param)
            bj.batches[lvl].trace()
            torch.cuda.synchronize()
            ctx.set_trace_mode("simple", 16)
            for i, j in enumerate(bj.jobs):
                v, fr, offs, refs = grids[i]
                for (sc, rays, hits, occ, width) in bj._items(j, lvl):
                    tag = f"[{mode} {param}] level {lvl} job {i}"
                    rh = rays.numpy()
                    pick = rng.choice(rays.n, min(512, rays.n), replace=False)
                    sub = {k: a[pick] for k, a in rh.items()}
                    if hits is not None:
                        got = hits.numpy()
                        alone = grt.HitBatch.empty(rays.n)
                        sc.trace_closest(rays, alone, image_width=width)
                        a = alone.numpy()
                        o = orc.trace_closest(fr, offs, refs, v, sub)
                        for k in ("id", "t", "u", "v"):
                            assert_same(got[k], a[k], f"hit.{k} batched vs alone {tag}")
                            assert_same(got[k][pick], o[k], f"hit.{k} vs oracle {tag}")
                    else:
                        got = occ.cpu().numpy()
                        alone = torch.empty(rays.n, dtype=torch.uint8, device="cuda:0")
                        sc.trace_any(rays, alone, image_width=width)
                        assert_same(got, alone.cpu().numpy(), f"occluded batched vs alone {tag}")
                        o = orc.trace_any(fr, offs, refs, v, sub)
                        assert_same(got[pick], o["occluded"], f"occluded vs oracle {tag}")
                    checked += 1
        assert checked >= 64 * 2
    finally:
        ctx.set_trace_mode("simple", 16)
        for b in bj.batches:
            b.close()
        for j in bj.jobs:
            j["scene"].close()


# ---------------------------------------------------------------------------
# serve-16
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def serve_scenes(gpu_ctx):
    scenes, tris = perf.build_serve_scenes(gpu_ctx)
    assert len(scenes) == 16 and tris > 10_000_000
    yield scenes
    for sc in scenes:
        sc.close()


def _serve(ctx, scenes, spec, batched, budget=2 << 30):
    outs, reqs = [], []
    for (s, w, h, d, t) in spec:
        cam = W.camera(w, h)
        hb = grt.HitBatch.empty(w * h)
        hb.f["id"].fill_(-5)
        occ = torch.full((w * h,), 5, dtype=torch.uint8, device="cuda:0")
        outs.append((hb, occ, cam, s))
        reqs.append((scenes[s], cam, d, t, (hb, occ)))
    torch.cuda.synchronize()
    res = grt.serve_run(ctx, reqs, batched=batched, memory_budget=budget, max_active=64)
    torch.cuda.synchronize()
    return outs, res


def test_serve16_staggered_arrivals_same_alone_and_in_mix(gpu_ctx, serve_scenes):
    """40 requests of the serving workload with Poisson arrivals at 300 per
    second: requests are admitted while earlier ones are on the GPU (checked
    from the admission and completion times), and each request's primary hits
    and shadow flags, and its active-ray count over all its steps, equal the
    same request served alone.  Three requests are checked against the
    oracle on a seeded pixel sample."""
    ctx = gpu_ctx
    spec = W.serve_requests(40, 300.0, len(serve_scenes), seed=W.SEED_SERVE + 11)
    assert len({s for (s, *_rest) in spec}) >= 10
    mix, (arrival, start, finish, rays, wall) = _serve(ctx, serve_scenes, spec, True)
    assert np.all(finish >= start) and np.all(start >= arrival - 1e-3)
    # some request was admitted while another was in flight
    overlap = sum(1 for a in range(len(spec)) for b in range(len(spec))
                  if a != b and start[a] < start[b] < finish[a])
    assert overlap > 0, "no request arrived while another was on the GPU"
    for k, req in enumerate(spec):
        alone, res = _serve(ctx, serve_scenes, [req[:4] + (0.0,)], False)
        assert res[3][0] == rays[k], f"request {k}: active rays alone vs in the mix"
        a, b = mix[k][0].numpy(), alone[0][0].numpy()
        for f in ("id", "t", "u", "v"):
            assert_same(a[f], b[f], f"request {k} primary hit.{f}")
        assert_same(mix[k][1].cpu().numpy(), alone[0][1].cpu().numpy(), f"request {k} shadow")
    rng = np.random.default_rng(5)
    for k in rng.choice(len(spec), 3, replace=False):
        hb, occ, cam, s = mix[k]
        v, nrm = W.sphere_scene(W.SEED_SERVE + 100 + s, W.serve_scene_sizes()[s])
        fr, offs, refs = _oracle_grid_equal(serve_scenes[s], v, 8.0, f"serve scene {s}")
        rays_all = orc.primary_rays(grt.camera_dict(cam))
        pick = rng.choice(len(rays_all["ox"]), 4096, replace=False)
        sub = {key: a[pick] for key, a in rays_all.items()}
        o = orc.trace_closest(fr, offs, refs, v, sub)
        g = hb.numpy()
        for f in ("id", "t", "u", "v"):
            assert_same(g[f][pick], o[f], f"request {k} hit.{f} vs oracle")
        sh = orc.gen_shadow(sub, o, nrm, W.LIGHT, 1e-4)
        assert_same(occ.cpu().numpy()[pick], orc.trace_any(fr, offs, refs, v, sh)["occluded"],
                    f"request {k} shadow vs oracle")


# ---------------------------------------------------------------------------
# ao-16M
# ---------------------------------------------------------------------------
def test_ao16m_grid_rebuild_and_ao_rays(gpu_ctx):
    """2^24 triangles at the AO frame's tuned lambda 4: the per-frame device
    rebuild equals the oracle's grid (6.7e7 cells, 8.4e7 references); the
    1024 x 512 primary hits, the 8 AO rays per pixel (tmax = 0.05 x scene
    extent) and all 4.2 M occlusion flags and counters equal the oracle's."""
    ctx = gpu_ctx
    lam, spp, width, height = 4.0, 8, 1024, 512
    v, nrm = W.ao_scene()
    assert v.shape[0] == 1 << 24
    sc = grt.Scene(ctx)
    try:
        sc.upload(v, nrm)
        sc.build_grid(lam)
        info = sc.grid_info()
        sc.reserve_grid(int(info["m"] * 1.1) + 4096, int(info["r"] * 1.1) + 4096)
        sc.rebuild_grid_async(lam)
        sc.check_grid()
        fr, offs, refs = _oracle_grid_equal(sc, v, lam, "ao-16M")
        assert fr["m"] > 60_000_000 and len(refs) > 80_000_000
        extent = float(np.max(np.asarray(fr["bmax"]) - np.asarray(fr["bmin"])))
        cam = W.ao_camera(width, height)
        n = width * height
        prim, hits = grt.RayBatch.empty(n), grt.HitBatch.empty(n)
        ctx.gen_primary(cam, prim)
        sc.trace_closest(prim, hits, image_width=width)
        ao = grt.RayBatch.empty(n * spp)
        sc.gen_ao(prim, hits, spp, W.SEED_AO, 1e-4, 0.05 * extent, ao)
        orays = orc.primary_rays(grt.camera_dict(cam))
        oh = orc.trace_closest(fr, offs, refs, v, orays)
        g = hits.numpy()
        for f in ("id", "t", "u", "v"):
            assert_same(g[f], oh[f], f"primary hit.{f}")
        assert (oh["id"] >= 0).mean() > 0.99
        oao = orc.gen_ao(orays, oh, nrm, spp, W.SEED_AO, 1e-4, 0.05 * extent)
        ga = ao.numpy()
        for k in ga:
            assert_same(ga[k], oao[k], f"ao rays.{k}")
        del ga
        oo = orc.trace_any(fr, offs, refs, v, oao, counters=True)
        for mode, param in (("simple", 16), ("pooled", 16), ("persistent", 16)):
            ctx.set_trace_mode(mode, param)
            occ = torch.full((n * spp,), 9, dtype=torch.uint8, device="cuda:0")
            cnt = torch.full((2 * n * spp,), -7, dtype=torch.int32, device="cuda:0")
            sc.trace_any(ao, occ, counters=cnt)
            gc = cnt.cpu().numpy().view(np.uint32).reshape(-1, 2)
            assert_same(occ.cpu().numpy(), oo["occluded"], f"ao occluded [{mode}]")
            assert_same(gc[:, 0], oo["cells"], f"ao cells [{mode}]")
            assert_same(gc[:, 1], oo["tests"], f"ao tests [{mode}]")
        assert 0.05 < oo["occluded"].mean() < 0.95
    finally:
        ctx.set_trace_mode("simple", 16)
        sc.close()
