"""Measurements of the batching, serving and ambient-occlusion workloads
(see grt.workloads for their definitions).  Every timed region starts with
scenes, grids and buffers resident in device memory.
"""
from __future__ import annotations

import time

import numpy as np

from . import (Batch, Context, Graph, HitBatch, KdTree, RayBatch, Scene, camera_lookat,
               gen_cluster_transforms, gen_stadium, serve_run, stream_create)
from . import workloads as W
from .autotune import Autotuner


def _torch():
    import torch
    return torch


# ---------------------------------------------------------------------------
# Serving: continuous batching vs the plain path
# ---------------------------------------------------------------------------
def build_serve_scenes(ctx: Context, lam: float = 8.0, n_scenes: int = 16, max_tris=None):
    scenes, tris = [], 0
    for i, nt in enumerate(W.serve_scene_sizes(n=n_scenes)):
        if max_tris:
            nt = min(nt, max_tris)
        v, nrm = W.sphere_scene(W.SEED_SERVE + 100 + i, nt)
        sc = Scene(ctx)
        sc.upload(v, nrm)
        sc.build_grid(lam)
        scenes.append(sc)
        tris += v.shape[0]
    return scenes, tris


def _requests(scenes, spec):
    return [(scenes[s], W.camera(w, h), d, t) for (s, w, h, d, t) in spec]


def _summary(arrival, start, finish, rays, wall):
    lat = finish - arrival
    return {"p50_ms": round(W.percentile_ms(lat, 50), 3),
            "p99_ms": round(W.percentile_ms(lat, 99), 3),
            "mean_ms": round(float(lat.mean() * 1e3), 3),
            "mrays_s": round(float(rays.sum()) / wall / 1e6, 2),
            "wall_s": round(wall, 3), "requests": int(len(lat))}


def serve_vs_plain(ctx: Context, scenes, n_requests: int = 120,
                   loads=(0.5, 0.8, 0.95, 1.0, 1.25, 1.5, 2.0, 2.25, 2.5), n_open: int = 400,
                   budget: int = 8 << 30,
                   max_active: int = 64, slo_ms: float = 50.0):
    """Serving (continuous batching) against the plain path on the mixed
    request stream.

    Saturation: n_requests all arriving at t = 0, for both paths -> each
    path's capacity.  Open loop: n_open Poisson arrivals at each `load` x the
    plain path's capacity (requests/s), the same stream for both paths ->
    p50/p99 latency per load, the p99 ratio plain/serve, and the highest load
    of the sweep at which each path keeps p99 within slo_ms.  The headline
    p99 headroom is the ratio at load 1.0 (arrivals at the plain path's own
    capacity), or at the highest load of the sweep when 1.0 is not in it."""
    torch = _torch()
    ctx.set_coop(16)
    warm = _requests(scenes, W.serve_requests(6, 0.0, len(scenes), seed=W.SEED_SERVE + 7))
    serve_run(ctx, warm, True, budget, max_active)
    serve_run(ctx, warm, False, budget, max_active)
    torch.cuda.synchronize()
    closed = _requests(scenes, W.serve_requests(n_requests, 0.0, len(scenes)))
    # Each path runs with the traversal schedule that saturates it best (the
    # batched path runs the split schedule as one ray per lane, so it tries
    # the first two only).
    modes = {"simple": 16, "pooled": 16, "split": 4}
    sat = {True: {}, False: {}}
    for mode, param in modes.items():
        ctx.set_trace_mode(mode, param)
        for batched in (False, True):
            if batched and mode == "split":
                continue
            sat[batched][mode] = _summary(*serve_run(ctx, closed, batched, budget, max_active))
    mode_of = {b: max(sat[b], key=lambda m: sat[b][m]["mrays_s"]) for b in sat}
    plain_sat, serve_sat = sat[False][mode_of[False]], sat[True][mode_of[True]]

    def run(reqs, batched):
        ctx.set_trace_mode(mode_of[batched], modes[mode_of[batched]])
        return _summary(*serve_run(ctx, reqs, batched, budget, max_active))

    capacity = n_requests / plain_sat["wall_s"]
    sweep = []
    for load in loads:
        rate = load * capacity
        opened = _requests(scenes, W.serve_requests(n_open, rate, len(scenes)))
        plain = run(opened, False)
        served = run(opened, True)
        sweep.append({"load_vs_plain_capacity": load, "arrival_rate_hz": round(rate, 2),
                      "serve": served, "plain": plain,
                      "p99_ratio": round(plain["p99_ms"] / max(served["p99_ms"], 1e-9), 2)})

    def max_load(path):
        ok = [p["load_vs_plain_capacity"] for p in sweep if p[path]["p99_ms"] <= slo_ms]
        return max(ok) if ok else 0.0

    top = next((p for p in sweep if p["load_vs_plain_capacity"] == 1.0), sweep[-1])
    slo = {"serve": max_load("serve"), "plain": max_load("plain")}
    # The floor under the p99 at that load: the p99 of the same request
    # stream with every request served alone on an idle GPU.
    ctx.set_trace_mode(mode_of[True], modes[mode_of[True]])
    stream = _requests(scenes, W.serve_requests(n_open, top["arrival_rate_hz"], len(scenes)))
    alone = []
    for req in stream:
        arr, st, fin, _, _ = serve_run(ctx, [req[:3] + (0.0,)], True, 1 << 30, max_active)
        alone.append(float((fin[0] - arr[0]) * 1e3))
    floor = {"p99_ms": round(float(np.percentile(alone, 99)), 3),
             "p50_ms": round(float(np.median(alone)), 3),
             "max_ms": round(max(alone), 3)}
    need = top["plain"]["p99_ms"] / 10.0
    reason = None
    if top["p99_ratio"] < 10.0:
        reason = (f"at load {top['load_vs_plain_capacity']:g} the plain path's p99 is "
                  f"{top['plain']['p99_ms']:.2f} ms, so 10x needs a serving p99 of "
                  f"{need:.2f} ms; the serving p99 there is {top['serve']['p99_ms']:.2f} ms, "
                  f"and the same {n_open} requests served one at a time on an idle GPU "
                  f"already have a p99 of {floor['p99_ms']:.2f} ms (max "
                  f"{floor['max_ms']:.2f} ms): the largest requests' own service time, "
                  f"not queueing, sets the serving p99")
    return {
        "same_requests_alone": floor,
        "p99_headroom_reason": reason,
        "plain_capacity_req_s": round(capacity, 2),
        "schedule": {"serve": mode_of[True], "plain": mode_of[False]},
        "saturation": {"serve": serve_sat, "plain": plain_sat},
        "saturation_mrays_x": round(serve_sat["mrays_s"] / max(plain_sat["mrays_s"], 1e-9), 2),
        "sweep": sweep,
        "p99_headroom_x": top["p99_ratio"],
        "p99_headroom_load": top["load_vs_plain_capacity"],
        "slo_p99_ms": slo_ms,
        "max_load_within_slo": slo,
        "slo_load_x": round(slo["serve"] / slo["plain"], 2) if slo["plain"] else None,
    }


# ---------------------------------------------------------------------------
# Batching: 64 scenes in one persistent launch per depth level, hipGraph
# ---------------------------------------------------------------------------
class BatchedJobs:
    """Buffers and per-level batches of the batched configuration."""

    def __init__(self, ctx: Context, jobs, lam: float):
        torch = _torch()
        self.ctx = ctx
        self.jobs = []
        for i, (tris, side, depth) in enumerate(jobs):
            v, nrm = W.sphere_scene(W.SEED_BATCH + 1000 + i, tris)
            sc = Scene(ctx)
            sc.upload(v, nrm)
            n = side * side
            self.jobs.append({"scene": sc, "side": side, "depth": depth, "n": n,
                              "cam": W.camera(side, side), "tris": v.shape[0],
                              "rays": [RayBatch.empty(n), RayBatch.empty(n)],
                              "hits": [HitBatch.empty(n), HitBatch.empty(n)],
                              "shadow": RayBatch.empty(n),
                              "occ": torch.empty(n, dtype=torch.uint8, device="cuda:0")})
        self.levels = 1 + max(j["depth"] for j in self.jobs)
        self.set_lambda(lam)

    def set_lambda(self, lam):
        for j in self.jobs:
            j["scene"].build_grid(lam)
        self.batches = []
        for lvl in range(self.levels):
            items = []
            for j in self.jobs:
                items += self._items(j, lvl)
            b = Batch(self.ctx, max(1, len(items)))
            b.set(items)
            self.batches.append(b)
        self.ctx.synchronize()

    @staticmethod
    def _items(j, lvl):
        if lvl > j["depth"]:
            return []
        if lvl == 0:
            return [(j["scene"], j["rays"][0], j["hits"][0], None, j["side"])]
        c = (lvl - 1) & 1
        items = [(j["scene"], j["shadow"], None, j["occ"], j["side"])]
        if lvl < j["depth"]:
            items.append((j["scene"], j["rays"][1 - c], j["hits"][1 - c], None, j["side"]))
        return items

    def _gen(self, j, lvl):
        sc = j["scene"]
        if lvl == 0:
            self.ctx.gen_primary(j["cam"], j["rays"][0])
            return
        c = (lvl - 1) & 1
        sc.gen_shadow(j["rays"][c], j["hits"][c], W.LIGHT, 1e-4, j["shadow"])
        if lvl < j["depth"]:
            sc.gen_reflection(j["rays"][c], j["hits"][c], 1e-4, j["rays"][1 - c])

    def run_batched(self):
        for lvl in range(self.levels):
            for j in self.jobs:
                if lvl <= j["depth"]:
                    self._gen(j, lvl)
            self.batches[lvl].trace()

    def run_plain(self):
        """One job after another, one launch per ray batch, with the context's
        current schedule."""
        for j in self.jobs:
            for lvl in range(j["depth"] + 1):
                self._gen(j, lvl)
                for (s, rays, hits, occ, width) in self._items(j, lvl):
                    if hits is not None:
                        s.trace_closest(rays, hits, image_width=width)
                    else:
                        s.trace_any(rays, occ, image_width=width)

    def active_rays(self):
        torch = _torch()
        total = 0
        for j in self.jobs:
            total += j["n"]
        # secondary rays: count once by running the levels eagerly
        for lvl in range(1, self.levels):
            for j in self.jobs:
                if lvl <= j["depth"]:
                    self._gen(j, lvl)
            self.batches[lvl].trace()
            for j in self.jobs:
                if lvl <= j["depth"]:
                    f = j["shadow"].f
                    total += int((f["tmin"] <= f["tmax"]).sum().item())
                    if lvl < j["depth"]:
                        f = j["rays"][1 - ((lvl - 1) & 1)].f
                        total += int((f["tmin"] <= f["tmax"]).sum().item())
        torch.cuda.synchronize()
        return total


def _time(fn, reps, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def batch_config(reps: int = 5):
    """Batched (one launch per level, hipGraph) vs plain; the autotuner picks
    the schedule (one ray per lane, pooled, persistent, or 2/4/8 lanes per ray,
    i.e. 64/32/16/8 rays per wave), lambda in {2,5,8}, block size and refill
    threshold.  The plain path gets its best of the single-scene schedules, the
    split one (4 lanes per ray) included."""
    torch = _torch()
    ctx = Context(0)
    stream = stream_create()
    torch.cuda.synchronize()
    ctx.set_stream(stream)
    jobs = W.batch_jobs()
    bj = BatchedJobs(ctx, jobs, 5.0)
    tuner = Autotuner()
    cands = [{"lam": lam, "block": blk, "mode": m, "refill": 16}
             for lam in (2.0, 5.0, 8.0) for blk in (64, 128, 256) for m in ("simple", "pooled")]
    cands += [{"lam": lam, "block": blk, "mode": "persistent", "refill": rf}
              for lam in (2.0, 5.0, 8.0) for blk in (64, 128, 256) for rf in (8, 16, 32)]
    # several lanes per ray: 32, 16 or 8 rays per wave ("refill" carries the lanes)
    cands += [{"lam": lam, "block": blk, "mode": "split", "refill": lanes}
              for lam in (2.0, 5.0, 8.0) for blk in (128, 256) for lanes in (2, 4, 8)]
    cur = {"lam": 5.0}

    def measure(c):
        if c["lam"] != cur["lam"]:
            bj.set_lambda(c["lam"])
            cur["lam"] = c["lam"]
        ctx.set_block_size(c["block"], c["block"])
        ctx.set_trace_mode(c["mode"], c["refill"])
        bj.run_batched()
        ctx.synchronize()
        t0 = time.perf_counter()
        bj.run_batched()
        ctx.synchronize()
        return time.perf_counter() - t0

    best = tuner.tune("batch-64", cands, measure)
    if best["lam"] != cur["lam"]:
        bj.set_lambda(best["lam"])
    ctx.set_block_size(best["block"], best["block"])
    ctx.set_trace_mode(best["mode"], best["refill"])
    rays = bj.active_rays()
    g = Graph(ctx)
    g.begin()
    bj.run_batched()
    g.end()
    g.launch()
    ctx.synchronize()

    def replay():
        g.launch()
        ctx.synchronize()

    t_graph = _time(replay, reps, torch)
    t_eager = _time(lambda: (bj.run_batched(), ctx.synchronize()), reps, torch)
    # the plain path at the same lambda, with its best schedule
    t_plain_by = {}
    for mode, param in (("simple", best["refill"]), ("pooled", best["refill"]),
                        ("persistent", best["refill"]), ("split", 4)):
        ctx.set_trace_mode(mode, param)
        t_plain_by[mode] = _time(lambda: (bj.run_plain(), ctx.synchronize()), reps, torch)
    t_plain = min(t_plain_by.values())
    ctx.set_trace_mode(best["mode"], best["refill"])
    g.close()
    out = {"jobs": len(jobs), "triangles": int(sum(j["tris"] for j in bj.jobs)),
           "primary_rays": int(sum(j["n"] for j in bj.jobs)), "rays_per_run": rays,
           "tuned": best,
           "rays_per_wave": 64 // best["refill"] if best["mode"] == "split" else 64,
           "batched_graph_ms": round(t_graph * 1e3, 3), "batched_eager_ms": round(t_eager * 1e3, 3),
           "plain_ms": round(t_plain * 1e3, 3),
           "plain_ms_by_schedule": {k: round(v * 1e3, 3) for k, v in t_plain_by.items()},
           "mrays_s": {"batched_graph": round(rays / t_graph / 1e6, 2),
                       "batched_eager": round(rays / t_eager / 1e6, 2),
                       "plain": round(rays / t_plain / 1e6, 2)},
           "autotune_top5": sorted(tuner.cache["batch-64"]["tried"],
                                   key=lambda r: r["cost"])[:5]}
    ctx.synchronize()
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    return out


# ---------------------------------------------------------------------------
# Ambient occlusion on 2^24 triangles
# ---------------------------------------------------------------------------
def ao_config(lams=(4.0, 8.0, 16.0), reps: int = 3, width: int = 1024, height: int = 512,
              spp: int = 8):
    """The incoherent large scene: 2^24 triangles, grid rebuilt every frame,
    8 cosine-weighted AO rays per pixel at 1024x512 from the primary hits (any
    hit, tmax = 0.05 x scene extent).  One frame = rebuild + primary rays +
    AO ray generation + AO rays; lambda is tuned on the frame.  Reports the
    build time, AO Mrays/s, the algorithmic bytes per AO ray (36 + 8 C + 52 T)
    and their HBM-roofline fraction."""
    torch = _torch()
    ctx = Context(0)
    v, nrm = W.ao_scene()
    sc = Scene(ctx)
    sc.upload(v, nrm)
    cam = W.ao_camera(width, height)
    n = width * height
    prim, hits = RayBatch.empty(n), HitBatch.empty(n)
    ao = RayBatch.empty(n * spp)
    occ = torch.empty(n * spp, dtype=torch.uint8, device="cuda:0")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    state = {}

    def frame(lam, extent, record=False):
        e = ev if record else [None] * 5
        if e[0] is not None:
            e[0].record()
        sc.rebuild_grid_async(lam)
        if e[1] is not None:
            e[1].record()
        ctx.gen_primary(cam, prim)
        sc.trace_closest(prim, hits, image_width=width)
        if e[2] is not None:
            e[2].record()
        sc.gen_ao(prim, hits, spp, W.SEED_AO, 1e-4, 0.05 * extent, ao)
        if e[3] is not None:
            e[3].record()
        sc.trace_any(ao, occ)
        if e[4] is not None:
            e[4].record()

    def setup(lam):
        sc.build_grid(lam)
        info = sc.grid_info()
        sc.reserve_grid(int(info["m"] * 1.1) + 4096, int(info["r"] * 1.1) + 4096)
        state["extent"] = float(np.max(info["bmax"] - info["bmin"]))
        state["lam"] = lam

    def measure(c):
        setup(c["lam"])
        best = {}
        for mode in ("simple", "pooled", "persistent"):
            ctx.set_trace_mode(mode, 16)
            frame(c["lam"], state["extent"])
            frame(c["lam"], state["extent"], record=True)
            torch.cuda.synchronize()
            best[mode] = sum(ev[i].elapsed_time(ev[i + 1]) for i in range(4))
        c["mode"] = min(best, key=best.get)
        return best[c["mode"]] * 1e-3

    tuner = Autotuner()
    cands = [{"lam": lam} for lam in lams]
    pick = tuner.tune("ao-16M", cands, measure)
    setup(pick["lam"])
    ctx.set_trace_mode(pick["mode"], 16)
    frame(pick["lam"], state["extent"])
    torch.cuda.synchronize()
    stage = np.zeros(4)
    for _ in range(reps):
        frame(pick["lam"], state["extent"], record=True)
        torch.cuda.synchronize()
        stage += [ev[i].elapsed_time(ev[i + 1]) for i in range(4)]
    stage /= reps
    sc.check_grid()
    active = int((ao.f["tmin"] <= ao.f["tmax"]).sum().item())
    cnt = torch.empty(2 * n * spp, dtype=torch.int32, device="cuda:0")
    sc.trace_any(ao, occ, counters=cnt)
    c = cnt.view(-1, 2).to(torch.int64)
    act = ao.f["tmin"] <= ao.f["tmax"]
    cells = float(c[act, 0].double().mean().item())
    tests = float(c[act, 1].double().mean().item())
    bytes_ray = 36 + 8 * cells + 52 * tests
    info = sc.grid_info()
    n_tris = v.shape[0]
    build_bytes = 36 * n_tris + 4 * info["r"] + 4 * (info["m"] + 1)
    build_ms, prim_ms, gen_ms, ao_ms = (float(x) for x in stage)
    gbs = active * bytes_ray / (ao_ms * 1e-3) / 1e9
    return {"triangles": n_tris, "ao_rays": active, "spp": spp, "lambda": pick["lam"],
            "schedule": pick["mode"],
            "primary_hit_rate": round(float((hits.f["id"] >= 0).float().mean().item()), 4),
            "grid": {"n": [int(x) for x in info["n"]], "cells": info["m"], "refs": info["r"]},
            "frame_ms": round(float(stage.sum()), 3),
            "stage_ms": {"build": round(build_ms, 3), "primary": round(prim_ms, 3),
                         "ao_gen": round(gen_ms, 3), "ao_trace": round(ao_ms, 3)},
            "ao_mrays_s": round(active / (ao_ms * 1e-3) / 1e6, 2),
            "mean_cells": round(cells, 2), "mean_tests": round(tests, 2),
            "algorithmic_bytes_per_ray": round(bytes_ray, 1),
            "ao_trace_gbs": round(gbs, 1), "ao_roofline_frac": round(gbs / 8000.0, 4),
            "build_algorithmic_gbs": round(build_bytes / (build_ms * 1e-3) / 1e9, 1),
            "autotune": tuner.cache["ao-16M"]["tried"]}


# ---------------------------------------------------------------------------
# kd-tree vs grid on the stadium frame
# ---------------------------------------------------------------------------
def kdtree_config(lam: float = 256.0, reps: int = 5, width: int = 1920, height: int = 1080,
                  two_level=(8.0, 4.0, 8)):
    """The kd-tree comparison accelerator against the uniform grid on the
    stadium frame (2^20 triangles in 256 clusters + the ground plane, frame 0's
    rigid cluster transforms): the same four traversal stages -- primary,
    shadow, mirror bounce, bounce shadow, rays generated from the grid's hits --
    timed per launch (HIP events, median of `reps`) through the one-level grid
    (best of one ray per lane, pooled tests and 2 / 16 lanes per ray), through
    the kd-tree and through the two-level grid, with the per-ray work of each
    (cells or leaves, triangle tests) and the build times (the grids' device
    rebuilds, the kd-tree's host build)."""
    torch = _torch()
    ctx = Context(0)
    seed = W.SEED_STADIUM
    v, nrm, centers = gen_stadium(seed, 256, 4096, (1 << 20) - 2)
    sc = Scene(ctx)
    sc.upload(v, nrm)
    sc.transform_clusters(gen_cluster_transforms(seed, 0, centers, 0.05), 4096, (1 << 20) - 2)
    sc.build_grid(lam)
    info = sc.grid_info()
    sc.reserve_grid(int(info["m"] * 1.25) + 4096, int(info["r"] * 1.5) + 4096)
    # the device rebuild into the reserved buffers (a grown reserve drops the
    # grid until the next build)
    ev_b = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    sc.rebuild_grid_async(lam)
    ev_b[0].record()
    sc.rebuild_grid_async(lam)
    ev_b[1].record()
    sc.check_grid()
    rebuild_ms = round(ev_b[0].elapsed_time(ev_b[1]), 3)
    kd = KdTree(sc)
    kinfo = kd.info()
    n = width * height
    cam = camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0.0, 1.0, 0.0), 60.0, width, height)
    prim, hits = RayBatch.empty(n), HitBatch.empty(n)
    ctx.gen_primary(cam, prim)
    sc.trace_closest(prim, hits, image_width=width)
    shadow, refl, rhits, shadow2 = RayBatch.empty(n), RayBatch.empty(n), HitBatch.empty(n), \
        RayBatch.empty(n)
    sc.gen_shadow(prim, hits, W.LIGHT, 1e-4, shadow)
    sc.gen_reflection(prim, hits, 1e-4, refl)
    sc.trace_closest(refl, rhits, image_width=width)
    sc.gen_shadow(refl, rhits, W.LIGHT, 1e-4, shadow2)
    occ = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    out_h = HitBatch.empty(n)
    cnt = torch.empty(2 * n, dtype=torch.int32, device="cuda:0")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        fn()
        ms = []
        for _ in range(reps):
            ev[0].record()
            fn()
            ev[1].record()
            torch.cuda.synchronize()
            ms.append(ev[0].elapsed_time(ev[1]))
        return float(np.median(ms))

    stages = {"primary": (prim, True), "shadow": (shadow, False), "bounce": (refl, True),
              "bounce_shadow": (shadow2, False)}
    out = {"triangles": int(v.shape[0]), "lambda": lam,
           "grid": {"n": [int(x) for x in info["n"]], "cells": info["m"], "refs": info["r"]},
           "kdtree": {k: kinfo[k] for k in ("nodes", "leaves", "empty_leaves", "elements",
                                            "depth")},
           "kdtree_host_build_ms": round(kinfo["build_ms"], 1), "stages": {}}
    out["grid_device_rebuild_ms"] = rebuild_ms
    tot = {"grid": 0.0, "kd": 0.0}
    n_rays = 0
    for name, (rays, closest) in stages.items():
        active = rays.f["tmin"] <= rays.f["tmax"]
        n_act = int(active.sum().item())
        n_rays += n_act
        grid_ms = {}
        for mode, param in (("simple", 16), ("pooled", 16), ("split", 2), ("split", 16)):
            ctx.set_trace_mode(mode, param)
            if closest:
                grid_ms[f"{mode}{param if mode == 'split' else ''}"] = timed(
                    lambda: sc.trace_closest(rays, out_h, image_width=width))
            else:
                grid_ms[f"{mode}{param if mode == 'split' else ''}"] = timed(
                    lambda: sc.trace_any(rays, occ, image_width=width))
        ctx.set_trace_mode("simple", 16)
        if closest:
            kd_ms = timed(lambda: kd.trace_closest(rays, out_h, image_width=width))
            kd.trace_closest(rays, out_h, image_width=width, counters=cnt)
        else:
            kd_ms = timed(lambda: kd.trace_any(rays, occ, image_width=width))
            kd.trace_any(rays, occ, image_width=width, counters=cnt)
        kc = cnt.view(-1, 2)[active].double().mean(0)
        if closest:
            sc.trace_closest(rays, out_h, image_width=width, counters=cnt)
        else:
            sc.trace_any(rays, occ, image_width=width, counters=cnt)
        gc = cnt.view(-1, 2)[active].double().mean(0)
        best = min(grid_ms, key=grid_ms.get)
        tot["grid"] += grid_ms[best]
        tot["kd"] += kd_ms
        out["stages"][name] = {
            "rays": n_act, "grid_ms": round(grid_ms[best], 3), "grid_schedule": best,
            "grid_ms_by_schedule": {k: round(x, 3) for k, x in grid_ms.items()},
            "kd_ms": round(kd_ms, 3),
            "grid_mrays_s": round(n_act / (grid_ms[best] * 1e-3) / 1e6, 1),
            "kd_mrays_s": round(n_act / (kd_ms * 1e-3) / 1e6, 1),
            "grid_cells_tests_per_ray": [round(float(gc[0]), 2), round(float(gc[1]), 2)],
            "kd_leaves_tests_per_ray": [round(float(kc[0]), 2), round(float(kc[1]), 2)]}
    # The same rays through the two-level grid (top lambda 8, sub-grids at
    # lambda2 4 in cells of >= 8 references), one ray per lane or 8 / 16 lanes
    # per ray, per stage the faster.
    lam1, lam2, min_refs = two_level
    sc.set_subgrid(lam2, min_refs)
    sc.build_grid(lam1)
    # an untimed rebuild first, like the one-level grid's above (the first
    # asynchronous two-level rebuild also sets up the context's second stream)
    sc.rebuild_grid_async(lam1)
    ev_b[0].record()
    sc.rebuild_grid_async(lam1)
    ev_b[1].record()
    sc.check_grid()
    out["two_level"] = {"grid": "%g/%g/%d" % two_level, "device_rebuild_ms":
                        round(ev_b[0].elapsed_time(ev_b[1]), 3)}
    tot["two_level"] = 0.0
    for name, (rays, closest) in stages.items():
        active = rays.f["tmin"] <= rays.f["tmax"]
        g2_ms = {}
        for mode, param in (("simple", 16), ("split", 8), ("split", 16)):
            ctx.set_trace_mode(mode, param)
            if closest:
                g2_ms[f"{mode}{param if mode == 'split' else ''}"] = timed(
                    lambda: sc.trace_closest(rays, out_h, image_width=width))
            else:
                g2_ms[f"{mode}{param if mode == 'split' else ''}"] = timed(
                    lambda: sc.trace_any(rays, occ, image_width=width))
        ctx.set_trace_mode("simple", 16)
        if closest:
            sc.trace_closest(rays, out_h, image_width=width, counters=cnt)
        else:
            sc.trace_any(rays, occ, image_width=width, counters=cnt)
        c2 = cnt.view(-1, 2)[active].double().mean(0)
        best = min(g2_ms, key=g2_ms.get)
        tot["two_level"] += g2_ms[best]
        e = out["stages"][name]
        e["two_level_ms"] = round(g2_ms[best], 3)
        e["two_level_schedule"] = best
        e["two_level_mrays_s"] = round(e["rays"] / (g2_ms[best] * 1e-3) / 1e6, 1)
        e["two_level_cells_tests_per_ray"] = [round(float(c2[0]), 2), round(float(c2[1]), 2)]
    sc.set_subgrid(0.0, 8)
    out["traversal_ms"] = {k: round(x, 3) for k, x in tot.items()}
    out["traversal_mrays_s"] = {k: round(n_rays / (x * 1e-3) / 1e6, 1) for k, x in tot.items()}
    kd.close()
    return out
