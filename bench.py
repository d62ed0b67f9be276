#!/usr/bin/env python3
"""Benchmark: Mrays/s of the uniform-grid tracer on one MI355X.

Workload ("teapot in a stadium", 2^20 binary32 triangles): 256 Gaussian
clusters of small triangles + a 2-triangle ground plane 100x the cluster
extent.  One step = one frame:
  1. per-cluster rigid transform of the rest pose (device),
  2. uniform-grid rebuild on the device (lambda = 5),
  3. 1920x1080 primary rays, closest hit,
  4. shadow rays toward a point light, any hit,
  5. one mirror bounce, closest hit,
  6. shadow rays from the bounce hits, any hit.
`value` counts every ray traced in the frame (inactive slots excluded) per
second of frame time, with scene, grid buffers and ray buffers resident in
HBM before the timed region.

Also reported: per-stage milliseconds (HIP events on the tracing stream), a
roofline for the dominant kernel (primary closest-hit traversal) from the
algorithmic bytes per ray 32 + 16 + 8*C + 52*T (C cells, T triangle tests,
counted on the device outside the timed region), and the CPU oracle timed on
the host cores on a bounded sample of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "grid-rt_amd"))

import numpy as np  # noqa: E402

SEED = 0x6772740000000002
N_CLUSTERS = 256
TRIS_PER_CLUSTER = 4096
N_CLUSTER_TRIS = (1 << 20) - 2
WIDTH, HEIGHT = 1920, 1080
LAMBDA = 5.0
LAMBDAS = (5.0, 8.0, 16.0, 32.0, 64.0, 128.0)
LIGHT = (1000.0, 1000.0, 1000.0)
OFFSET = 1e-4
EYE, CENTER, UP = (0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0.0, 1.0, 0.0)
FOVY = 60.0
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="approximate CPU time budget of the baseline sample")
    p.add_argument("--block", type=int, default=0, help="traversal block size (0 = default)")
    p.add_argument("--sched", default=None,
                   help="traversal schedule per stage (primary, shadow, bounce, bounce "
                        "shadow): s = one ray per lane, p = persistent refill; default: tuned")
    p.add_argument("--refill", type=int, default=16)
    p.add_argument("--lam", type=float, default=None,
                   help="grid resolution factor; default: tuned over %s" % (LAMBDAS,))
    p.add_argument("--mode", default="frame", choices=["frame", "serve", "batch", "ao"],
                   help="frame: the stadium frame (default, headline line, with a short "
                        "serve-vs-plain latency section); serve / batch / ao: the serving, "
                        "batching and ambient-occlusion workloads on their own")
    p.add_argument("--no-serve", action="store_true", help="frame mode without the latency run")
    p.add_argument("--serve-requests", type=int, default=120)
    return p.parse_args()


class Frame:
    """Device buffers and the per-frame pipeline of the stadium workload."""

    def __init__(self, grt, torch, lam=LAMBDA, sched="spsp", refill=16):
        self.grt, self.torch = grt, torch
        self.lam = lam
        self.sched = sched
        self.refill = refill
        self.ctx = grt.Context(0)
        v, n, centers = grt.gen_stadium(SEED, N_CLUSTERS, TRIS_PER_CLUSTER, N_CLUSTER_TRIS)
        self.verts, self.normals, self.centers = v, n, centers
        self.scene = grt.Scene(self.ctx)
        self.scene.upload(v, n)
        self.cam = grt.camera_lookat(EYE, CENTER, UP, FOVY, WIDTH, HEIGHT)
        self.n = WIDTH * HEIGHT
        self.prim = grt.RayBatch.empty(self.n)
        self.hits = grt.HitBatch.empty(self.n)
        self.shadow = grt.RayBatch.empty(self.n)
        self.occ = torch.empty(self.n, dtype=torch.uint8, device="cuda:0")
        self.refl = grt.RayBatch.empty(self.n)
        self.rhits = grt.HitBatch.empty(self.n)
        self.shadow2 = grt.RayBatch.empty(self.n)
        self.occ2 = torch.empty(self.n, dtype=torch.uint8, device="cuda:0")
        # Grid buffers sized with headroom by one synchronous build of frame 0.
        self.xf = {}
        self.set_lambda(lam)
        names = ["transform", "build", "primary", "shadow", "bounce", "bounce_shadow"]
        self.stage_names = names

    def set_lambda(self, lam):
        """Synchronous build of frame 0 at this resolution; reserves headroom
        so the per-frame rebuilds never reallocate."""
        self.lam = lam
        self.scene.transform_clusters(self.transforms(0), TRIS_PER_CLUSTER, N_CLUSTER_TRIS)
        self.scene.build_grid(lam)
        info = self.scene.grid_info()
        self.scene.reserve_grid(int(info["m"] * 1.25) + 4096, int(info["r"] * 1.5) + 4096)

    def time_stages(self, frames=2):
        """Stage milliseconds of the last of `frames` frames."""
        ev = [self.torch.cuda.Event(enable_timing=True) for _ in range(7)]
        for f in range(frames):
            self.run(2000 + f, ev if f == frames - 1 else None)
        self.torch.cuda.synchronize()
        return [ev[i].elapsed_time(ev[i + 1]) for i in range(6)]

    def transforms(self, frame):
        if frame not in self.xf:
            self.xf[frame] = self.grt.gen_cluster_transforms(SEED, frame, self.centers, 0.05)
        return self.xf[frame]

    def run(self, frame, events=None):
        sc, W = self.scene, WIDTH
        ev = events if events is not None else [None] * 7

        def mark(i):
            if ev[i] is not None:
                ev[i].record()

        mark(0)
        sc.transform_clusters(self.transforms(frame), TRIS_PER_CLUSTER, N_CLUSTER_TRIS)
        mark(1)
        sc.rebuild_grid_async(self.lam)
        mark(2)
        self.ctx.gen_primary(self.cam, self.prim)
        self.mode(0)
        sc.trace_closest(self.prim, self.hits, image_width=W)
        mark(3)
        sc.gen_shadow(self.prim, self.hits, LIGHT, OFFSET, self.shadow)
        self.mode(1)
        sc.trace_any(self.shadow, self.occ, image_width=W)
        mark(4)
        sc.gen_reflection(self.prim, self.hits, OFFSET, self.refl)
        self.mode(2)
        sc.trace_closest(self.refl, self.rhits, image_width=W)
        mark(5)
        sc.gen_shadow(self.refl, self.rhits, LIGHT, OFFSET, self.shadow2)
        self.mode(3)
        sc.trace_any(self.shadow2, self.occ2, image_width=W)
        mark(6)

    def mode(self, stage):
        """Traversal schedule of a stage: 's' one ray per lane, 'p' persistent."""
        self.ctx.set_trace_mode(self.sched[stage] == "p", self.refill)

    def active(self, rb):
        f = rb.f
        return int((f["tmin"] <= f["tmax"]).sum().item())


def algorithmic_bytes(frame, torch):
    """Bytes per launch of the primary closest-hit kernel from device counters."""
    cnt = torch.empty(2 * frame.n, dtype=torch.int32, device="cuda:0")
    frame.mode(0)
    frame.scene.trace_closest(frame.prim, frame.hits, image_width=WIDTH, counters=cnt)
    torch.cuda.synchronize()
    c = cnt.view(-1, 2).to(torch.int64)
    cells = int(c[:, 0].sum().item())
    tests = int(c[:, 1].sum().item())
    n = frame.n
    return 48 * n + 8 * cells + 52 * tests, cells / n, tests / n


def autotune(frame, lams, sched):
    """Pick lambda and the per-stage traversal schedule on the frame itself
    (outside the timed region): for each lambda, time every stage under both
    schedules and keep the faster per stage."""
    from grt.autotune import Autotuner, per_stage_schedule
    tuner = Autotuner()
    table = {}

    def measure(c):
        frame.set_lambda(c["lam"])
        if sched:
            frame.sched = sched
            ms = frame.time_stages()
            table[c["lam"]] = {"sched": sched, "frame_ms": round(sum(ms), 3)}
            return sum(ms)
        frame.sched = "ssss"
        ms_s = frame.time_stages()
        frame.sched = "pppp"
        ms_p = frame.time_stages()
        best = per_stage_schedule(ms_s[2:], ms_p[2:])
        total = sum(ms_s[:2]) / 2 + sum(ms_p[:2]) / 2 + sum(min(a, b) for a, b in
                                                             zip(ms_s[2:], ms_p[2:]))
        table[c["lam"]] = {"sched": best, "frame_ms": round(total, 3)}
        return total

    best = tuner.tune("stadium-1M", [{"lam": lam} for lam in lams], measure)
    return best["lam"], table[best["lam"]]["sched"], table


def stage_counters(frame, torch):
    """Mean/max cells and triangle tests over the active rays of each stage
    (device counters, outside the timed region)."""
    out = {}
    cnt = torch.empty(2 * frame.n, dtype=torch.int32, device="cuda:0")
    stages = [("primary", frame.prim, True), ("shadow", frame.shadow, False),
              ("bounce", frame.refl, True), ("bounce_shadow", frame.shadow2, False)]
    scratch_h = frame.grt.HitBatch.empty(frame.n)
    scratch_o = torch.empty(frame.n, dtype=torch.uint8, device="cuda:0")
    for i, (name, rays, closest) in enumerate(stages):
        frame.mode(i)
        if closest:
            frame.scene.trace_closest(rays, scratch_h, image_width=WIDTH, counters=cnt)
        else:
            frame.scene.trace_any(rays, scratch_o, image_width=WIDTH, counters=cnt)
        active = rays.f["tmin"] <= rays.f["tmax"]
        c = cnt.view(-1, 2)[active].to(torch.int64)
        if c.shape[0] == 0:
            continue
        out[name] = {"rays": int(c.shape[0]),
                     "mean_cells": round(float(c[:, 0].double().mean().item()), 1),
                     "mean_tests": round(float(c[:, 1].double().mean().item()), 1),
                     "max_tests": int(c[:, 1].max().item())}
    return out


def cpu_baseline(frame, torch, budget_s):
    """Oracle (C, OpenMP over rows, schedule(dynamic)) on a row sample of the
    frame's primary + shadow rays, and its single-thread grid build."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or orc.max_threads()
    threads = max(1, min(threads, 16))
    v, nrm = frame.scene.download_geometry()
    t0 = time.perf_counter()
    fr = orc.grid_frame(v, frame.lam)
    _, offs, refs = orc.grid_build(v, fr)
    build_s = time.perf_counter() - t0
    rays = orc.primary_rays(frame.grt.camera_dict(frame.cam))

    def run_rows(stride):
        rows = np.arange(0, HEIGHT, stride)
        idx = (rows[:, None] * WIDTH + np.arange(WIDTH)[None, :]).ravel()
        sub = {k: a[idx] for k, a in rays.items()}
        t = time.perf_counter()
        # MultiThreadRenderer.cpp:23-27: OpenMP schedule(dynamic, 16) over rows.
        h = orc.trace_closest(fr, offs, refs, v, sub, threads=threads, chunk=16 * WIDTH)
        sh = orc.gen_shadow(sub, h, nrm, LIGHT, OFFSET)
        orc.trace_any(fr, offs, refs, v, sh, threads=threads, chunk=16 * WIDTH)
        dt = time.perf_counter() - t
        return len(idx) + int((sh["tmin"] <= sh["tmax"]).sum()), dt, len(rows)

    n_probe, dt_probe, _ = run_rows(120)
    rate = n_probe / max(dt_probe, 1e-6)
    stride = int(np.clip(np.ceil(HEIGHT * WIDTH * 2 / max(rate * budget_s, 1.0)), 1, 120))
    n_rays, spent, n_rows = run_rows(stride)
    return {"value": round(n_rays / spent / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "sample": (f"CPU oracle on {n_rows} image rows (every {stride}th row) of the same "
                       f"frame: primary closest-hit + shadow any-hit, {n_rays} rays in "
                       f"{spent:.1f} s, OpenMP schedule(dynamic) over 16-row chunks; "
                       f"single-thread grid build of the frame {build_s * 1e3:.0f} ms")}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of "
                          "it (rays shard by image tile with the scene replicated, see DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import grt
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    torch.cuda.set_device(0)
    if args.mode != "frame":
        from grt import perf
        if args.mode == "serve":
            ctx = grt.Context(0)
            scenes, tris = perf.build_serve_scenes(ctx)
            res = perf.serve_vs_plain(ctx, scenes, n_requests=args.serve_requests)
            res["triangles"] = tris
        elif args.mode == "batch":
            res = perf.batch_config()
        else:
            res = perf.ao_config()
        print(json.dumps({"mode": args.mode, **res}))
        return
    frame = Frame(grt, torch, args.lam or LAMBDA, args.sched or "spsp", args.refill)
    if args.block:
        frame.ctx.set_block_size(args.block, args.block)
    lams = [args.lam] if args.lam else list(LAMBDAS)
    lam, sched, tune_table = autotune(frame, lams, args.sched)
    frame.sched = sched
    frame.set_lambda(lam)
    for w in range(args.warmup):
        frame.run(1000 + w)
    frame.scene.check_grid()
    torch.cuda.synchronize()

    n_ev = 7
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(n_ev)]
              for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        frame.run(s, events[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    frame.scene.check_grid()

    stages = {name: 0.0 for name in frame.stage_names}
    for s in range(args.steps):
        for i, name in enumerate(frame.stage_names):
            stages[name] += events[s][i].elapsed_time(events[s][i + 1])
    stages = {k: v / args.steps for k, v in stages.items()}
    ms_per_step = wall * 1e3 / args.steps

    rays = {"primary": frame.n, "shadow": frame.active(frame.shadow),
            "bounce": frame.active(frame.refl), "bounce_shadow": frame.active(frame.shadow2)}
    total_rays = sum(rays.values())
    value = total_rays / (ms_per_step * 1e-3) / 1e6
    info = frame.scene.grid_info()

    bytes_launch, mean_c, mean_t = algorithmic_bytes(frame, torch)
    prim_ms = stages["primary"]
    achieved = bytes_launch / (prim_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "trace_closest_stadium_traffic.md")
    if os.path.exists(tpath):
        for line in open(tpath):
            if line.startswith("hbm_bytes_per_launch:"):
                traffic = float(line.split(":", 1)[1])

    out = {
        "metric": "Mrays/s (primary+shadow) at 1 GPU; p50/p99 request latency, serve vs plain",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded stadium scene, random per-frame rigid cluster transforms)",
        "config": {
            "workload": "stadium-1M: 2^20 tris (256 Gaussian clusters + 2-tri ground plane), "
                        "1920x1080 primary + shadow + 1 mirror bounce + bounce shadow, "
                        f"per-frame transform + device grid rebuild (lambda={lam:g}, autotuned)",
            "rays_per_frame": rays,
            "grid": {"n": [int(x) for x in info["n"]], "cells": info["m"], "refs": info["r"]},
            "stage_ms": {k: round(v, 3) for k, v in stages.items()},
            "mrays_primary_shadow_only": round((rays["primary"] + rays["shadow"]) /
                                               ((stages["primary"] + stages["shadow"]) * 1e-3) /
                                               1e6, 2),
            "mean_cells_per_primary_ray": round(mean_c, 2),
            "mean_tests_per_primary_ray": round(mean_t, 2),
            "stage_counters": stage_counters(frame, torch),
            "traversal": {"schedule_per_stage": sched, "refill": args.refill,
                          "block": args.block or 128, "lambda": lam},
            "autotune_frame_ms_by_lambda": tune_table,
            "parallelism": "single GPU",
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_trace_closest (primary)",
                     "bytes_per_launch": bytes_launch},
    }
    if not args.no_serve:
        # p50/p99 request latency, continuous batching vs the plain path, on
        # the serving workload (16 resident sphere scenes, mixed requests).
        from grt import perf
        scenes, tris = perf.build_serve_scenes(frame.ctx)
        lat = perf.serve_vs_plain(frame.ctx, scenes, n_requests=args.serve_requests)
        lat["triangles"] = tris
        out["latency"] = lat
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(frame, torch, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
