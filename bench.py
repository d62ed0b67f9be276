#!/usr/bin/env python3
"""Benchmark: Mrays/s of the uniform-grid tracer on one MI355X.

Workload ("teapot in a stadium", 2^20 binary32 triangles): 256 Gaussian
clusters of small triangles + a 2-triangle ground plane 100x the cluster
extent.  One step = one frame:
  1. per-cluster rigid transform of the rest pose (device),
  2. uniform-grid rebuild on the device -- one level (lambda 8..512) or two
     (a coarse top grid with uniform sub-grids in its dense cells), the
     configuration chosen by the autotuner before the timed region,
  3. 1920x1080 primary rays, closest hit,
  4. shadow rays toward a point light, any hit,
  5. one mirror bounce, closest hit,
  6. shadow rays from the bounce hits, any hit.
`value` counts every ray traced in the frame (inactive slots excluded) per
second of frame time, with scene, grid buffers and ray buffers resident in
HBM before the timed region.

Also reported: per-launch-group milliseconds (HIP events on the tracing
stream), a roofline for the dominant traversal kernel from the algorithmic
bytes per active ray -- 32 in + 16 out (closest) or 4 out (any) + 8*C + 52*T,
C cells visited and T triangle tests counted on the device outside the timed
region -- and the CPU oracle timed on the host cores on a sample of the same
frame.
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
LAMBDAS = (8.0, 32.0, 64.0, 128.0, 256.0, 512.0)
# Two-level grids: (top lambda, sub-grid lambda, minimum references of a cell
# that gets a sub-grid).
TWO_LEVEL = ((8.0, 4.0, 8), (16.0, 4.0, 8), (16.0, 4.0, 4))
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
                        "shadow): s = one ray per lane, w = one ray per lane with pooled "
                        "tests, p = persistent refill, 2/4/8 = lanes per ray, x = 16 lanes "
                        "per ray; default: tuned")
    p.add_argument("--refill", type=int, default=16)
    p.add_argument("--ray-lists", nargs="?", const="ccc", default=None,
                   help="ray lists of the three secondary stages (shadow, bounce, bounce "
                        "shadow), one letter each: - = none, c = compacted in slot order; "
                        "the flag alone means ccc; default: tuned with the schedule "
                        "(--- when --sched is given)")
    p.add_argument("--coop", type=int, default=16,
                   help="cooperative cell tests in drained persistent waves: minimum "
                        "references left (0 = off)")
    p.add_argument("--lam", type=float, default=None,
                   help="grid resolution factor; default: tuned over %s" % (LAMBDAS,))
    p.add_argument("--sub", default=None,
                   help="second grid level 'LAMBDA2,MIN_REFS' or 'off' (with --lam); "
                        "default: tuned over %s" % (TWO_LEVEL,))
    p.add_argument("--mode", default="frame", choices=["frame", "serve", "batch", "ao", "kdtree"],
                   help="frame: the stadium frame (default, headline line, with a short "
                        "serve-vs-plain latency section); serve / batch / ao: the serving, "
                        "batching and ambient-occlusion workloads on their own; kdtree: the "
                        "kd-tree comparison accelerator against the grid on the stadium frame")
    p.add_argument("--no-serve", action="store_true", help="frame mode without the latency run")
    p.add_argument("--sequential", action="store_true",
                   help="timed frames run every stage on one stream (default: the shadow and "
                        "bounce branches run on two streams)")
    p.add_argument("--serve-requests", type=int, default=120)
    p.add_argument("--profile", action="store_true",
                   help="only the warmup and timed frames (with --lam and --sched: no "
                        "tuning), for a profiler")
    p.add_argument("--copy-calibration", action="store_true",
                   help="first copy 1 GiB on the device three times (counter calibration)")
    return p.parse_args()


def intervals(ev):
    return [ev[i].elapsed_time(ev[i + 1]) for i in range(len(ev) - 1)]


# Traversal stages of a frame: (name, closest hit?).
STAGES = (("primary", True), ("shadow", False), ("bounce", True), ("bounce_shadow", False))


def grid_tag(lam, sub):
    """'LAMBDA' for a one-level grid, 'LAMBDA/LAMBDA2/MIN_REFS' for two levels."""
    return "%g" % lam if sub is None else "%g/%g/%d" % (lam, sub[0], sub[1])


def kernel_name(closest, sched_char, two_level=False):
    """The traversal kernel a stage launches under a schedule character (a
    two-level grid is walked by its own kernel under every schedule)."""
    if two_level:
        kern = "k_trace2_split" if sched_char in "248x" else "k_trace2"
        return "%s<%s>" % (kern, "false" if closest else "true")
    if sched_char == "p":
        return "k_trace_persistent<%s, false>" % ("false" if closest else "true")
    if sched_char in "248x":
        return "k_trace_split<%s, false>" % ("false" if closest else "true")
    pooled = "true" if sched_char == "w" else "false"
    return ("k_trace_closest<false, %s>" if closest else "k_trace_any<false, %s>") % pooled


class Frame:
    """Device buffers and the per-frame pipeline of the stadium workload."""

    def __init__(self, grt, torch, lam=LAMBDA, sched="spsp", refill=16, sub=None):
        self.grt, self.torch = grt, torch
        self.lam = lam
        self.sub = sub
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
        # ray lists of the three secondary stages (active ray indices + count)
        self.lists = [torch.empty(self.n, dtype=torch.int32, device="cuda:0") for _ in range(3)]
        self.counts = [torch.zeros(1, dtype=torch.int32, device="cuda:0") for _ in range(3)]
        self.lists_kind = "---"
        self.rgb = torch.empty((self.n, 3), dtype=torch.float32, device="cuda:0")
        self.rgba = torch.empty(self.n, dtype=torch.int32, device="cuda:0")
        self.phong = grt.phong({"light": LIGHT})
        # Grid buffers sized with headroom by one synchronous build of frame 0.
        self.xf = {}
        self.set_grid(lam, sub)
        names = ["transform", "build", "primary", "shadow", "bounce", "bounce_shadow"]
        self.stage_names = names

    def set_grid(self, lam, sub=None):
        """Synchronous build of frame 0 at this resolution (sub: the second
        level's (lambda2, min_refs), or None for one level); reserves headroom
        so the per-frame rebuilds never reallocate."""
        self.lam, self.sub = lam, sub
        self.scene.transform_clusters(self.transforms(0), TRIS_PER_CLUSTER, N_CLUSTER_TRIS)
        if sub is None:
            self.scene.set_subgrid(0.0, 8)
        else:
            self.scene.set_subgrid(float(sub[0]), int(sub[1]))
        self.scene.build_grid(lam)
        info = self.scene.grid_info()
        self.scene.reserve_grid(int(info["m"] * 1.25) + 4096, int(info["r"] * 1.5) + 4096)

    N_EVENTS = 12

    def time_stages(self, frames=2):
        """Milliseconds of transform, build and the four ray stages (ray
        generation + traversal each): per stage, the median over frames 2..
        `frames` (the first frame warms the caches)."""
        evs = [[self.torch.cuda.Event(enable_timing=True) for _ in range(self.N_EVENTS)]
               for _ in range(frames - 1)]
        for f in range(frames):
            self.run(2000 + f, evs[f - 1] if f > 0 else None)
        self.torch.cuda.synchronize()
        d = np.median([intervals(ev) for ev in evs], axis=0)
        return [d[0], d[1], d[2] + d[3], d[4] + d[5], d[6] + d[7], d[8] + d[9]]

    def time_frames(self, frames=5, concurrent=False):
        """Wall milliseconds per frame over `frames` frames after one warm-up."""
        self.run(3000, concurrent=concurrent)
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            self.run(3001 + f, concurrent=concurrent)
        self.torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / frames

    def transforms(self, frame):
        if frame not in self.xf:
            self.xf[frame] = self.grt.gen_cluster_transforms(SEED, frame, self.centers, 0.05)
        return self.xf[frame]

    def concurrent_ok(self):
        """Whether the shadow and bounce stages may run on two streams (the
        persistent schedules keep their slot counter per stream, the ray-list
        compaction its block sums per context: the frame compacts before the
        branches split)."""
        return True

    def run(self, frame, events=None, concurrent=False):
        """One frame.  Events (if given, N_EVENTS of them, each recorded on
        the stream of the launches it brackets) bracket every launch group:
        transform, build, for each of the four ray stages its generation
        kernel and its traversal kernel, and the Phong shading of both levels
        into the RGBA8 frame.  Concurrent: the shadow rays are traced on the
        context's stream while the bounce rays, their shadow rays and both
        traces run on a second stream (the two branches depend only on the
        primary hits); the streams join before shading.  Both ray lists of
        the first two secondary stages are made before the branches split
        (the compaction's block sums live in the context)."""
        if concurrent:
            return self.run_concurrent(frame, events)
        sc, W = self.scene, WIDTH
        ev = events if events is not None else [None] * self.N_EVENTS

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
        mark(3)
        sc.trace_closest(self.prim, self.hits, image_width=W)
        mark(4)
        L, C, K = self.lists, self.counts, self.lists_kind
        sc.gen_shadow(self.prim, self.hits, LIGHT, OFFSET, self.shadow)
        self.make_list(0, self.shadow, occluded=self.occ)
        self.mode(1)
        mark(5)
        if K[0] != "-":
            sc.trace_any_list(self.shadow, self.occ, L[0], C[0])
        else:
            sc.trace_any(self.shadow, self.occ, image_width=W)
        mark(6)
        sc.gen_reflection(self.prim, self.hits, OFFSET, self.refl)
        self.make_list(1, self.refl, miss_hits=self.rhits)
        self.mode(2)
        mark(7)
        if K[1] != "-":
            sc.trace_closest_list(self.refl, self.rhits, L[1], C[1])
        else:
            sc.trace_closest(self.refl, self.rhits, image_width=W)
        mark(8)
        sc.gen_shadow(self.refl, self.rhits, LIGHT, OFFSET, self.shadow2)
        self.make_list(2, self.shadow2, occluded=self.occ2)
        self.mode(3)
        mark(9)
        if K[2] != "-":
            sc.trace_any_list(self.shadow2, self.occ2, L[2], C[2])
        else:
            sc.trace_any(self.shadow2, self.occ2, image_width=W)
        mark(10)
        sc.shade_phong(self.prim, self.hits, self.rgb, self.occ, 0, self.phong)
        sc.shade_phong(self.refl, self.rhits, self.rgb, self.occ2, 1, self.phong, rgba=self.rgba)
        mark(11)

    def run_concurrent(self, frame, events=None):
        sc, W, torch = self.scene, WIDTH, self.torch
        main = torch.cuda.current_stream()
        if not hasattr(self, "side"):
            self.side = torch.cuda.Stream()      # default priority
        side = self.side
        ev = events if events is not None else [None] * self.N_EVENTS

        def mark(i, stream):
            if ev[i] is not None:
                ev[i].record(stream)

        L, C, K = self.lists, self.counts, self.lists_kind
        mark(0, main)
        sc.transform_clusters(self.transforms(frame), TRIS_PER_CLUSTER, N_CLUSTER_TRIS)
        mark(1, main)
        sc.rebuild_grid_async(self.lam)
        mark(2, main)
        self.ctx.gen_primary(self.cam, self.prim)
        self.mode(0)
        mark(3, main)
        sc.trace_closest(self.prim, self.hits, image_width=W)
        mark(4, main)
        sc.gen_shadow(self.prim, self.hits, LIGHT, OFFSET, self.shadow)
        self.make_list(0, self.shadow, occluded=self.occ)
        sc.gen_reflection(self.prim, self.hits, OFFSET, self.refl)
        self.make_list(1, self.refl, miss_hits=self.rhits)
        side.wait_stream(main)
        self.mode(1)
        mark(5, main)
        if K[0] != "-":
            sc.trace_any_list(self.shadow, self.occ, L[0], C[0])
        else:
            sc.trace_any(self.shadow, self.occ, image_width=W)
        mark(6, main)
        self.ctx.set_stream(side.cuda_stream)
        try:
            self.mode(2)
            mark(7, side)
            if K[1] != "-":
                sc.trace_closest_list(self.refl, self.rhits, L[1], C[1])
            else:
                sc.trace_closest(self.refl, self.rhits, image_width=W)
            mark(8, side)
            sc.gen_shadow(self.refl, self.rhits, LIGHT, OFFSET, self.shadow2)
            self.make_list(2, self.shadow2, occluded=self.occ2)
            self.mode(3)
            mark(9, side)
            if K[2] != "-":
                sc.trace_any_list(self.shadow2, self.occ2, L[2], C[2])
            else:
                sc.trace_any(self.shadow2, self.occ2, image_width=W)
            mark(10, side)
        finally:
            self.ctx.set_stream(main.cuda_stream)
        main.wait_stream(side)
        sc.shade_phong(self.prim, self.hits, self.rgb, self.occ, 0, self.phong)
        sc.shade_phong(self.refl, self.rhits, self.rgb, self.occ2, 1, self.phong, rgba=self.rgba)
        mark(11, main)

    def make_list(self, k, rays, **out):
        kind = self.lists_kind[k]
        if kind == "c":
            self.ctx.compact_rays(rays, WIDTH, self.lists[k], self.counts[k], **out)
        elif kind != "-":
            raise ValueError("ray list kind %r (expected - or c)" % kind)

    SCHEDULES = {"s": "simple", "w": "pooled", "p": "persistent", "2": "split", "4": "split",
                 "8": "split", "x": "split"}
    SPLIT_LANES = {"2": 2, "4": 4, "8": 8, "x": 16}

    def mode(self, stage):
        """Traversal schedule of a stage: 's' one ray per lane, 'w' one ray per
        lane with the wave's tests pooled, 'p' persistent, '2'/'4'/'8' that
        many lanes per ray, 'x' 16 lanes per ray."""
        ch = self.sched[stage]
        self.ctx.set_trace_mode(self.SCHEDULES[ch], self.SPLIT_LANES.get(ch, self.refill))

    def active(self, rb):
        f = rb.f
        return int((f["tmin"] <= f["tmax"]).sum().item())


def autotune(frame, grids, sched, lists):
    """Pick the grid, the per-stage traversal schedule and the secondary stages'
    ray lists on the frame itself (outside the timed region).  grids lists
    (lambda, sub) pairs, sub = None for one level or (lambda2, min_refs).  For
    a one-level grid, every stage is timed under each schedule, with and
    without ray lists (each stage's time includes its ray generation and list
    compaction), and the fastest per stage is kept.  A two-level grid has two
    traversal kernels, one ray per lane ('s') and 2..16 lanes per ray ('2',
    '4', '8', 'x'), chosen per stage the same way."""
    from grt.autotune import Autotuner
    tuner = Autotuner()
    table = {}

    def measure(c):
        sub = tuple(c["sub"]) if c["sub"] is not None else None
        tag = grid_tag(c["lam"], sub)
        frame.set_grid(c["lam"], sub)
        if sched:
            frame.sched = sched
            frame.lists_kind = lists or "---"
            ms = frame.time_stages()
            table[tag] = {"sched": sched, "lists": frame.lists_kind,
                          "frame_ms": round(sum(ms), 3), "lam": c["lam"], "sub": c["sub"]}
            return sum(ms)
        by = {}
        # on a two-level grid 's', 'w' and 'p' launch the same one-ray-per-lane
        # kernel; '2'..'x' its lanes-per-ray variant
        for ch in (Frame.SCHEDULES if sub is None else "s248x"):
            for kind in (("-",) if lists else ("-", "c")):
                frame.sched = ch * 4
                frame.lists_kind = lists or kind * 3
                by[(ch, kind)] = frame.time_stages(frames=3)
        # primary rays are never listed; each secondary stage takes its best
        # pair, a ray list only when it beats the same schedule without one by 2%
        picks = []
        for i in range(4):
            k = min((k for k in by if i > 0 or k[1] == "-"), key=lambda k: by[k][2 + i])
            if k[1] == "c" and by[(k[0], "-")][2 + i] <= 1.02 * by[k][2 + i]:
                k = (k[0], "-")
            picks.append(k)
        best = "".join(ch for ch, _ in picks)
        best_lists = lists or "".join(kind for _, kind in picks[1:])
        total = sum(sum(ms[:2]) for ms in by.values()) / len(by) + \
            sum(by[k][2 + i] for i, k in enumerate(picks))
        table[tag] = {"sched": best, "lists": best_lists, "frame_ms": round(total, 3),
                      "lam": c["lam"], "sub": c["sub"]}
        return total

    cands = [{"lam": lam, "sub": (list(sub) if sub is not None else None)} for lam, sub in grids]
    best = tuner.tune("stadium-1M", cands, measure)
    best_tag = grid_tag(best["lam"], tuple(best["sub"]) if best["sub"] is not None else None)
    if not sched and len(grids) > 1:
        # The two best grids are often within 1%: time each whole frame with
        # its tuned schedule over more frames and keep the faster.
        top = sorted(table, key=lambda k: table[k]["frame_ms"])[:2]
        final = {}
        for tag in top:
            e = table[tag]
            frame.set_grid(e["lam"], tuple(e["sub"]) if e["sub"] is not None else None)
            frame.sched, frame.lists_kind = e["sched"], e["lists"]
            final[tag] = sum(frame.time_stages(frames=7))
            e["frame_ms_final"] = round(final[tag], 3)
        best_tag = min(final, key=final.get)
        best = {"lam": table[best_tag]["lam"], "sub": table[best_tag]["sub"]}
    entry = table[best_tag]
    sub = tuple(best["sub"]) if best["sub"] is not None else None
    return best["lam"], sub, entry["sched"], entry["lists"], table


def stage_counters(frame, torch):
    """Per traversal stage, from device counters (outside the timed region):
    active rays, mean/max cells visited and triangle tests, and the stage's
    algorithmic bytes -- per active ray 32 B ray in + 16 B hit out (closest) or
    4 B flag out (any) + 8 B per cell visited (offsets[c], offsets[c+1]; on a
    two-level grid the 8-byte cell record of a top cell and the offset pair of
    a sub-cell, both counted as cells) + 52 B per triangle test (4 B reference
    + 48 B triangle record)."""
    out = {}
    cnt = torch.empty(2 * frame.n, dtype=torch.int32, device="cuda:0")
    rays_of = {"primary": frame.prim, "shadow": frame.shadow, "bounce": frame.refl,
               "bounce_shadow": frame.shadow2}
    scratch_h = frame.grt.HitBatch.empty(frame.n)
    scratch_o = torch.empty(frame.n, dtype=torch.uint8, device="cuda:0")
    for i, (name, closest) in enumerate(STAGES):
        rays = rays_of[name]
        frame.mode(i)
        if closest:
            frame.scene.trace_closest(rays, scratch_h, image_width=WIDTH, counters=cnt)
        else:
            frame.scene.trace_any(rays, scratch_o, image_width=WIDTH, counters=cnt)
        active = rays.f["tmin"] <= rays.f["tmax"]
        c = cnt.view(-1, 2)[active].to(torch.int64)
        n_act = int(c.shape[0])
        cells = int(c[:, 0].sum().item()) if n_act else 0
        tests = int(c[:, 1].sum().item()) if n_act else 0
        per_ray = 48 if closest else 36
        out[name] = {"rays": n_act,
                     "mean_cells": round(cells / max(n_act, 1), 2),
                     "mean_tests": round(tests / max(n_act, 1), 2),
                     "max_tests": int(c[:, 1].max().item()) if n_act else 0,
                     "algorithmic_bytes": per_ray * n_act + 8 * cells + 52 * tests}
    return out


def host_description():
    """CPU model, logical CPUs of the host (nproc) and CPUs this process may run
    on (its affinity mask)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "allowed_cpus": allowed}


def cpu_baseline(frame, torch, budget_s, lam):
    """The oracle (C, OpenMP over rows, schedule(dynamic, 16 rows), mirroring
    MultiThreadRenderer.cpp:23-27) on the frame's primary + shadow rays with
    every CPU this process may use (OMP_NUM_THREADS when the host sets it):
    a row sample when the whole frame would take longer than budget_s, else
    the whole frame repeated until about budget_s has been spent.  Also the
    same trace on one thread (SingleThreadRenderer.cpp:8-33) on a row sample
    of about budget_s / 3, and the single-thread grid build of the frame
    (UniformGridAccStructBuilder::buildGeometry).  The grid is the reference's
    one-level grid at `lam`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    host = host_description()
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, env_threads or host["allowed_cpus"])
    v, nrm = frame.scene.download_geometry()
    t0 = time.perf_counter()
    fr = orc.grid_frame(v, lam)
    _, offs, refs = orc.grid_build(v, fr)
    build_s = time.perf_counter() - t0
    rays = orc.primary_rays(frame.grt.camera_dict(frame.cam))

    def run_rows(stride, n_threads, chunk_rows=16):
        rows = np.arange(0, HEIGHT, stride)
        idx = (rows[:, None] * WIDTH + np.arange(WIDTH)[None, :]).ravel()
        sub = {k: a[idx] for k, a in rays.items()} if stride > 1 else rays
        t = time.perf_counter()
        h = orc.trace_closest(fr, offs, refs, v, sub, threads=n_threads,
                              chunk=chunk_rows * WIDTH)
        sh = orc.gen_shadow(sub, h, nrm, LIGHT, OFFSET)
        orc.trace_any(fr, offs, refs, v, sh, threads=n_threads, chunk=chunk_rows * WIDTH)
        dt = time.perf_counter() - t
        return len(idx) + int((sh["tmin"] <= sh["tmax"]).sum()), dt, len(rows)

    # Probe on every 16th row in one-row chunks (enough chunks for every
    # thread) to size the sample.
    run_rows(64, threads, 1)
    n_probe, dt_probe, _ = run_rows(16, threads, 1)
    full_s = dt_probe * 16
    if full_s > budget_s:
        stride = int(np.ceil(full_s / budget_s))
        n_rays, spent, n_rows = run_rows(stride, threads)
        what = f"{n_rows} image rows (every {stride}th row)"
    else:
        reps = max(1, int(round(budget_s / max(full_s, 1e-3))))
        n_rays, spent = 0, 0.0
        for _ in range(reps):
            n, dt, n_rows = run_rows(1, threads)
            n_rays += n
            spent += dt
        what = f"the whole frame ({HEIGHT} rows) {reps} times"
    # One thread: a row sample sized from the multi-threaded rate.
    one_budget = budget_s / 3.0
    est_one_s = (spent / max(n_rays, 1)) * threads * (2 * WIDTH * HEIGHT)
    stride1 = max(1, int(np.ceil(est_one_s / one_budget)))
    n1, dt1, rows1 = run_rows(stride1, 1)
    return {"value": round(n_rays / spent / 1e6, 3), "unit": "Mrays/s", "cores": threads,
            "kind": "port",
            "single_thread": {"value": round(n1 / dt1 / 1e6, 3), "unit": "Mrays/s",
                              "sample": f"{rows1} image rows (every {stride1}th row), "
                                        f"{n1} rays in {dt1:.1f} s"},
            "host": {**host, "omp_num_threads_env": env_threads or None},
            "grid_build_single_thread_ms": round(build_s * 1e3, 1),
            "sample": (f"CPU oracle on {what}: primary closest-hit + shadow any-hit, "
                       f"{n_rays} rays in {spent:.1f} s on {threads} threads (every CPU "
                       f"the process may use, or OMP_NUM_THREADS when set), OpenMP "
                       f"schedule(dynamic) over 16-row chunks, same grid (lambda "
                       f"{lam:g}, one level); single-thread grid build of the frame "
                       f"{build_s * 1e3:.0f} ms")}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU only; with more GPUs the "
                          "rays would shard by image tile or request over replicated scenes, "
                          "which is not implemented here", "n_gpus": args.gpus}))
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
        elif args.mode == "kdtree":
            res = perf.kdtree_config(lam=args.lam or 256.0)
        else:
            res = perf.ao_config()
        print(json.dumps({"mode": args.mode, **res}))
        return
    sub_arg = None
    if args.sub is not None and args.sub != "off":
        l2, mr = args.sub.split(",")
        sub_arg = (float(l2), int(mr))
    frame = Frame(grt, torch, args.lam or LAMBDA, args.sched or "spsp", args.refill, sub_arg)
    if args.ray_lists is not None and (len(args.ray_lists) != 3 or
                                       set(args.ray_lists) - set("-c")):
        raise SystemExit("--ray-lists takes three letters of - and c (shadow, bounce, "
                         "bounce shadow)")
    if args.block:
        frame.ctx.set_block_size(args.block, args.block)
    frame.ctx.set_coop(args.coop)
    if args.copy_calibration:
        copy_calibration(torch)
    if args.lam:
        grids = [(args.lam, sub_arg)]
    else:
        grids = [(lam, None) for lam in LAMBDAS] + [(l1, (l2, mr)) for l1, l2, mr in TWO_LEVEL]
    lam, sub, sched, lists, tune_table = autotune(frame, grids, args.sched, args.ray_lists)
    frame.sched = sched
    frame.lists_kind = lists
    frame.set_grid(lam, sub)
    tag = grid_tag(lam, sub)
    for w in range(args.warmup):
        frame.run(1000 + w)
    frame.scene.check_grid()
    torch.cuda.synchronize()

    concurrent = not args.sequential and frame.concurrent_ok()
    for w in range(2):
        frame.run(1100 + w, concurrent=concurrent)
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(Frame.N_EVENTS)]
              for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        frame.run(s, events[s], concurrent=concurrent)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    frame.scene.check_grid()
    ms_per_step = wall * 1e3 / args.steps
    # Mean over the timed frames of each launch group (HIP events recorded on
    # the stream each launch runs on).  With the two branches concurrent, a
    # traversal launch's interval is its own duration while sharing the GPU.
    def span(ev, a, b):
        return ev[a].elapsed_time(ev[b])
    d = {k: float(np.mean([span(ev, a, b) for ev in events]))
         for k, (a, b) in {"transform": (0, 1), "build": (1, 2), "gen0": (2, 3), "primary": (3, 4),
                           "gen1": (4, 5), "shadow": (5, 6), "gen2": (6, 7), "bounce": (7, 8),
                           "gen3": (8, 9), "bounce_shadow": (9, 10), "shade": (10, 11)}.items()}
    if concurrent:
        d["gen2"] = 0.0    # the reflection rays are made before the branches split
    trace_ms = {k: d[k] for k in ("primary", "shadow", "bounce", "bounce_shadow")}
    stage_ms = {"transform": d["transform"], "build": d["build"],
                "raygen": d["gen0"] + d["gen1"] + d["gen2"] + d["gen3"],
                **trace_ms, "shade": d["shade"]}
    if args.profile:
        print(json.dumps({"mode": "profile", "grid": tag, "schedule": sched, "coop": args.coop,
                          "lists": frame.lists_kind, "concurrent_branches": concurrent,
                          "ms_per_step": round(ms_per_step, 3),
                          "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()}}))
        return

    rays = {"primary": frame.n, "shadow": frame.active(frame.shadow),
            "bounce": frame.active(frame.refl), "bounce_shadow": frame.active(frame.shadow2)}
    value = sum(rays.values()) / (ms_per_step * 1e-3) / 1e6
    info = frame.scene.grid_info()
    sub_info = frame.scene.subgrid_info() if sub is not None else None
    counters = stage_counters(frame, torch)

    # Roofline of the dominant traversal launch: the stage whose traversal
    # launch takes the most time per frame (one launch per stage and frame);
    # achieved = that stage's algorithmic bytes over its launch duration.
    dom_stage = max(trace_ms, key=trace_ms.get)
    si = [name for name, _ in STAGES].index(dom_stage)
    dom = kernel_name(STAGES[si][1], sched[si], sub is not None)
    dom_bytes = counters[dom_stage]["algorithmic_bytes"]
    achieved = dom_bytes / (trace_ms[dom_stage] * 1e-3) / 1e9
    # DRAM bytes of that launch at this lambda, from the committed
    # FETCH_SIZE / WRITE_SIZE profile (null if it has no such line).
    traffic_of = {name: profile_traffic(tag, name,
                                        kernel_name(closest, sched[i], sub is not None),
                                        "-" if i == 0 else frame.lists_kind[i - 1])
                  for i, (name, closest) in enumerate(STAGES)}
    traffic = traffic_of[dom_stage]
    per_stage_roof = {}
    for name, _ in STAGES:
        gbs = counters[name]["algorithmic_bytes"] / (trace_ms[name] * 1e-3) / 1e9
        per_stage_roof[name] = {"gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if traffic_of[name] is not None:
            mgbs = traffic_of[name] / (trace_ms[name] * 1e-3) / 1e9
            per_stage_roof[name].update({"measured_gbs": round(mgbs, 1),
                                         "measured_frac": round(mgbs / HBM_PEAK_GBS, 4)})
    ps_bytes = counters["primary"]["algorithmic_bytes"] + counters["shadow"]["algorithmic_bytes"]
    ps_ms = trace_ms["primary"] + trace_ms["shadow"]
    ps_gbs = ps_bytes / (ps_ms * 1e-3) / 1e9
    roof_ps = {"gbs": round(ps_gbs, 1), "frac": round(ps_gbs / HBM_PEAK_GBS, 4)}
    if traffic_of["primary"] is not None and traffic_of["shadow"] is not None:
        mgbs = (traffic_of["primary"] + traffic_of["shadow"]) / (ps_ms * 1e-3) / 1e9
        roof_ps.update({"measured_gbs": round(mgbs, 1),
                        "measured_frac": round(mgbs / HBM_PEAK_GBS, 4)})

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
                        "Phong shading of both levels into an RGBA8 frame, "
                        f"per-frame transform + device grid rebuild ({grid_desc(lam, sub)}, "
                        f"autotuned)",
            "rays_per_frame": rays,
            "grid": {"n": [int(x) for x in info["n"]], "cells": info["m"], "refs": info["r"],
                     "lambda": lam,
                     "subgrid": (None if sub_info is None else
                                 {"lambda2": sub[0], "min_refs": sub[1],
                                  "sub_cells": sub_info["cells"],
                                  "sub_refs": sub_info["refs"]})},
            "stage_ms": {k: round(v, 3) for k, v in stage_ms.items()},
            "value_definition": "every ray the frame traces -- closest-hit rays (primary, "
                                "mirror bounce) and shadow rays of both levels, inactive "
                                "slots excluded -- per second of whole-frame time "
                                "(transform, grid rebuild, ray generation, traversal, "
                                "Phong shading)",
            "mrays_primary_shadow_stages": round((rays["primary"] + rays["shadow"]) /
                                                 (ps_ms * 1e-3) / 1e6, 2),
            "mrays_primary_shadow_per_frame_time": round((rays["primary"] + rays["shadow"]) /
                                                         (ms_per_step * 1e-3) / 1e6, 2),
            "roofline_primary_shadow": roof_ps,
            "roofline_by_stage": per_stage_roof,
            "stage_counters": counters,
            "traversal": {"schedule_per_stage": sched, "refill": args.refill,
                          "block": args.block or 128, "grid": tag, "coop": args.coop,
                          "secondary_ray_lists": frame.lists_kind,
                          "concurrent_branches": concurrent},
            "stage_ms_note": ("shadow and bounce branches run on two streams: a traversal "
                              "stage's time is its launch duration while sharing the GPU, "
                              "and the stages overlap" if concurrent else
                              "stages run one after another on one stream"),
            "autotune_frame_ms_by_grid": tune_table,
            "parallelism": "single GPU",
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "measured_frac": (None if traffic is None else
                                       round(traffic / (trace_ms[dom_stage] * 1e-3) / 1e9 /
                                             HBM_PEAK_GBS, 4)),
                     "kernel": "grt::" + dom, "stage": dom_stage,
                     "avg_launch_ms": round(trace_ms[dom_stage], 3),
                     "algorithmic_bytes_per_launch": dom_bytes},
    }
    if sub is not None:
        # The same roofline on the fastest one-level grid of the tuning table:
        # the reference's structure, more bytes per ray, a longer frame.
        one = {k: e for k, e in tune_table.items() if e.get("sub") is None}
        if one:
            tag1 = min(one, key=lambda k: one[k].get("frame_ms_final", one[k]["frame_ms"]))
            e1 = one[tag1]
            frame.set_grid(e1["lam"], None)
            frame.sched, frame.lists_kind = e1["sched"], e1["lists"]
            ms1 = frame.time_stages(frames=6)
            c1 = stage_counters(frame, torch)
            st1 = {"primary": ms1[2], "shadow": ms1[3]}
            b1 = c1["primary"]["algorithmic_bytes"] + c1["shadow"]["algorithmic_bytes"]
            gbs1 = b1 / ((st1["primary"] + st1["shadow"]) * 1e-3) / 1e9
            out["config"]["one_level_comparison"] = {
                "grid": tag1, "schedule_per_stage": e1["sched"], "lists": e1["lists"],
                "frame_ms": round(sum(ms1), 3),
                "frame_ms_wall": round(frame.time_frames(5, concurrent and
                                                         frame.concurrent_ok()), 3),
                "stage_ms": {n: round(v, 3) for n, v in zip(frame.stage_names, ms1)},
                "primary_shadow_algorithmic_bytes": b1,
                "roofline_primary_shadow": {"gbs": round(gbs1, 1),
                                            "frac": round(gbs1 / HBM_PEAK_GBS, 4)},
                "mean_cells_tests": {n: [c1[n]["mean_cells"], c1[n]["mean_tests"]]
                                     for n in c1},
                "note": "stage times include ray generation and list compaction, as in "
                        "the tuning table; frame_ms_wall: whole frames timed like the "
                        "benchmarked ones (same branch concurrency when the schedule "
                        "allows it)"}
            frame.set_grid(lam, sub)
            frame.sched, frame.lists_kind = sched, lists
    if not args.no_serve:
        # p50/p99 request latency, continuous batching vs the plain path, on
        # the serving workload (16 resident sphere scenes, mixed requests).
        from grt import perf
        sctx = grt.Context(0)   # the serving run's own schedule and buffers
        scenes, tris = perf.build_serve_scenes(sctx)
        lat = perf.serve_vs_plain(sctx, scenes, n_requests=args.serve_requests)
        lat["triangles"] = tris
        out["latency"] = lat
    if not args.no_cpu_baseline:
        # the reference's one-level grid: at the tuned lambda, or for a
        # two-level pick at the fastest one-level lambda of the tuning table
        one = {k: e for k, e in tune_table.items() if e.get("sub") is None and "/" not in k}
        cpu_lam = lam if sub is None else (
            min(one.values(), key=lambda e: e.get("frame_ms_final", e["frame_ms"]))["lam"]
            if one and all("lam" in e for e in one.values()) else 256.0)
        out["cpu_baseline"] = cpu_baseline(frame, torch, args.cpu_seconds, cpu_lam)
    print(json.dumps(out))


def grid_desc(lam, sub):
    if sub is None:
        return f"one level, lambda={lam:g}"
    return (f"two levels: top lambda={lam:g}, sub-grids at lambda2={sub[0]:g} in cells of "
            f">= {sub[1]} references")


def profile_traffic(tag, stage, kernel, listed):
    """HBM bytes per launch of a stage's traversal kernel on this grid (its
    grid_tag), whole image ('-') or compacted list ('c'), from the committed
    FETCH_SIZE / WRITE_SIZE profile (None when the profile has no such line)."""
    tpath = os.path.join(ROOT, "profiles", "trace_stadium_traffic.md")
    if not os.path.exists(tpath):
        return None
    key = "hbm_bytes_per_launch[grid=%s][%s][%s][list=%s]:" % (tag, stage, kernel, listed)
    for line in open(tpath):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def copy_calibration(torch, n=1 << 28, reps=3):
    """A plain device copy of n float32 (1 GiB read + 1 GiB write), so a
    counter run can scale FETCH_SIZE / WRITE_SIZE to bytes."""
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y


if __name__ == "__main__":
    main()
