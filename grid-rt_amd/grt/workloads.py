"""Seeded workloads of the benchmark configurations.

  spheres-16k   64 UV-spheres x 256 triangles, 512x512 primary + shadow
  stadium-1M    2^20 triangles in 256 Gaussian clusters + a ground plane
                100x the cluster extent (bench.py's frame)
  batch-64      64 sphere scenes, 2k..500k triangles (log-uniform), square
                images of 64..1024 pixels a side (log-uniform), depth 1..3
  serve-16      16 resident sphere scenes of 10k..4M triangles (log-uniform);
                requests draw a scene uniformly, a resolution from
                {256^2, 512^2, 1024^2, 1920x1080} and 1..3 bounces; Poisson
                arrivals
  ao-16M        4096 clusters x 16 spheres x 256 = 2^24 triangles, 8 cosine-
                weighted ambient-occlusion rays per pixel at 1024x512

Scene geometry comes from libgrt's generators (splitmix64 -> xoshiro256**);
workload parameters (sizes, resolutions, arrival gaps) are drawn with numpy's
PCG64 from the same seeds.
"""
from __future__ import annotations

import math

import numpy as np

from . import camera_lookat, gen_sphere_clusters, gen_spheres

SEED_SPHERES = 0x6772740000000001
SEED_STADIUM = 0x6772740000000002
SEED_BATCH = 0x6772740000000003
SEED_SERVE = 0x6772740000000004
SEED_AO = 0x6772740000000005

EYE, CENTER, UP = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0)
FOVY = 60.0
LIGHT = (1000.0, 1000.0, 1000.0)


def sphere_scene(seed: int, n_tris: int):
    """Sphere scene of about n_tris triangles with the spheres-16k density:
    radii U(0.02, 0.08) scaled by cbrt(64 / spheres)."""
    n_spheres = max(1, int(round(n_tris / 256)))
    s = (64.0 / n_spheres) ** (1.0 / 3.0)
    return gen_spheres(seed, n_spheres, 0.02 * s, 0.08 * s)


def camera(width: int, height: int):
    return camera_lookat(EYE, CENTER, UP, FOVY, width, height)


def log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def batch_jobs(seed: int = SEED_BATCH, n: int = 64):
    """(n_tris, side, depth) of the batched configuration."""
    rng = np.random.default_rng(seed)
    tris = log_uniform(rng, 2_000, 500_000, n).astype(np.int64)
    side = np.round(log_uniform(rng, 64, 1024, n)).astype(np.int64)
    depth = rng.integers(1, 4, n)
    return [(int(t), int(s), int(d)) for t, s, d in zip(tris, side, depth)]


SERVE_RESOLUTIONS = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]


def serve_scene_sizes(seed: int = SEED_SERVE, n: int = 16):
    rng = np.random.default_rng(seed)
    return [int(x) for x in log_uniform(rng, 10_000, 4_000_000, n)]


def serve_requests(n: int, rate_hz: float, n_scenes: int, seed: int = SEED_SERVE + 1):
    """(scene index, width, height, depth, arrival_s): Poisson arrivals at
    rate_hz (rate_hz <= 0: all at t = 0); depth = 1 + bounces, bounces 1..3."""
    rng = np.random.default_rng(seed)
    if rate_hz > 0:
        t = np.cumsum(rng.exponential(1.0 / rate_hz, n))
        t -= t[0]
    else:
        t = np.zeros(n)
    scene = rng.integers(0, n_scenes, n)
    res = rng.integers(0, len(SERVE_RESOLUTIONS), n)
    bounces = rng.integers(1, 4, n)
    return [(int(scene[i]), SERVE_RESOLUTIONS[res[i]][0], SERVE_RESOLUTIONS[res[i]][1],
             int(1 + bounces[i]), float(t[i])) for i in range(n)]


def ao_scene(seed: int = SEED_AO, clusters: int = 4096, per_cluster: int = 16):
    """2^24 triangles: clustered spheres (sigma 0.02, radii U(0.003, 0.012))."""
    return gen_sphere_clusters(seed, clusters, per_cluster, 0.02, 0.003, 0.012)


def ao_camera(width: int, height: int):
    """From the middle of the scene box's front face (z = 0) looking in: every
    primary ray crosses the whole box, so every pixel has a hit to start AO rays
    from."""
    return camera_lookat((0.5, 0.5, 0.0), (0.5, 0.5, 1.0), UP, FOVY, width, height)


def percentile_ms(lat_s, q):
    return float(np.percentile(np.asarray(lat_s) * 1e3, q)) if len(lat_s) else float("nan")
