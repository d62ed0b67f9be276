"""Binary32 traversal against float64 brute force, per workload (CPU only).

The HIP kernels equal the oracle bit for bit (tests/test_gpu_parity.py), so the
oracle's error against float64 brute-force intersection is the HIP path's
error.  For sampled rays of each workload:

- closest hit: a ray whose triangle differs from the float64 nearest must be a
  near-tie (second-nearest float64 t within 1e-6 relative of the nearest) or
  graze an edge (a float64 barycentric within 1e-5 of zero); both kinds are
  counted.  On equal triangles the relative t error and the absolute u, v
  error are measured (t is asserted within 1e-5 relative; u and v of small,
  distant triangles lose more to cancellation in binary32 and are reported);
- any hit: a flag that differs from float64 must have its nearest float64
  occluder within 1e-5 relative of the ray's tmax.

The figures are printed (pytest -s) and quoted in the design notes.
"""
import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import SEED_C1, SEED_C2, rng

STADIUM = (256, 4096, (1 << 20) - 2)
STADIUM_CAM = ((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0.0, 1.0, 0.0), 60.0, 1920, 1080)
C1_CAM = ((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 60.0, 512, 512)


def _sample(rays, idx):
    return {k: np.ascontiguousarray(a[idx]) for k, a in rays.items()}


def _edge(u, v):
    w = np.stack([u, v, 1 - u - v], 1)
    return np.abs(w).min(1) < 1e-5


def closest_report(v, f, offs, refs, rays):
    o = orc.trace_closest(f, offs, refs, v, rays)
    b = orc.brute_closest_f64(v, rays)
    mism = o["id"] != b["id"]
    with np.errstate(invalid="ignore"):
        tie = np.abs(b["t2"] - b["t1"]) <= 1e-6 * np.abs(b["t1"])
    edge = (b["id"] >= 0) & _edge(b["u"], b["v"])
    ok = (o["id"] >= 0) & ~mism
    t1 = b["t1"][ok]
    rel = np.abs(o["t"][ok] - t1) / np.abs(t1)
    duv = np.maximum(np.abs(o["u"][ok] - b["u"][ok]), np.abs(o["v"][ok] - b["v"][ok]))
    return o, {"rays": int(len(mism)), "hits": int((b["id"] >= 0).sum()),
               "id_mismatch": int(mism.sum()), "near_tie": int((mism & tie).sum()),
               "edge": int((mism & edge & ~tie).sum()),
               "unexplained": int((mism & ~tie & ~edge).sum()),
               "t_rel_max": float(rel.max()) if rel.size else 0.0,
               "t_rel_p99": float(np.quantile(rel, 0.99)) if rel.size else 0.0,
               "uv_abs_max": float(duv.max()) if duv.size else 0.0}


def any_report(v, f, offs, refs, rays):
    o = orc.trace_any(f, offs, refs, v, rays)
    b = orc.brute_any_f64(v, rays)
    mism = o["occluded"] != b["occluded"]
    tn = b["t_near"]
    bound = np.isfinite(tn) & (np.abs(tn - rays["tmax"]) <= 1e-5 * np.abs(tn))
    return {"rays": int(len(mism)), "occluded": int(b["occluded"].sum()),
            "flag_mismatch": int(mism.sum()), "at_tmax": int((mism & bound).sum()),
            "unexplained": int((mism & ~bound).sum())}


def _check(rep):
    assert rep["unexplained"] == 0, rep
    if "t_rel_max" in rep:
        assert rep["t_rel_max"] < 1e-5, rep


def test_fp64_error_c1():
    """512x512 spheres scene (lambda 5): every 16th pixel's primary ray and its
    shadow ray to a point light."""
    v, n = grt.gen_spheres(SEED_C1, 64, 0.02, 0.08)
    f = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, f)
    rays = _sample(orc.primary_rays(orc.camera_lookat(*C1_CAM)), slice(None, None, 16))
    hits, prim = closest_report(v, f, offs, refs, rays)
    sh = orc.gen_shadow(rays, hits, n, (3.0, 4.0, -2.0), 1e-4)
    shadow = any_report(v, f, offs, refs, sh)
    print("\nC1 primary", prim, "\nC1 shadow", shadow)
    _check(prim)
    _check(shadow)


def test_fp64_error_stadium():
    """2^20-triangle stadium (lambda 256, the tuned frame's grid): 2048 random
    pixels' primary rays, their shadow rays and their mirror bounces."""
    v, n, _ = grt.gen_stadium(SEED_C2, *STADIUM)
    f = orc.grid_frame(v, 256.0)
    _, offs, refs = orc.grid_build(v, f)
    cam = orc.camera_lookat(*STADIUM_CAM)
    idx = np.sort(rng(5).choice(1920 * 1080, 2048, replace=False))
    rays = _sample(orc.primary_rays(cam), idx)
    hits, prim = closest_report(v, f, offs, refs, rays)
    sh = orc.gen_shadow(rays, hits, n, (1000.0, 1000.0, 1000.0), 1e-4)
    shadow = any_report(v, f, offs, refs, sh)
    refl = orc.gen_reflection(rays, hits, n, 1e-4)
    _, bounce = closest_report(v, f, offs, refs, refl)
    print("\nstadium primary", prim, "\nstadium shadow", shadow, "\nstadium bounce", bounce)
    _check(prim)
    _check(shadow)
    _check(bounce)


def _merge(reports):
    out = {}
    for r in reports:
        for k, x in r.items():
            if k.startswith("t_rel") or k == "uv_abs_max":
                out[k] = max(out.get(k, 0.0), x)
            else:
                out[k] = out.get(k, 0) + x
    return out


def _scene_rows(v, n, lam, cam, n_rays, seed, light=(1000.0, 1000.0, 1000.0)):
    """Primary, shadow and mirror-bounce reports on n_rays random pixels."""
    f = orc.grid_frame(v, lam)
    _, offs, refs = orc.grid_build(v, f)
    rays_all = orc.primary_rays(grt.camera_dict(cam))
    idx = np.sort(rng(seed).choice(len(rays_all["ox"]), n_rays, replace=False))
    rays = _sample(rays_all, idx)
    hits, prim = closest_report(v, f, offs, refs, rays)
    sh = orc.gen_shadow(rays, hits, n, light, 1e-4)
    shadow = any_report(v, f, offs, refs, sh)
    refl = orc.gen_reflection(rays, hits, n, 1e-4)
    _, bounce = closest_report(v, f, offs, refs, refl)
    return prim, shadow, bounce


def test_fp64_error_batch64():
    """The 64 batched scenes (2k..500k triangles, lambda 2, the tuned batch
    grid): 48 random pixels per scene of its own square image, with their
    shadow rays and mirror bounces."""
    from grt import workloads as W
    rows = [[], [], []]
    for i, (tris, side, _depth) in enumerate(W.batch_jobs()):
        v, n = W.sphere_scene(W.SEED_BATCH + 1000 + i, tris)
        for r, rep in zip(rows, _scene_rows(v, n, 2.0, W.camera(side, side), 48, i)):
            r.append(rep)
    prim, shadow, bounce = (_merge(r) for r in rows)
    print("\nbatch-64 primary", prim, "\nbatch-64 shadow", shadow, "\nbatch-64 bounce", bounce)
    _check(prim)
    _check(shadow)
    _check(bounce)


def test_fp64_error_serve16():
    """The 16 resident serving scenes (10k..4M triangles, lambda 8): 64 random
    pixels of a 1920 x 1080 request per scene, with shadow rays and bounces."""
    from grt import workloads as W
    rows = [[], [], []]
    for i, nt in enumerate(W.serve_scene_sizes()):
        v, n = W.sphere_scene(W.SEED_SERVE + 100 + i, nt)
        for r, rep in zip(rows, _scene_rows(v, n, 8.0, W.camera(1920, 1080), 64, 50 + i)):
            r.append(rep)
    prim, shadow, bounce = (_merge(r) for r in rows)
    print("\nserve-16 primary", prim, "\nserve-16 shadow", shadow, "\nserve-16 bounce", bounce)
    _check(prim)
    _check(shadow)
    _check(bounce)


def test_fp64_error_ao16m():
    """2^24 triangles (lambda 4, the tuned AO grid): 256 random pixels' primary
    rays of the 1024 x 512 frame, and 512 random AO rays of the frame's
    8 per pixel (tmax = 0.05 x scene extent)."""
    from grt import workloads as W
    v, n = W.ao_scene()
    f = orc.grid_frame(v, 4.0)
    _, offs, refs = orc.grid_build(v, f)
    rays_all = orc.primary_rays(grt.camera_dict(W.ao_camera(1024, 512)))
    idx = np.sort(rng(21).choice(len(rays_all["ox"]), 256, replace=False))
    _, prim = closest_report(v, f, offs, refs, _sample(rays_all, idx))
    hits_all = orc.trace_closest(f, offs, refs, v, rays_all)
    extent = float(np.max(np.asarray(f["bmax"]) - np.asarray(f["bmin"])))
    ao_all = orc.gen_ao(rays_all, hits_all, n, 8, W.SEED_AO, 1e-4, 0.05 * extent)
    aidx = np.sort(rng(22).choice(len(ao_all["ox"]), 512, replace=False))
    ao = any_report(v, f, offs, refs, _sample(ao_all, aidx))
    print("\nao-16M primary", prim, "\nao-16M AO", ao)
    _check(prim)
    _check(ao)
