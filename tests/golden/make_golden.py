"""Writes tests/golden/c1_small.npz: a small seeded scene of the config-1
generator (6 UV-spheres, 1536 triangles) with the oracle's grid, primary /
shadow / reflection / ambient-occlusion rays and their results.

The reference ships no tests or golden vectors for this path, so these
fixtures freeze the oracle's answers on inputs the library's own generators
make; test_golden.py checks the oracle (CPU) and the HIP path (GPU) against
them.  Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "grid-rt_amd"), os.path.join(ROOT, "oracle")]

import grt  # noqa: E402  (host-only generators and camera)
import pyoracle as orc  # noqa: E402

SEED = 0x6772740000000001
LAMBDA = 5.0
WIDTH, HEIGHT = 32, 24
LIGHT = (1000.0, 1000.0, 1000.0)
OFFSET = 1e-3
AO_SPP, AO_SEED, AO_TMAX = 4, 77, 0.5


def make():
    v, nrm = grt.gen_spheres(SEED, 6, 0.2, 0.35)
    fr = orc.grid_frame(v, LAMBDA)
    counts, offs, refs = orc.grid_build(v, fr)
    cam = grt.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0, 1, 0), 60.0, WIDTH, HEIGHT)
    rays = orc.primary_rays(grt.camera_dict(cam))
    hit = orc.trace_closest(fr, offs, refs, v, rays, counters=True)
    sh = orc.gen_shadow(rays, hit, nrm, LIGHT, OFFSET)
    occ = orc.trace_any(fr, offs, refs, v, sh)["occluded"]
    rf = orc.gen_reflection(rays, hit, nrm, OFFSET)
    rhit = orc.trace_closest(fr, offs, refs, v, rf)
    ao = orc.gen_ao(rays, hit, nrm, AO_SPP, AO_SEED, OFFSET, AO_TMAX)
    aocc = orc.trace_any(fr, offs, refs, v, ao)["occluded"]
    out = {"verts": v, "normals": nrm, "counts": counts, "offsets": offs, "refs": refs,
           "frame_f": np.concatenate([np.float32(fr[k]) for k in ("bmin", "bmax", "cell",
                                                                   "inv_cell")]),
           "frame_n": np.int32(fr["n"]),
           "camera": np.concatenate([np.float32(list(getattr(cam, k)))
                                     for k in ("pos", "base", "nu", "nv")]),
           "hit_counters": np.stack([hit["cells"], hit["tests"]], 1).astype(np.uint32),
           "shadow_occluded": occ, "ao_occluded": aocc}
    for k in ("id", "t", "u", "v"):
        out["hit_" + k] = hit[k]
        out["refl_hit_" + k] = rhit[k]
    for k in orc.RAY_FIELDS:
        out["ray_" + k] = rays[k]
        out["shadow_" + k] = sh[k]
        out["refl_" + k] = rf[k]
        out["ao_" + k] = ao[k]
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "c1_small.npz"), **make())
    print("wrote", os.path.join(HERE, "c1_small.npz"))
