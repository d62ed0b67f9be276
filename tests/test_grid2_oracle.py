"""CPU pins of the oracle's two-level grid (orc_grid2_plan / count / fill and
orc_grid2_trace_closest / any): the sub-grid of every dense top cell from its
geometry, the resolution rule from the mathematics it states, every empty
cell's skip radius from a brute-force scan of the cubes around it, and the
traces against float64 brute force over all triangles and the one-level
oracle.
The GPU suite then holds the HIP build and traversal to this oracle bit for
bit (tests/test_gpu_grid2.py)."""
import math

import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import SEED_C1, rng

LAM, LAM2, MIN_REFS, RMAX = 0.5, 4.0, 8, 16
SKIP_MAX = 7
EMPTY = np.uint32(0x80000000)


def has_subgrid(cs):
    """Top cells with a sub-grid (plan word non-zero, without the empty-cell flag)."""
    return (cs[:, 1] != 0) & ((cs[:, 1] & EMPTY) == 0)


@pytest.fixture(scope="module")
def two_level():
    v, _ = grt.gen_spheres(SEED_C1, 16, 0.05, 0.15)
    fr = orc.grid_frame(v, LAM)
    _, offs, refs = orc.grid_build(v, fr)
    g2 = orc.grid2_build(v, fr, offs, refs, MIN_REFS, LAM2, RMAX)
    return v, fr, offs, refs, g2


def _cell_coords(fr, c):
    nx, ny = int(fr["n"][0]), int(fr["n"][1])
    return c % nx, (c // nx) % ny, c // (nx * ny)


def test_resolution_is_the_smallest_per_axis(two_level):
    """A top cell of n >= min_refs references gets, per axis, the smallest
    r with r x (V / (lambda2 n))^(1/3) >= cell extent (capped at 16), i.e.
    sub-cells about as large as lambda2 n cubes filling the cell; cells below
    min_refs, and cells where that gives 1 x 1 x 1, have no sub-grid."""
    v, fr, offs, refs, (cs, so, sr) = two_level
    cell = [float(x) for x in fr["cell"]]
    vol = cell[0] * cell[1] * cell[2]
    counts = np.diff(offs.astype(np.int64))
    dense = 0
    for c in range(int(fr["m"])):
        n = int(counts[c])
        p = int(cs[c, 1])
        if n == 0:
            assert p >> 31 == 1 and cs[c, 0] == 0, c
            continue
        if n < MIN_REFS:
            assert p == 0 and cs[c, 0] == 0, c
            continue
        f = (LAM2 * n / vol) ** (1.0 / 3.0)
        want = [cell[a] * f for a in range(3)]
        if p == 0:
            assert all(w <= 1.0 + 1e-5 for w in want), (c, want)
            continue
        dense += 1
        r = [p & 255, (p >> 8) & 255, (p >> 16) & 255]
        for a in range(3):
            if r[a] == RMAX:
                assert want[a] > (RMAX - 1) * (1 - 1e-5), (c, a, r, want)
            else:
                assert r[a] >= want[a] * (1 - 1e-5) and r[a] - 1 < want[a] * (1 + 1e-5), \
                    (c, a, r, want)
    assert dense > 20


def test_sub_cells_hold_the_triangles_that_overlap_them(two_level):
    """Every sub-grid spans its top cell, sub-cells list ascending unique ids,
    and a triangle of the top cell is listed in exactly the sub-cells its
    float64 AABB overlaps (up to binary32 rounding at sub-cell boundaries,
    clamped into the sub-grid like the top grid's voxel rule)."""
    v, fr, offs, refs, (cs, so, sr) = two_level
    v9 = v.reshape(-1, 9).astype(np.float64)
    bmin = np.array(fr["bmin"], np.float64)
    cell = np.array(fr["cell"], np.float64)
    assert so[0] == 0 and np.all(np.diff(so.astype(np.int64)) >= 0) and so[-1] == len(sr)
    checked = 0
    for c in np.nonzero(has_subgrid(cs))[0]:
        p = int(cs[c, 1])
        r = np.array([p & 255, (p >> 8) & 255, (p >> 16) & 255])
        base = int(cs[c, 0])
        corner = bmin + np.array(_cell_coords(fr, int(c)), np.float64) * cell
        size = cell / r
        tris = refs[offs[c]:offs[c + 1]]
        listed = {int(t): set() for t in tris}
        m = int(r.prod())
        for s in range(m):
            ids = sr[so[base + s]:so[base + s + 1]]
            assert np.all(np.diff(ids.astype(np.int64)) > 0), (c, s)
            for t in ids:
                assert int(t) in listed, (c, s, int(t))
                listed[int(t)].add(s)
        eps = 1e-4 * size
        for t in tris:
            x = v9[t].reshape(3, 3)
            mn, mx = x.min(0), x.max(0)
            def rng_idx(lo_v, hi_v):
                lo = np.clip(np.floor((lo_v - corner) / size), 0, r - 1).astype(int)
                hi = np.clip(np.floor((hi_v - corner) / size), 0, r - 1).astype(int)
                return lo, hi
            ilo, ihi = rng_idx(mn + eps, mx - eps)     # certainly covered
            olo, ohi = rng_idx(mn - eps, mx + eps)     # possibly covered
            got = listed[int(t)]
            for s in got:
                q = np.array([s % r[0], (s // r[0]) % r[1], s // (r[0] * r[1])])
                assert np.all(q >= olo) and np.all(q <= ohi), (c, int(t), q, olo, ohi)
            for qz in range(ilo[2], ihi[2] + 1):
                for qy in range(ilo[1], ihi[1] + 1):
                    for qx in range(ilo[0], ihi[0] + 1):
                        assert qx + r[0] * (qy + r[1] * qz) in got, (c, int(t), qx, qy, qz)
            checked += 1
    assert checked > 1000


def _skip_radii_brute(occ):
    """Largest r <= SKIP_MAX with the (2r+1)^3 cube around each cell empty
    (cells outside the grid are empty), by scanning growing cubes."""
    n2, n1, n0 = occ.shape
    pad = SKIP_MAX + 1
    big = np.zeros((n2 + 2 * pad, n1 + 2 * pad, n0 + 2 * pad), bool)
    big[pad:pad + n2, pad:pad + n1, pad:pad + n0] = occ
    r = np.full(occ.shape, -1, np.int64)
    alive = ~occ
    for k in range(SKIP_MAX + 1):
        hit = np.zeros(occ.shape, bool)
        for dz in range(-k, k + 1):
            for dy in range(-k, k + 1):
                for dx in range(-k, k + 1):
                    if max(abs(dz), abs(dy), abs(dx)) != k:
                        continue
                    hit |= big[pad + dz:pad + dz + n2, pad + dy:pad + dy + n1,
                               pad + dx:pad + dx + n0]
        alive &= ~hit
        r[alive] = k
    return r


@pytest.mark.parametrize("scene", ["spheres", "sparse"])
def test_skip_radii_are_the_largest_empty_cubes(two_level, scene):
    """Every empty top cell carries the largest r <= 7 such that no cell within
    r cells of it on every axis holds a reference, checked against a scan of
    the growing cubes around every cell; occupied cells carry no radius."""
    if scene == "spheres":
        v, fr, offs, refs, (cs, _, _) = two_level
    else:
        v, _ = grt.gen_spheres(SEED_C1 + 3, 3, 0.02, 0.05)
        fr = orc.grid_frame(v, 4.0)
        _, offs, refs = orc.grid_build(v, fr)
        cs, _, _ = orc.grid2_build(v, fr, offs, refs, MIN_REFS, LAM2, RMAX)
    n0, n1, n2 = (int(x) for x in fr["n"])
    counts = np.diff(offs.astype(np.int64))
    occ = (counts > 0).reshape(n2, n1, n0)
    want = _skip_radii_brute(occ).ravel()
    empty = counts == 0
    assert np.all((cs[empty, 1] & EMPTY) != 0)
    got = (cs[empty, 1] & 0xff).astype(np.int64)
    assert np.array_equal(got, want[empty])
    assert np.all((cs[~empty, 1] & EMPTY) == 0)
    # radius 0 next to geometry; the cap in the sparse scene's open space
    assert (got == 0).any() and got.max() >= (SKIP_MAX if scene == "sparse" else 4), \
        np.bincount(got)


def _rays(n, seed, spread=3.0):
    r = rng(seed)
    o = ((r.random((n, 3)) - 0.5) * spread + 0.5).astype(np.float32)
    target = r.random((n, 3)).astype(np.float32)
    d = (target - o).astype(np.float32)
    d[::7] = r.normal(size=(len(d[::7]), 3))
    return {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
            "dz": d[:, 2], "tmin": np.full(n, 2e-4, np.float32),
            "tmax": np.full(n, np.inf, np.float32)}


def test_two_level_closest_matches_brute_force_and_one_level(two_level):
    """Closest hits through both levels equal float64 brute force over all
    triangles (rays whose two nearest hits lie within 1e-6 relative, or that
    graze an edge, are excused) and the one-level oracle's except on such
    near-ties; t within 1e-5 relative."""
    v, fr, offs, refs, g2 = two_level
    rays = _rays(6000, 41)
    o = orc.grid2_trace_closest(fr, offs, refs, g2, v, rays, counters=True)
    b = orc.brute_closest_f64(v, rays)
    assert (b["id"] >= 0).sum() > 1000
    with np.errstate(invalid="ignore"):   # misses: inf - inf
        near_tie = np.abs(b["t2"] - b["t1"]) <= 1e-6 * np.abs(b["t1"])
    w = np.stack([b["u"], b["v"], 1 - b["u"] - b["v"]], 1)
    excused = near_tie | (np.abs(w).min(1) < 1e-5)
    bad = (o["id"] != b["id"]) & ~excused
    assert bad.sum() == 0, np.nonzero(bad)[0][:10]
    ok = (o["id"] >= 0) & (o["id"] == b["id"])
    assert np.max(np.abs(o["t"][ok] - b["t1"][ok]) / b["t1"][ok]) < 1e-5
    o1 = orc.trace_closest(fr, offs, refs, v, rays, counters=True)
    diff = (o["id"] != o1["id"]) | (o["t"].view(np.uint32) != o1["t"].view(np.uint32))
    assert (diff & ~excused).sum() == 0
    # the second level visits more (smaller) cells and tests fewer triangles
    assert o["tests"].sum() < o1["tests"].sum()


def test_two_level_any_matches_brute_force(two_level):
    v, fr, offs, refs, g2 = two_level
    rays = _rays(6000, 43)
    rays["tmin"][:] = 0.0
    rays["tmax"] = (rng(5).random(6000) * 2).astype(np.float32)
    o = orc.grid2_trace_any(fr, offs, refs, g2, v, rays)
    b = orc.brute_any_f64(v, rays)
    tn = b["t_near"]
    boundary = np.isfinite(tn) & (np.abs(tn - rays["tmax"]) <= 1e-5 * np.abs(tn))
    assert ((o["occluded"] != b["occluded"]) & ~boundary).sum() == 0
    assert 100 < o["occluded"].sum() < 5900
    one = orc.trace_any(fr, offs, refs, v, rays)["occluded"]
    assert math.isclose(float((o["occluded"] == one).mean()), 1.0)


def test_rays_lying_in_a_subgrid_cell_plane(two_level):
    """Rays with a zero direction component whose origin lies exactly on the
    lower boundary plane of a cell with a sub-grid (binary32 `cx cell + bmin`):
    every plane t on that axis is 0 x inf = NaN or infinite, so the walk must
    neither cross nor drop the axis.  Aimed through triangles straddling the
    plane, their closest hits equal the one-level oracle's and float64 brute
    force."""
    v, fr, offs, refs, (cs, so, sr) = two_level
    f32 = np.float32
    v9 = v.reshape(-1, 9)
    nx, ny = int(fr["n"][0]), int(fr["n"][1])
    cell, bmin = np.asarray(fr["cell"], f32), np.asarray(fr["bmin"], f32)
    ox, oy, oz, dz = [], [], [], []
    for c in np.nonzero(has_subgrid(cs))[0]:
        cx = int(c) % nx
        if cx == 0:
            continue
        plane = f32(f32(cx) * cell[0]) + bmin[0]
        for t in refs[offs[c]:offs[c + 1]]:
            tri = v9[t].reshape(3, 3)
            if not (tri[:, 0].min() < plane < tri[:, 0].max()):
                continue
            ox.append(plane)
            oy.append(tri[:, 1].mean())
            oz.append(tri[:, 2].min() - f32(1.0))
            dz.append(f32(1.0))
        if len(ox) > 400:
            break
    n = len(ox)
    assert n > 50
    rays = {"ox": np.array(ox, f32), "oy": np.array(oy, f32), "oz": np.array(oz, f32),
            "dx": np.zeros(n, f32), "dy": np.zeros(n, f32), "dz": np.array(dz, f32),
            "tmin": np.zeros(n, f32), "tmax": np.full(n, np.inf, f32)}
    o = orc.grid2_trace_closest(fr, offs, refs, (cs, so, sr), v, rays)
    o1 = orc.trace_closest(fr, offs, refs, v, rays)
    b = orc.brute_closest_f64(v, rays)
    assert (o1["id"] >= 0).sum() > n // 4
    assert np.array_equal(o["id"], o1["id"]) and np.array_equal(o["t"].view(np.uint32),
                                                              o1["t"].view(np.uint32))
    with np.errstate(invalid="ignore"):
        near_tie = np.abs(b["t2"] - b["t1"]) <= 1e-6 * np.abs(b["t1"])
    assert ((o["id"] != b["id"]) & ~near_tie).sum() == 0
