"""Pins for the CPU oracle (oracle/grt_oracle.c), CPU only.

The reference ships no tests or golden vectors for this path and cannot be
compiled here (it needs the un-vendored vrbase library), so every oracle
function is pinned by checks that do not restate its formula: closed forms,
known answers, FP64 brute force and geometric invariants.  Each check is
chosen so that a plausible slip (swapped u/v, dropped cell-exit bound, wrong
truncation, wrong step order) fails it.
"""
import math

import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import SEED_C1, SEED_C2, rng


# ---------------------------------------------------------------------------
# a2: frame
# ---------------------------------------------------------------------------
def test_cbrtf_is_correctly_rounded_cube_root():
    xs = np.concatenate([np.geomspace(1e-30, 1e30, 20000), rng(1).random(5000) * 1e4,
                         [1.0, 8.0, 27.0, 64.0, 125.0, 1e6, 0.125]]).astype(np.float32)
    got = np.array([orc.cbrtf(float(x)) for x in xs], np.float32)
    exact = np.cbrt(xs.astype(np.float64)).astype(np.float32)
    # exact cubes
    for c in (1.0, 2.0, 3.0, 4.0, 5.0, 100.0, 0.5):
        assert orc.cbrtf(float(np.float32(c) ** 3)) == np.float32(c)
    ulp = np.abs(got.view(np.int32).astype(np.int64) - exact.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    assert (ulp == 0).mean() > 0.999


@pytest.mark.parametrize("lam", [2.0, 5.0, 8.0])
def test_frame_closed_form(lam):
    r = rng(int(lam))
    for _ in range(200):
        mn = (r.random(3) * 10 - 5).astype(np.float32)
        ext = (r.random(3) * 10 + 0.01).astype(np.float32)
        if r.random() < 0.2:
            ext[r.integers(3)] = 0.0  # planar axis
        mx = (mn + ext).astype(np.float32)
        n_tris = int(r.integers(1, 2_000_000))
        f = orc.grid_frame_from_box(mn, mx, n_tris, lam)
        ext64 = mx.astype(np.float64) - mn.astype(np.float64)
        pad = np.where(ext64 == 0, 0.01, 0.01 * ext64)
        assert np.allclose(f["bmin"], mn - pad, rtol=1e-6, atol=1e-6)
        assert np.allclose(f["bmax"], mx + pad, rtol=1e-6, atol=1e-6)
        diag = f["bmax"].astype(np.float64) - f["bmin"].astype(np.float64)
        fac = (lam * n_tris / np.prod(diag)) ** (1.0 / 3.0)
        want = np.maximum(np.ceil(diag * fac), 1)
        near_int = np.abs(diag * fac - np.round(diag * fac)) < 1e-4
        assert np.all((f["n"] == want) | near_int), (f["n"], want, diag * fac)
        assert np.allclose(f["cell"] * f["n"], diag, rtol=2e-6)
        assert np.allclose(f["cell"] * f["inv_cell"], 1.0, rtol=2e-7)
        assert f["m"] == int(np.prod(f["n"].astype(np.int64)))


def test_frame_of_vertices_is_padded_bbox():
    v, _ = grt.gen_spheres(SEED_C1, 64)
    f = orc.grid_frame(v, 5.0)
    pts = v.reshape(-1, 3)
    lo, hi = pts.min(0), pts.max(0)
    ext = hi.astype(np.float64) - lo
    assert np.allclose(f["bmin"], lo - 0.01 * ext, rtol=1e-6)
    assert np.allclose(f["bmax"], hi + 0.01 * ext, rtol=1e-6)
    # M is close to lambda*N for a near-cubic box
    assert 0.9 * 5 * len(v) < f["m"] < 1.4 * 5 * len(v)


# ---------------------------------------------------------------------------
# a3-a5: grid build
# ---------------------------------------------------------------------------
def _geo_cells(frame, tri, tol):
    """Cells whose box intersects the triangle's AABB, in float64; with tol>0
    the overlap must exceed tol*cell on every axis (strict), tol<0 allows a
    gap of |tol|*cell (closed, with slack)."""
    p = tri.reshape(3, 3).astype(np.float64)
    lo, hi = p.min(0), p.max(0)
    bmin = frame["bmin"].astype(np.float64)
    cell = (frame["bmax"].astype(np.float64) - bmin) / frame["n"]
    out = []
    rngs = []
    for d in range(3):
        k = np.arange(frame["n"][d])
        c0 = bmin[d] + k * cell[d]
        c1 = c0 + cell[d]
        ov = np.minimum(c1, hi[d]) - np.maximum(c0, lo[d])
        rngs.append(k[ov > tol * cell[d]])
    for z in rngs[2]:
        for y in rngs[1]:
            for x in rngs[0]:
                out.append(int(x + frame["n"][0] * (y + frame["n"][1] * z)))
    return set(out)


@pytest.mark.parametrize("seed", range(6))
def test_grid_build_matches_geometry(seed):
    r = rng(100 + seed)
    n = int(r.integers(1, 64))
    centers = r.random((n, 3))
    v = (np.repeat(centers, 3, axis=0) + 0.3 * (r.random((3 * n, 3)) - 0.5)).reshape(n, 9)
    v = v.astype(np.float32)
    f = orc.grid_frame(v, 0.5 + seed)
    assert f["m"] <= 8 ** 3 * 4
    counts, offs, refs = orc.grid_build(v, f)
    # offsets are the exclusive scan of counts, R is the total
    assert offs[0] == 0 and np.array_equal(np.diff(offs.astype(np.int64)), counts)
    assert offs[-1] == len(refs) == counts.sum()
    cells_of = {t: set() for t in range(n)}
    for c in range(f["m"]):
        seg = refs[offs[c]:offs[c + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0), "refs ascending, no duplicates"
        for t in seg:
            cells_of[int(t)].add(c)
    for t in range(n):
        strict = _geo_cells(f, v[t], 1e-4)
        closed = _geo_cells(f, v[t], -1e-4)
        assert strict <= cells_of[t] <= closed, t
        lo, hi = orc.tri_cell_range(f, v[t])
        assert len(cells_of[t]) == int(np.prod(hi - lo + 1))


def test_grid_build_brute_force_tiny():
    """Cell membership by brute force over all (cell, triangle) pairs."""
    v = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0],
                  [0.5, 0.5, 0.5, 0.9, 0.5, 0.5, 0.5, 0.9, 0.6],
                  [0.1, 0.8, 0.2, 0.15, 0.85, 0.25, 0.12, 0.9, 0.3]], np.float32)
    f = orc.grid_frame(v, 20.0)
    counts, offs, refs = orc.grid_build(v, f)
    for c in range(f["m"]):
        seg = set(int(x) for x in refs[offs[c]:offs[c + 1]])
        want = {t for t in range(len(v)) if c in _geo_cells(f, v[t], 1e-4)}
        maybe = {t for t in range(len(v)) if c in _geo_cells(f, v[t], -1e-4)}
        assert want <= seg <= maybe


# ---------------------------------------------------------------------------
# a6: camera
# ---------------------------------------------------------------------------
def test_camera_closed_form():
    W, H = 640, 480
    cam = orc.camera_lookat((1, 2, 3), (1, 2, 10), (0, 1, 0), 60.0, W, H)
    base, nu, nv = (np.array(cam[k], np.float64) for k in ("base", "nu", "nv"))
    sy = math.tan(math.radians(30.0))
    sx = sy * W / H
    # centre of the image looks straight ahead at distance zNear
    assert np.allclose(base + 0.5 * nu + 0.5 * nv, [0, 0, 1], atol=1e-6)
    assert np.isclose(np.linalg.norm(nu), 2 * sx, rtol=1e-6)
    assert np.isclose(np.linalg.norm(nv), 2 * sy, rtol=1e-6)
    assert abs(nu @ nv) < 1e-6 and abs(nu @ [0, 0, 1]) < 1e-6
    assert nv[1] > 0          # image y grows upward
    rays = orc.primary_rays(cam)
    d = np.stack([rays["dx"], rays["dy"], rays["dz"]], 1).astype(np.float64)
    assert np.allclose(d[0], base, atol=1e-7)       # pixel (0,0) = corner = baseDir
    y, x = np.divmod(np.arange(W * H), W)
    want = base + np.outer(x / W, nu) + np.outer(y / H, nv)
    assert np.allclose(d, want, atol=1e-6)
    assert np.all(rays["tmin"] == np.float32(2e-4)) and np.all(np.isinf(rays["tmax"]))


# ---------------------------------------------------------------------------
# a9: Moller-Trumbore
# ---------------------------------------------------------------------------
TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)


def test_mt_known_answers():
    hit, t, u, v = orc.mt(TRI, (1 / 3, 1 / 3, 1.0), (0, 0, -1))
    assert hit and abs(t - 1) < 1e-6 and abs(u - 1 / 3) < 1e-6 and abs(v - 1 / 3) < 1e-6
    assert orc.mt(TRI, (0, 0, 2), (0, 0, -1))[1:] == (2.0, 0.0, 0.0)      # vertex 0
    assert orc.mt(TRI, (1, 0, 2), (0, 0, -1))[1:] == (2.0, 1.0, 0.0)      # vertex 1: u=1
    assert orc.mt(TRI, (0, 1, 2), (0, 0, -1))[1:] == (2.0, 0.0, 1.0)      # vertex 2: v=1
    assert not orc.mt(TRI, (0.2, 0.2, 0.0), (1, 0, 0))[0]                 # in-plane ray
    assert not orc.mt(TRI, (0.2, 0.2, -1.0), (0, 0, -1), 0.0)[0]          # behind origin
    assert orc.mt(TRI, (0.2, 0.2, -1.0), (0, 0, 1))[0]                    # back face hit
    assert not orc.mt(TRI, (0.6, 0.6, 1.0), (0, 0, -1))[0]                # outside u+v>1
    assert not orc.mt(TRI, (0.2, 0.2, 1.0), (0, 0, -1), 0.0, 0.5)[0]      # beyond tmax


def test_mt_barycentrics_reconstruct_hit_point():
    r = rng(5)
    for _ in range(500):
        tri = r.random(9).astype(np.float32)
        w = r.dirichlet([1, 1, 1])
        p = w[0] * tri[0:3] + w[1] * tri[3:6] + w[2] * tri[6:9]
        o = (p + r.normal(size=3)).astype(np.float32)
        d = (p - o).astype(np.float32)
        hit, t, u, v = orc.mt(tri, o, d, -1.0, 10.0)
        if min(w) < 1e-3:
            continue
        assert hit
        q = (1 - u - v) * tri[0:3] + u * tri[3:6] + v * tri[6:9]
        assert np.allclose(q, o + t * d, atol=2e-5)
        # binary32 MT loses ~|o-v0|/edge ulps; a u/v swap would be off by ~0.1
        assert abs(u - w[1]) < 2e-3 and abs(v - w[2]) < 2e-3


# ---------------------------------------------------------------------------
# a7/a8: clip + DDA
# ---------------------------------------------------------------------------
def _frame_unit(n):
    """A grid over [0,1]^3 built from a box whose padding we undo by hand."""
    f = orc.grid_frame_from_box([0, 0, 0], [1, 1, 1], n ** 3, 1.0)
    return f


def test_dda_axis_aligned_sequence():
    f = _frame_unit(8)
    nx, ny, nz = f["n"]
    cy = cz = 3
    y = f["bmin"][1] + (cy + 0.5) * f["cell"][1]
    z = f["bmin"][2] + (cz + 0.5) * f["cell"][2]
    cells = orc.dda_cells(f, (-1.0, y, z), (1.0, 0.0, 0.0))
    want = [x + nx * (cy + ny * cz) for x in range(nx)]
    assert list(cells) == want
    cells = orc.dda_cells(f, (2.0, y, z), (-1.0, -0.0, 0.0))
    assert list(cells) == want[::-1]


def _segment_cells(f, o, d, t0, t1, tol):
    """Cells whose box the FP64 segment crosses with overlap length > tol."""
    bmin = f["bmin"].astype(np.float64)
    cell = (f["bmax"].astype(np.float64) - bmin) / f["n"]
    out = set()
    for z in range(f["n"][2]):
        for y in range(f["n"][1]):
            for x in range(f["n"][0]):
                lo = bmin + cell * (x, y, z)
                hi = lo + cell
                a, b = t0, t1
                for k in range(3):
                    if d[k] == 0:
                        if not (lo[k] <= o[k] <= hi[k]):
                            a, b = 1, 0
                        continue
                    ta, tb = (lo[k] - o[k]) / d[k], (hi[k] - o[k]) / d[k]
                    a, b = max(a, min(ta, tb)), min(b, max(ta, tb))
                if b - a > tol:
                    out.add(int(x + f["n"][0] * (y + f["n"][1] * z)))
    return out


def test_dda_visits_cells_of_segment():
    f = _frame_unit(6)
    r = rng(9)
    nx, ny = f["n"][0], f["n"][1]
    for _ in range(300):
        o = r.random(3).astype(np.float32)
        d = r.normal(size=3).astype(np.float32)
        cells = orc.dda_cells(f, o, d, 0.0, np.inf)
        # consecutive cells are face neighbours
        xyz = np.stack([cells % nx, (cells // nx) % ny, cells // (nx * ny)], 1)
        assert np.all(np.abs(np.diff(xyz, axis=0)).sum(1) == 1)
        assert len(set(cells.tolist())) == len(cells)
        o64, d64 = o.astype(np.float64), d.astype(np.float64)
        tb = [((f["bmin"][k] if d64[k] < 0 else f["bmax"][k]) - o64[k]) / d64[k] for k in range(3)]
        t_exit = min(tb)
        strict = _segment_cells(f, o64, d64, 0.0, t_exit, 1e-6)
        closed = _segment_cells(f, o64, d64, 0.0, t_exit, -1e-6)
        assert strict <= set(cells.tolist()) <= closed


# ---------------------------------------------------------------------------
# closest / any hit vs FP64 brute force
# ---------------------------------------------------------------------------
def _rays(n, seed, spread=3.0):
    r = rng(seed)
    o = ((r.random((n, 3)) - 0.5) * spread + 0.5).astype(np.float32)
    target = r.random((n, 3)).astype(np.float32)
    d = (target - o).astype(np.float32)
    d[::7] = r.normal(size=(len(d[::7]), 3))
    return {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
            "dz": d[:, 2], "tmin": np.full(n, 2e-4, np.float32),
            "tmax": np.full(n, np.inf, np.float32)}


def _edge_ambiguous(b):
    w = np.stack([b["u"], b["v"], 1 - b["u"] - b["v"]], 1)
    return np.abs(w).min(1) < 1e-5


@pytest.mark.parametrize("lam", [1.0, 5.0])
def test_trace_closest_matches_brute_force(lam):
    v, _ = grt.gen_spheres(SEED_C1, 16, 0.05, 0.15)
    f = orc.grid_frame(v, lam)
    _, offs, refs = orc.grid_build(v, f)
    rays = _rays(6000, int(lam))
    o = orc.trace_closest(f, offs, refs, v, rays)
    b = orc.brute_closest_f64(v, rays)
    assert (b["id"] >= 0).sum() > 1000
    near_tie = np.abs(b["t2"] - b["t1"]) <= 1e-6 * np.abs(b["t1"])
    excused = near_tie | _edge_ambiguous(b)
    bad = (o["id"] != b["id"]) & ~excused
    assert bad.sum() == 0, np.nonzero(bad)[0][:10]
    ok = (o["id"] >= 0) & (o["id"] == b["id"])
    assert np.max(np.abs(o["t"][ok] - b["t1"][ok]) / b["t1"][ok]) < 1e-5


def test_trace_any_matches_brute_force():
    v, _ = grt.gen_spheres(SEED_C1, 16, 0.05, 0.15)
    f = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, f)
    rays = _rays(6000, 77)
    rays["tmin"][:] = 0.0
    rays["tmax"] = (rng(3).random(6000) * 2).astype(np.float32)
    o = orc.trace_any(f, offs, refs, v, rays)
    b = orc.brute_any_f64(v, rays)
    tn = b["t_near"]
    boundary = np.isfinite(tn) & (np.abs(tn - rays["tmax"]) <= 1e-5 * np.abs(tn))
    bad = (o["occluded"] != b["occluded"]) & ~boundary
    assert bad.sum() == 0
    assert 100 < o["occluded"].sum() < 5900


def test_closest_needs_cell_exit_bound():
    """Two triangles: the far one is listed in the first cell the ray visits,
    the near one only in a later cell.  The cell-exit bound makes the traversal
    return the near one (reference intersectTriangles, ray.tfar = cell exit)."""
    far = [0.05, 0.0, 0.9, 0.05, 1.0, 0.9, 0.05, 0.0, -0.9]   # spans x=0.05 plane, long in z
    near = [0.6, 0.4, 0.45, 0.6, 0.6, 0.45, 0.6, 0.5, 0.55]
    v = np.array([far, near, [0, 0, 0, 1, 1, 1, 1, 0, 1]], np.float32)
    f = orc.grid_frame(v, 4.0)
    _, offs, refs = orc.grid_build(v, f)
    rays = {"ox": np.float32([0.9]), "oy": np.float32([0.5]), "oz": np.float32([0.5]),
            "dx": np.float32([-1.0]), "dy": np.float32([0.0]), "dz": np.float32([0.0]),
            "tmin": np.float32([0.0]), "tmax": np.float32([np.inf])}
    o = orc.trace_closest(f, offs, refs, v, rays)
    b = orc.brute_closest_f64(v, rays)
    assert o["id"][0] == b["id"][0] == 1


# ---------------------------------------------------------------------------
# a10-a12: secondary rays
# ---------------------------------------------------------------------------
def _one_hit_scene():
    v = np.array([[-1, 0, -1, 1, 0, -1, 0, 0, 1]], np.float32)      # y=0 plane
    n = np.array([[0, 1, 0, 0, 1, 0, 0, 1, 0]], np.float32)
    rays = {"ox": np.float32([0.1, 0.2]), "oy": np.float32([1.0, 1.0]),
            "oz": np.float32([0.0, 0.1]), "dx": np.float32([0.3, 0.0]),
            "dy": np.float32([-1.0, 1.0]), "dz": np.float32([0.2, 0.0]),
            "tmin": np.float32([0, 0]), "tmax": np.float32([np.inf, np.inf])}
    f = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, f)
    hits = orc.trace_closest(f, offs, refs, v, rays)
    return v, n, rays, hits


def test_reflection_closed_form():
    v, n, rays, hits = _one_hit_scene()
    assert hits["id"][0] == 0 and hits["id"][1] == -1
    r = orc.gen_reflection(rays, hits, n, 1e-3)
    d = np.array([0.3, -1.0, 0.2])
    rd = np.array([r["dx"][0], r["dy"][0], r["dz"][0]])
    assert np.allclose(rd, [0.3, 1.0, 0.2], atol=1e-6)        # d - 2(d.n)n
    assert np.isclose(np.linalg.norm(rd), np.linalg.norm(d), rtol=1e-6)
    h = np.array([0.1, 0.0, 0.0]) + d * hits["t"][0] + np.array([0, 0, 0])
    assert np.allclose([r["ox"][0], r["oy"][0], r["oz"][0]], [0.4, 1e-3, 0.2], atol=1e-6)
    assert r["tmin"][1] > r["tmax"][1]                          # miss -> inactive


def test_shadow_closed_form():
    v, n, rays, hits = _one_hit_scene()
    light = (3.0, 5.0, -2.0)
    s = orc.gen_shadow(rays, hits, n, light, 1e-3)
    hp = np.array([0.4, 0.0, 0.2])
    L = np.array(light) - hp
    sd = np.array([s["dx"][0], s["dy"][0], s["dz"][0]])
    assert np.allclose(sd, L / np.linalg.norm(L), atol=1e-6)
    assert np.isclose(s["tmax"][0], np.linalg.norm(L), rtol=1e-6) and s["tmin"][0] == 0
    # light below the surface: no shadow ray
    s2 = orc.gen_shadow(rays, hits, n, (0.0, -5.0, 0.0), 1e-3)
    assert s2["tmin"][0] > s2["tmax"][0]


def test_shading_normal_interpolates():
    n9 = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    assert np.allclose(orc.shading_normal(n9, 0, 0), [1, 0, 0])
    assert np.allclose(orc.shading_normal(n9, 1, 0), [0, 1, 0])
    assert np.allclose(orc.shading_normal(n9, 0, 1), [0, 0, 1])
    assert np.allclose(orc.shading_normal(n9, 1 / 3, 1 / 3), np.ones(3) / math.sqrt(3), atol=1e-6)


def test_sincos_turns():
    t = np.linspace(0, 1, 4001, endpoint=False).astype(np.float32)
    sc = np.array([orc.sincos_turns(float(x)) for x in t])
    assert np.max(np.abs(sc[:, 0] - np.sin(2 * np.pi * t.astype(np.float64)))) < 2e-6
    assert np.max(np.abs(sc[:, 1] - np.cos(2 * np.pi * t.astype(np.float64)))) < 2e-6


def test_ao_cosine_weighted():
    v, n, rays, hits = _one_hit_scene()
    spp = 20000
    a = orc.gen_ao({k: x[:1] for k, x in rays.items()}, {k: x[:1] for k, x in hits.items()}, n,
                   spp, 42, 1e-3, 0.05)
    d = np.stack([a["dx"], a["dy"], a["dz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-6)
    cos = d[:, 1]
    assert cos.min() >= 0.0
    assert abs(cos.mean() - 2 / 3) < 0.01          # E[cos] of a cosine lobe
    assert abs((cos ** 2).mean() - 0.5) < 0.01     # E[cos^2] = 1/2
    assert abs(d[:, 0].mean()) < 0.02 and abs(d[:, 2].mean()) < 0.02
    assert np.all(a["tmax"] == np.float32(0.05))


def test_uniform24_is_uniform():
    u = np.array([orc.uniform24(7, i, 0) for i in range(20000)])
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert np.all(u * 2 ** 24 == np.floor(u * 2 ** 24))


def test_transform_is_rigid():
    v, nrm, centers = grt.gen_stadium(SEED_C2, 4, 64, 4 * 64 - 2)
    xf = grt.gen_cluster_transforms(SEED_C2, 3, centers, 0.05)
    tv, tn = orc.transform(v, nrm, xf, 64, 4 * 64 - 2)
    m = xf.reshape(-1, 3, 4).astype(np.float64)
    for t in (0, 70, 200, 253):
        c = t // 64
        p = v[t].reshape(3, 3).astype(np.float64)
        want = p @ m[c, :, :3].T + m[c, :, 3]
        assert np.allclose(tv[t].reshape(3, 3), want, atol=1e-6)
        assert np.allclose(np.linalg.norm(p[1] - p[0]),
                           np.linalg.norm(tv[t].reshape(3, 3)[1] - tv[t].reshape(3, 3)[0]),
                           rtol=1e-4)
    assert np.array_equal(tv[-2:], v[-2:])          # ground plane untouched


# ---------------------------------------------------------------------------
# Phong shading (shading_FRAG.glsl:116-173): known answers on hand-built hits
# ---------------------------------------------------------------------------
def test_pow_u32_is_the_integer_power():
    assert orc.pow_u32(2.0, 10) == 1024.0
    assert orc.pow_u32(0.75, 1) == np.float32(0.75)
    assert orc.pow_u32(0.3, 0) == 1.0
    r = rng(11)
    for _ in range(2000):
        x = float(np.float32(r.uniform(0.5, 1.0)))   # x^e stays a normal float
        e = int(r.integers(1, 100))
        want = x ** e
        # each of the <= 2 log2(e) roundings adds at most 2^-24 relative error,
        # squared terms double theirs: 2 e ulps bound the whole ladder
        assert abs(orc.pow_u32(x, e) - want) <= 2.0 * e * 2.0 ** -24 * want, (x, e)


def _flat_scene():
    """One triangle in the plane y = 0, vertex normals (0, 1, 0)."""
    nrm = np.tile(np.float32([0, 1, 0]), 3).reshape(1, 9)
    return nrm


def _shade_one(o, d, light, kd=(0.2, 0.4, 0.6), occluded=0, spec_exp=50, t=1.0, depth=0,
               rgb=None, hit=True):
    rays = {k: np.float32([x]) for k, x in zip(("ox", "oy", "oz", "dx", "dy", "dz"), (*o, *d))}
    hits = {"id": np.int32([0 if hit else -1]), "t": np.float32([t]),
            "u": np.float32([0.25]), "v": np.float32([0.25])}
    params = {"light": light, "spec_exp": spec_exp}
    rgb, img = orc.shade_phong(rays, hits, _flat_scene(), np.uint8([occluded]), params, depth,
                               rgb, np.float32([kd]))
    return rgb[0], img[0]


F = np.float32


def test_phong_head_on_light_gives_diffuse_plus_full_specular():
    # view straight down, light straight up: N.L = 1 and the mirror vector is L
    c, img = _shade_one((0.25, 1, 0.25), (0, -1, 0), (0.25, 11, 0.25))
    assert list(c) == [F(0.2) + F(0.5), F(0.4) + F(0.5), F(0.6) + F(0.5)]
    # RGBA8 truncates c * 255 and clamps: 0.7 -> 178, 0.9 -> 229, 1.1 -> 255
    assert img == 178 | (229 << 8) | (255 << 16) | (255 << 24)
    # without the specular term only the diffuse colour is left
    c, _ = _shade_one((0.25, 1, 0.25), (0, -1, 0), (0.25, 11, 0.25), spec_exp=0)
    assert list(c) == [F(0.2), F(0.4), F(0.6)]


def test_phong_shadow_dims_diffuse_and_drops_specular():
    c, _ = _shade_one((0.25, 1, 0.25), (0, -1, 0), (0.25, 11, 0.25), occluded=1)
    assert list(c) == [F(0.2) * F(0.2), F(0.4) * F(0.2), F(0.6) * F(0.2)]
    # a light below the surface shadows the point: N.L < 0, diffuse clamps to 0
    c, _ = _shade_one((0.25, 1, 0.25), (0, -1, 0), (0.25, -10, 0.25))
    assert list(c) == [0.0, 0.0, 0.0]


def test_phong_specular_follows_the_mirror_direction():
    # 45-degree view from -x: the mirror direction is (1, 1, 0)/sqrt 2; a light
    # along it gives the full specular, N.L = 1/sqrt 2 scales the diffuse
    c, _ = _shade_one((-0.75, 1, 0.25), (1, -1, 0), (10.25, 10, 0.25))
    s = 1.0 / math.sqrt(2.0)
    for k, kd in enumerate((0.2, 0.4, 0.6)):
        assert abs(c[k] - (kd * s + 0.5)) < 1e-6
    # the light on the other side (along the incoming ray): no specular
    c, _ = _shade_one((-0.75, 1, 0.25), (1, -1, 0), (-9.75, 10, 0.25))
    for k, kd in enumerate((0.2, 0.4, 0.6)):
        assert abs(c[k] - kd * s) < 1e-6
    # a light 45 degrees off the mirror direction: (cos 45)^50 = 2^-25
    c, _ = _shade_one((0.25, 1, 0.25), (0, -1, 0), (10.25, 10, 0.25), kd=(0, 0, 0))
    assert all(abs(x / (0.5 * 2.0 ** -25) - 1.0) < 1e-4 for x in c)


def test_phong_normal_faces_the_ray():
    # seen from below, the surface normal is turned to (0, -1, 0)
    c, _ = _shade_one((0.25, -1, 0.25), (0, 1, 0), (0.25, -11, 0.25))
    assert list(c) == [F(0.2) + F(0.5), F(0.4) + F(0.5), F(0.6) + F(0.5)]


def test_phong_misses_and_reflection_accumulation():
    c, img = _shade_one((0, 0, 0), (0, 0, 1), (0, 10, 0), hit=False)
    assert list(c) == [1.0, 1.0, 1.0] and img == 0xffffffff
    base = np.float32([[0.1, 0.2, 0.3]])
    c, _ = _shade_one((0, 0, 0), (0, 0, 1), (0, 10, 0), hit=False, depth=1, rgb=base)
    assert list(c) == list(base[0])
    col, _ = _shade_one((0.25, 1, 0.25), (0, -1, 0), (0.25, 11, 0.25))
    for depth, w in ((1, F(0.4)), (2, F(0.4) * F(0.4))):
        c, _ = _shade_one((0.25, 1, 0.25), (0, -1, 0), (0.25, 11, 0.25), depth=depth, rgb=base)
        assert list(c) == [base[0][k] + w * col[k] for k in range(3)]


def test_headlight_known_answers():
    """RayTracing_kernel.cu:104-120: colour = (ambient + (diffuse - ambient) *
    nDotD) * diffuse with ambient 0.1, diffuse (0, 0, 1), nDotD = -N.d/|d|,
    packed r | g << 8 | b << 16 | 255 << 24 from c * 255 truncated."""
    facing = np.tile(np.float32([0, 0, -1]), 3)
    # head-on: nDotD = 1 -> blue 1.0
    assert orc.headlight((0, 0, 2.0), facing, 0, 0.3, 0.3) == 0xFFFF0000
    # 60 degrees off the normal: nDotD = 0.5 -> blue 0.55 -> 140
    d = (0.0, math.sqrt(3.0) / 2.0, 0.5)
    assert orc.headlight(d, facing, 0, 0.3, 0.3) == 0xFF000000 | (140 << 16)
    # seen from behind: nDotD = -1 -> blue clamps to 0, red and green stay 0
    assert orc.headlight((0, 0, -1.0), facing, 0, 0.3, 0.3) == 0xFF000000
    # a miss is white
    assert orc.headlight((0, 0, 1.0), facing, -1, 0.0, 0.0) == 0xFFFFFFFF


# ---------------------------------------------------------------------------
# The reference CPU tracer's Wald test and the specification's Moller-Trumbore
# ---------------------------------------------------------------------------
def _tri_accel(v0, v1, v2):
    """TriAccel::buildFrom (Triangle.cpp:11-55) in binary32."""
    f = np.float32
    b = (v2 - v0).astype(f)
    c = (v1 - v0).astype(f)
    n = np.array([f(c[1] * b[2]) - f(c[2] * b[1]), f(c[2] * b[0]) - f(c[0] * b[2]),
                  f(c[0] * b[1]) - f(c[1] * b[0])], f)
    if abs(n[0]) > abs(n[1]) and abs(n[0]) > abs(n[2]):
        k = 0
    elif abs(n[1]) > abs(n[2]):
        k = 1
    else:
        k = 2
    u, w = (k + 1) % 3, (k + 2) % 3
    acc = {"k": k, "n_u": f(n[u] / n[k]), "n_v": f(n[w] / n[k])}
    nl = (n * f(f(1.0) / n[k])).astype(f)
    acc["n_d"] = f(f(f(v0[0] * nl[0]) + f(v0[1] * nl[1])) + f(v0[2] * nl[2]))
    inv = f(f(1.0) / f(f(b[u] * c[w]) - f(b[w] * c[u])))
    acc["b_nu"], acc["b_nv"] = f(-b[w] * inv), f(b[u] * inv)
    acc["b_d"] = f(f(f(b[w] * v0[u]) - f(b[u] * v0[w])) * inv)
    acc["c_nu"], acc["c_nv"] = f(c[w] * inv), f(-c[u] * inv)
    acc["c_d"] = f(f(f(c[u] * v0[w]) - f(c[w] * v0[u])) * inv)
    return acc


def _hit_wald(acc, o, d):
    """RayTriIntersection::hitWald (RayTriIntersection.cpp:12-48) without the
    interval test: (f, lambda, mu), or None when outside the triangle."""
    f = np.float32
    k = acc["k"]
    u, w = (k + 1) % 3, (k + 2) % 3
    nd = f(f(1.0) / f(f(d[k] + f(acc["n_u"] * d[u])) + f(acc["n_v"] * d[w])))
    dist = f(nd * f(f(f(acc["n_d"] - o[k]) - f(acc["n_u"] * o[u])) - f(acc["n_v"] * o[w])))
    hu = f(o[u] + f(dist * d[u]))
    hv = f(o[w] + f(dist * d[w]))
    lam = f(f(f(hu * acc["b_nu"]) + f(hv * acc["b_nv"])) + acc["b_d"])
    mue = f(f(f(hu * acc["c_nu"]) + f(hv * acc["c_nv"])) + acc["c_d"])
    if lam < 0 or mue < 0 or f(f(1.0) - lam) - mue < 0:
        return None
    return dist, lam, mue


def test_wald_barycentrics_are_the_mt_weights():
    """The reference CPU grid path records hitWald's lambda, mu as the hit's
    v1Coord, v2Coord (RayTriIntersection.cpp:44-47, called from
    UniformGridAccStruct.cpp:373); the specification's Moller-Trumbore returns
    u, v as the weights of vertices 1 and 2.  On random well-shaped triangles
    hit at random interior points, at least 17 degrees off grazing (|cos| >=
    0.3, where neither binary32 test loses more than a few ulps to
    conditioning), both agree with the float64 solution of o + t d = v0 +
    u e1 + v e2 to 2e-5 and with each other to 3e-5 (99% within 5e-6); t
    agrees to 1e-5 relative."""
    r = rng(77)
    dl, dm = [], []
    for _ in range(3000):
        v = r.uniform(-1, 1, (3, 3)).astype(np.float32)
        e1, e2 = v[1] - v[0], v[2] - v[0]
        cr = np.cross(e1.astype(np.float64), e2.astype(np.float64))
        if np.linalg.norm(cr) < 0.05 or min(np.linalg.norm(e1), np.linalg.norm(e2)) < 0.2:
            continue
        a, b = r.uniform(0.05, 0.9, 2)
        if a + b > 0.95:
            a, b = 0.95 - b, 0.95 - a
            if a < 0.05 or b < 0.05:
                continue
        p = v[0].astype(np.float64) + a * e1 + b * e2
        o = (p + r.normal(size=3) * 2.0).astype(np.float32)
        d = (p - o).astype(np.float32)
        cosang = abs(np.dot(cr / np.linalg.norm(cr), d / np.linalg.norm(d)))
        if cosang < 0.3:
            continue
        ok, t, u, vv = orc.mt(v.reshape(9), o, d)
        w = _hit_wald(_tri_accel(v[0], v[1], v[2]), o, d)
        assert ok and w is not None, (v, o, d)
        f, lam, mue = w
        V, O, D = v.astype(np.float64), o.astype(np.float64), d.astype(np.float64)
        t64, u64, v64 = np.linalg.solve(np.stack([-D, V[1] - V[0], V[2] - V[0]], 1), O - V[0])
        assert abs(f - t) <= 1e-5 * abs(t), (f, t)
        assert max(abs(u - u64), abs(vv - v64)) <= 2e-5, (u, vv, u64, v64)
        assert max(abs(lam - u64), abs(mue - v64)) <= 2e-5, (lam, mue, u64, v64)
        assert abs(lam - u) <= 3e-5 and abs(mue - vv) <= 3e-5, (lam, u, mue, vv)
        dl.append(abs(lam - u))
        dm.append(abs(mue - vv))
    assert len(dl) > 1000
    assert np.percentile(dl, 99) <= 5e-6 and np.percentile(dm, 99) <= 5e-6


def test_wald_weights_swapped_would_fail():
    """The pin above is not vacuous: on an asymmetric hit, lambda and mu
    differ, so a swapped (v1Coord, v2Coord) assignment would be caught."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    o = np.array([0.2, 0.6, -1.0], np.float32)
    d = np.array([0, 0, 1], np.float32)
    ok, t, u, vv = orc.mt(v.reshape(9), o, d)
    assert ok
    f, lam, mue = _hit_wald(_tri_accel(v[0], v[1], v[2]), o, d)
    assert (u, vv) == (np.float32(0.2), np.float32(0.6))
    assert (lam, mue) == (u, vv) and lam != mue and f == t == 1.0
