"""The kd-tree comparison accelerator on the CPU: the host-built tree's
structure and completeness, the surface-area-heuristic choice at the root,
and the oracle's kd traversal against float64 brute force and against the
grid traversal (the same closest triangle, t, u and v).

Reference: TriangleTreeBuilder.cpp:13-358 (build), KdTreeAccStruct.cpp:122-240
(traversal), KdTreeAccStructBuilder.cpp:39-97 (node layout).
"""
import numpy as np
import pytest

import grt
import pyoracle as orc
from conftest import SEED_C1, SEED_C2, rng

MAX_DEPTH = 48
CT, CI = np.float32(1.0), np.float32(1.4)


def _tree(v):
    kd = grt.KdTree.from_host(v)
    info = kd.info()
    nodes, elements = kd.download()
    kd.close()
    return info, nodes, elements


def _scenes():
    r = rng(3)
    out = {}
    out["spheres"] = grt.gen_spheres(SEED_C1, 16, 0.03, 0.1)[0]
    planar = r.random((300, 9), dtype=np.float32)
    planar[:, 1::3] = 0.5
    out["planar"] = np.vstack([planar, r.random((100, 9), dtype=np.float32)])
    dup = np.repeat(r.random((60, 9), dtype=np.float32), 3, axis=0)
    out["duplicates"] = dup
    degen = r.random((300, 9), dtype=np.float32)
    degen[::5, 3:6] = degen[::5, 0:3]
    degen[::5, 6:9] = degen[::5, 0:3]
    out["degenerate"] = degen
    out["stadium"] = grt.gen_stadium(SEED_C2, 4, 512, 4 * 512 - 2)[0]
    # axis-aligned boxes of triangles meeting on shared planes
    q = []
    for x in range(4):
        for y in range(3):
            x0, y0 = np.float32(x * 0.25), np.float32(y * 0.25)
            q.append([x0, y0, 0, x0 + 0.25, y0, 0, x0, y0 + 0.25, 0])
            q.append([x0, y0, 0.5, x0 + 0.25, y0, 0.5, x0, y0 + 0.25, 0.5])
            q.append([x0, y0, 0, x0, y0 + 0.25, 0, x0, y0, 0.5])
    out["grid_planes"] = np.array(q, np.float32)
    return out


SCENES = list(_scenes())


def _walk(nodes):
    """(leaf node index, depth, box-free path) for every leaf, and checks the
    breadth-first layout: children adjacent, every node reached once."""
    seen = np.zeros(len(nodes), bool)
    stack = [(0, 0)]
    leaves = []
    while stack:
        k, d = stack.pop()
        assert not seen[k], f"node {k} reached twice"
        seen[k] = True
        w = int(nodes[k, 1])
        if w & 0x80000000:
            leaves.append((k, d))
        else:
            left = w >> 2
            assert left > k and left + 1 < len(nodes), (k, left)
            stack += [(left, d + 1), (left + 1, d + 1)]
    assert seen.all(), "unreachable nodes"
    return leaves


@pytest.mark.parametrize("name", SCENES)
def test_kd_structure(name):
    v = _scenes()[name]
    info, nodes, elements = _tree(v)
    assert info["nodes"] == len(nodes) and info["elements"] == len(elements)
    leaves = _walk(nodes)
    assert info["leaves"] == len(leaves)
    assert info["depth"] == max(d for _, d in leaves) <= MAX_DEPTH
    assert info["empty_leaves"] == sum(1 for k, _ in leaves if (nodes[k, 1] & 0x7fffffff) == 0)
    # leaf element ranges tile [0, elements) in breadth-first order, each
    # ascending in triangle id
    ranges = sorted((int(nodes[k, 0]), int(nodes[k, 1] & 0x7fffffff)) for k, _ in leaves)
    pos = 0
    for first, cnt in ranges:
        assert first == pos
        seg = elements[first:first + cnt]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        pos += cnt
    assert pos == len(elements)
    assert elements.max() < len(v)
    np.testing.assert_array_equal(info["bmin"], v.reshape(-1, 3).min(axis=0))
    np.testing.assert_array_equal(info["bmax"], v.reshape(-1, 3).max(axis=0))


def _leaves_of_point(nodes, p):
    """Leaves whose region holds point p (both sides of a plane it lies on)."""
    out, stack = [], [0]
    while stack:
        k = stack.pop()
        w = int(nodes[k, 1])
        if w & 0x80000000:
            out.append(k)
            continue
        axis, left = w & 3, w >> 2
        split = float(nodes[k, 0:1].view(np.float32)[0])
        if p[axis] <= split:
            stack.append(left)
        if p[axis] >= split:
            stack.append(left + 1)
    return out


@pytest.mark.parametrize("name", SCENES)
def test_kd_every_triangle_point_is_in_its_leaf(name):
    """Completeness: random interior points of every non-degenerate triangle
    (float64) lie in leaves that list the triangle."""
    v = _scenes()[name].astype(np.float64).reshape(-1, 3, 3)
    info, nodes, elements = _tree(v.astype(np.float32).reshape(-1, 9))
    members = {}
    for k in range(len(nodes)):
        w = int(nodes[k, 1])
        if w & 0x80000000:
            first, cnt = int(nodes[k, 0]), w & 0x7fffffff
            members[k] = set(elements[first:first + cnt].tolist())
    r = rng(9)
    area = np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    checked = 0
    for t in np.nonzero(area > 1e-12)[0][:400]:
        cen = v[t].mean(axis=0)
        for _ in range(6):
            a, b = r.random(2)
            if a + b > 1.0:
                a, b = 1.0 - a, 1.0 - b
            q = v[t, 0] + a * (v[t, 1] - v[t, 0]) + b * (v[t, 2] - v[t, 0])
            p = cen + 0.96 * (q - cen)   # strictly inside the triangle
            ls = _leaves_of_point(nodes, p)
            assert any(t in members[k] for k in ls), f"triangle {t} missing at {p}"
            checked += 1
    assert checked > 100


def _sah_cost(lo, hi, k, p, nl, np_, nr):
    """The builder's plane cost in binary32 (the reference's sah())."""
    f = np.float32

    def area(a, b):
        x, y, z = f(b[0] - a[0]), f(b[1] - a[1]), f(b[2] - a[2])
        return f(f(f(x * y) + f(x * z)) + f(y * z))
    llo, lhi, rlo, rhi = lo.copy(), hi.copy(), lo.copy(), hi.copy()
    lhi[k] = min(p, hi[k])
    rlo[k] = max(p, lo[k])
    inv_area = f(f(1.0) / area(lo, hi))
    pl, pr = f(area(llo, lhi) * inv_area), f(area(rlo, rhi) * inv_area)
    inv_ext = f(f(1.0) / f(hi[k] - lo[k]))
    lext, rext = f(f(lhi[k] - llo[k]) * inv_ext), f(f(rhi[k] - rlo[k]) * inv_ext)

    def one(a, b):
        c = f(CT + f(CI * f(f(pl * f(a)) + f(pr * f(b)))))
        if a == 0:
            c = f(c * f(f(0.8) + f(f(0.2) * rext)))
        elif b == 0:
            c = f(c * f(f(0.8) + f(f(0.2) * lext)))
        return c
    return min(one(nl + np_, nr), one(nl, np_ + nr))


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_kd_root_split_minimises_sah_by_brute_force(seed):
    """On small scenes the root plane has the least cost over every candidate
    position, counted directly (not by the event sweep): triangles left of,
    in and right of the plane."""
    r = rng(100 + seed)
    c = r.random((40, 1, 3), dtype=np.float32)
    v = (c + (r.random((40, 3, 3), dtype=np.float32) - 0.5) * 0.2).astype(np.float32)
    if seed == 4:
        v[::4, :, 1] = np.float32(0.5)   # planar triangles on one plane
    v = v.reshape(-1, 9)
    info, nodes, _ = _tree(v)
    w = int(nodes[0, 1])
    assert not (w & 0x80000000), "root should split"
    lo, hi = info["bmin"].astype(np.float32), info["bmax"].astype(np.float32)
    tb = v.reshape(-1, 3, 3)
    tlo, thi = tb.min(axis=1), tb.max(axis=1)
    best = np.float32(np.inf)
    for k in range(3):
        planar = tlo[:, k] == thi[:, k]
        for p in np.unique(np.concatenate([tlo[:, k], thi[:, k]])):
            nl = int(np.sum(~planar & (tlo[:, k] < p)) + np.sum(planar & (tlo[:, k] < p)))
            npl = int(np.sum(planar & (tlo[:, k] == p)))
            nr = int(np.sum(~planar & (thi[:, k] > p)) + np.sum(planar & (tlo[:, k] > p)))
            best = min(best, _sah_cost(lo, hi, k, np.float32(p), nl, npl, nr))
    axis = w & 3
    p = np.float32(nodes[0, 0:1].view(np.float32)[0])
    planar = tlo[:, axis] == thi[:, axis]
    nl = int(np.sum(~planar & (tlo[:, axis] < p)) + np.sum(planar & (tlo[:, axis] < p)))
    npl = int(np.sum(planar & (tlo[:, axis] == p)))
    nr = int(np.sum(~planar & (thi[:, axis] > p)) + np.sum(planar & (tlo[:, axis] > p)))
    chosen = _sah_cost(lo, hi, axis, p, nl, npl, nr)
    assert chosen == best, (chosen, best)
    assert chosen < CI * np.float32(len(v))


def _random_rays(n, seed, lo=-0.5, hi=1.5):
    r = rng(seed)
    o = r.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = r.normal(size=(n, 3)).astype(np.float32)
    d[::9, 1:] = 0.0
    d[1::11, 0] = -0.0
    return {"ox": o[:, 0], "oy": o[:, 1], "oz": o[:, 2], "dx": d[:, 0], "dy": d[:, 1],
            "dz": d[:, 2], "tmin": np.full(n, 1e-4, np.float32),
            "tmax": np.where(np.arange(n) % 4 == 0, np.float32(0.5), np.float32(np.inf))}


@pytest.mark.parametrize("name", SCENES)
def test_kd_oracle_equals_grid_and_brute_force(name):
    """The oracle's kd traversal returns the grid traversal's closest hit
    (triangle, t, u, v bit for bit) and any-hit flag, and the float64 brute
    force's triangle, on random rays; rays whose two nearest float64 hits lie
    within 1e-6 relative are reported, not compared."""
    v = _scenes()[name]
    info, nodes, elements = _tree(v)
    box = (info["bmin"], info["bmax"])
    rays = _random_rays(6000, 17)
    c = v.reshape(-1, 3).mean(axis=0)
    for k, ax in enumerate(("dx", "dy", "dz")):   # half of them aimed at the scene
        rays[ax][::2] = (c[k] - rays["o" + ax[1]][::2]).astype(np.float32)
    kd = orc.kd_trace_closest(box, nodes, elements, v, rays)
    fr = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, fr)
    gr = orc.trace_closest(fr, offs, refs, v, rays)
    bf = orc.brute_closest_f64(v, rays)
    tie = np.abs(bf["t2"] - bf["t1"]) <= 1e-6 * np.abs(bf["t1"])
    assert (kd["id"] >= 0).sum() > 500
    for f in ("id", "t", "u", "v"):
        a, b = kd[f][~tie], gr[f][~tie]
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b), f
    assert np.array_equal(kd["id"][~tie], bf["id"][~tie])
    ka = orc.kd_trace_any(box, nodes, elements, v, rays)["occluded"]
    ga = orc.trace_any(fr, offs, refs, v, rays)["occluded"]
    assert np.array_equal(ka, ga)


def test_kd_oracle_visits_fewer_leaves_than_grid_cells():
    """On the config-1 spheres the kd-tree visits fewer leaves and tests fewer
    triangles per primary ray than the lambda-5 grid visits cells and tests
    (the comparison the accelerator exists for)."""
    v, _ = grt.gen_spheres(SEED_C1, 64, 0.02, 0.08)
    info, nodes, elements = _tree(v)
    cam = grt.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0, 1, 0), 60.0, 128, 128)
    rays = orc.primary_rays(grt.camera_dict(cam))
    kd = orc.kd_trace_closest((info["bmin"], info["bmax"]), nodes, elements, v, rays,
                              counters=True)
    fr = orc.grid_frame(v, 5.0)
    _, offs, refs = orc.grid_build(v, fr)
    gr = orc.trace_closest(fr, offs, refs, v, rays, counters=True)
    assert np.array_equal(kd["id"], gr["id"])
    assert kd["leaves"].mean() < gr["cells"].mean()
    assert kd["tests"].mean() < gr["tests"].mean()


def test_kd_build_rejects_bad_input():
    with pytest.raises(grt.GrtError):
        grt.KdTree.from_host(np.array([[0, 0, 0, 1, 0, 0, np.nan, 1, 0]], np.float32))
    with pytest.raises(grt.GrtError):
        grt.KdTree.from_host(np.zeros((0, 9), np.float32))
