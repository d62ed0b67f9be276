"""Scene ingestion (grt_mesh_*, host only): primitive assembly, the degenerate
filter with id compaction, transforms, OBJ/MTL loading and the procedural
shapes.  Expected values are worked out by hand from the reference's rules
(PrimitiveBuilder.cpp:96-219, Triangle.cpp:11-63, GpuObjLoader.cpp:138-151)."""
import math

import numpy as np
import pytest

import grt

P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 2, 0], [1, 2, 0], [0, 3, 0]],
             np.float32)


def _tris(m):
    v, n, kd, src = m.arrays()
    return v.reshape(-1, 3, 3), n.reshape(-1, 3, 3), kd, src


def _as_indices(tris):
    """Each triangle's corners as indices into P."""
    out = []
    for t in tris:
        out.append(tuple(int(np.nonzero((P == c).all(axis=1))[0][0]) for c in t))
    return out


def test_primitive_assembly_rules():
    m = grt.Mesh()
    m.primitive(grt.TRIANGLE_STRIP, P[:5])
    t, _, _, _ = _tris(m)
    assert _as_indices(t) == [(0, 1, 2), (3, 2, 1), (2, 3, 4)]
    m = grt.Mesh()
    m.primitive(grt.TRIANGLE_FAN, P[:5])
    m.primitive(grt.POLYGON, P[:4])
    t, _, _, _ = _tris(m)
    assert _as_indices(t) == [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 1, 2), (0, 2, 3)]
    m = grt.Mesh()
    m.primitive(grt.TRIANGLES, P)   # 7 vertices: two triangles, the last vertex unused
    t, _, _, _ = _tris(m)
    assert _as_indices(t) == [(0, 1, 2), (3, 4, 5)]
    assert m.info() == (2, 0)


def test_degenerate_triangles_are_dropped_and_ids_compacted():
    m = grt.Mesh()
    nan = float("nan")
    m.primitive(grt.TRIANGLES, [[0, 0, 0], [1, 0, 0], [0, 1, 0],       # 0 kept
                                [0, 0, 0], [1, 1, 1], [2, 2, 2],       # 1 collinear
                                [5, 5, 5], [5, 5, 5], [6, 5, 5],       # 2 repeated corner
                                [0, 0, 0], [nan, 0, 0], [0, 1, 0],     # 3 not finite
                                [0, 0, 1e-3], [1e-3, 0, 1e-3], [0, 1e-3, 1e-3],   # 4 kept (small)
                                [1e30, 0, 0], [0, 1e30, 0], [0, 0, 1e30]])        # 5 overflows
    t, _, _, src = _tris(m)
    assert m.info() == (2, 4)
    assert list(src) == [0, 4]
    assert t[1][0][2] == np.float32(1e-3)
    # a strip through a repeated vertex loses exactly the triangles using it twice
    m = grt.Mesh()
    m.primitive(grt.TRIANGLE_STRIP, [P[0], P[1], P[1], P[2], P[3]])
    assert m.info() == (1, 2)
    assert list(m.arrays()[3]) == [2]


def test_filter_matches_the_area_on_random_triangles():
    """The accelerator constants are finite exactly when the triangle has a
    non-zero normal, for coordinates of moderate size."""
    r = np.random.default_rng(3)
    tris = r.uniform(-1, 1, (400, 3, 3)).astype(np.float32)
    tris[::4, 2] = tris[::4, 0] + 0.5 * (tris[::4, 1] - tris[::4, 0])   # on an edge's line?
    tris[1::4, 2] = tris[1::4, 1]                                     # repeated corner
    m = grt.Mesh()
    m.primitive(grt.TRIANGLES, tris.reshape(-1, 3))
    _, _, _, src = _tris(m)
    kept = np.zeros(len(tris), bool)
    kept[src] = True
    e1 = tris[:, 1].astype(np.float32) - tris[:, 0]
    e2 = tris[:, 2].astype(np.float32) - tris[:, 0]
    c = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1],
                  e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], 1)
    nonzero = np.abs(c).max(axis=1) > 0
    assert np.array_equal(kept, nonzero)
    assert 0 < kept.sum() < len(tris)


def test_transform_of_vertices_and_normals():
    m = grt.Mesh()
    # scale x by 2 and translate by (1, 2, 3)
    m.set_transform([2, 0, 0, 1, 0, 1, 0, 2, 0, 0, 1, 3])
    s = 1.0 / math.sqrt(2.0)
    m.normal(s, s, 0)
    m.primitive(grt.TRIANGLES, P[:3])
    t, n, _, _ = _tris(m)
    assert np.array_equal(t[0], np.float32([[1, 2, 3], [3, 2, 3], [1, 3, 3]]))
    # normals take the inverse transpose diag(1/2, 1, 1): (1/2, 1, 0) normalised
    want = np.array([0.5, 1.0, 0.0]) / math.sqrt(1.25)
    assert np.allclose(n[0], want, atol=1e-6)
    with pytest.raises(grt.GrtError):
        m.set_transform([0] * 12)   # singular


def test_colour_follows_the_triangles():
    m = grt.Mesh()
    m.primitive(grt.TRIANGLES, P[:3])
    m.set_color(0.25, 0.5, 0.75)
    m.primitive(grt.TRIANGLES, P[3:6])
    _, _, kd, _ = _tris(m)
    assert kd.tolist() == [[1, 1, 1], [0.25, 0.5, 0.75]]


def test_immediate_mode_misuse_is_refused():
    m = grt.Mesh()
    with pytest.raises(grt.GrtError):
        m.vertex(0, 0, 0)          # no open primitive
    with pytest.raises(grt.GrtError):
        m.begin(0x1234)            # unknown primitive
    m.begin(grt.TRIANGLES)
    with pytest.raises(grt.GrtError):
        m.begin(grt.TRIANGLES)     # nested
    m.end()
    with pytest.raises(grt.GrtError):
        m.end()


def test_obj_faces_normals_and_materials(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl red\nKd 0.9 0.1 0.05\nnewmtl blue\nKd 0 0 1\n")
    (tmp_path / "m.obj").write_text(
        "# a quad, a triangle by negative indices, a degenerate face\n"
        "mtllib m.mtl\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
        "vn 0 0 1\nvn 0 0 -1\n"
        "usemtl red\n"
        "f 1//1 2//1 3//2 4//2\n"
        "usemtl nosuch\n"
        "f -4/7 -3 -2\n"
        "usemtl blue\n"
        "f 1 1 2\n")
    m = grt.Mesh()
    m.load_obj(tmp_path / "m.obj")
    t, n, kd, src = _tris(m)
    V = np.float32([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    assert m.info() == (4, 0)
    assert np.array_equal(t[0], V[[0, 1, 2]]) and np.array_equal(t[1], V[[0, 2, 3]])
    assert np.array_equal(n[0], np.float32([[0, 0, 1], [0, 0, 1], [0, 0, -1]]))
    assert np.array_equal(n[1], np.float32([[0, 0, 1], [0, 0, -1], [0, 0, -1]]))
    assert np.array_equal(t[2], V[[0, 1, 2]]) and not n[2].any()   # no normals given
    assert kd.tolist()[:3] == [[np.float32(0.9), np.float32(0.1), np.float32(0.05)]] * 2 + [[1, 1, 1]]
    assert kd[3].tolist() == [0, 0, 1]
    m2 = grt.Mesh()
    m2.load_obj(tmp_path / "m.obj", drop_degenerate=True)
    assert m2.info() == (3, 1)
    (tmp_path / "bad.obj").write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(grt.GrtError):
        grt.Mesh().load_obj(tmp_path / "bad.obj")
    with pytest.raises(grt.GrtError):
        grt.Mesh().load_obj(tmp_path / "missing.obj")


def test_obj_loader_survives_mutated_files(tmp_path):
    """Random byte mutations of a valid OBJ either load (every vertex of every
    triangle a finite coordinate triple taken from the file's `v` lines, or
    dropped) or are refused with an error; the loader never crashes and a
    refused load leaves the mesh as it was."""
    base = ("mtllib m.mtl\nusemtl red\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
            "f 1//1 2//1 3//1 4//1\nf -1 -2 -3\n")
    (tmp_path / "m.mtl").write_text("newmtl red\nKd 1 0 0\n")
    r = np.random.default_rng(3)
    alphabet = list("0123456789 -/\nfvn#e.")
    loaded = refused = 0
    for k in range(300):
        s = list(base)
        for _ in range(int(r.integers(1, 8))):
            s[int(r.integers(len(s)))] = alphabet[int(r.integers(len(alphabet)))]
        path = tmp_path / f"f{k}.obj"
        path.write_text("".join(s))
        m = grt.Mesh()
        m.begin(grt.TRIANGLES)
        for p in ((5, 5, 5), (6, 5, 5), (5, 6, 5)):
            m.vertex(*p)
        m.end()
        try:
            m.load_obj(path, drop_degenerate=bool(k & 1))
        except grt.GrtError:
            refused += 1
            assert m.info() == (1, 0), "a refused load changed the mesh"
            continue
        loaded += 1
        t, n, kd, src = _tris(m)
        assert np.isfinite(t).all()
        assert len(t) == m.info()[0] >= 1
    assert loaded > 0 and refused > 0


def test_procedural_sphere_and_cube():
    m = grt.Mesh()
    m.add_sphere(16, 16)
    kept, dropped = m.info()
    assert kept + dropped == 2 * 16 * 8     # stacks/2 strips of 2*slices triangles
    t, n, _, _ = _tris(m)
    r = np.linalg.norm(t.reshape(-1, 3), axis=1)
    assert np.allclose(r, 1.0, atol=1e-6)
    assert np.allclose(n.reshape(-1, 3), t.reshape(-1, 3), atol=1e-6)   # normal = position
    c = grt.Mesh()
    c.set_transform([1, 0, 0, 10, 0, 1, 0, 0, 0, 0, 1, 0])
    c.add_cube()
    t, n, _, _ = _tris(c)
    assert c.info() == (12, 0)
    assert t[..., 0].min() == 9.5 and t[..., 0].max() == 10.5
    for k in range(12):
        e = np.cross(t[k, 1] - t[k, 0], t[k, 2] - t[k, 0])
        assert np.allclose(e / np.linalg.norm(e), n[k, 0])           # counter-clockwise outside
        assert np.dot(n[k, 0], t[k].mean(axis=0) - [10, 0, 0]) > 0   # normals point out


_BUILDER_PROG = r"""
#include <cstdio>
#include "grt/rt_compat.h"
// rt::Context-style immediate mode through the plugin-side PrimitiveBuilder.
int main()
{
    grt::ref_ptr<grt::Geometry> g(new grt::Geometry);
    grt::PrimitiveBuilder b;
    const float xf[12] = {2, 0, 0, 1, 0, 1, 0, 0, 0, 0, 1, 0};
    b.setMatrix(xf);
    b.begin(grt::RT_TRIANGLE_STRIP, g.get());
    b.setNormal(1, 1, 0);
    const float p[5][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
    for (auto& q : p)
        b.addVertex(q[0], q[1], q[2]);
    b.end();
    b.setMatrix(nullptr);
    b.begin(grt::RT_POLYGON, g.get());
    b.addVertex(0, 0, 5);
    b.addVertex(1, 0, 5);
    b.addVertex(1, 1, 5);
    b.addVertex(0, 1, 5);
    b.end();
    std::printf("%zu %zu\n", g->vertices.size(), g->triDesc.size());
    for (auto& t : g->triDesc)
        std::printf("%u %u %u\n", t.v[0], t.v[1], t.v[2]);
    std::printf("%a %a %a %a\n", g->vertices[1].x, g->normals[0].x, g->normals[0].y,
                g->normals[0].z);
    return 0;
}
"""


def test_plugin_primitive_builder(tmp_path):
    import os
    import subprocess
    from conftest import ROOT
    src = tmp_path / "pb.cpp"
    src.write_text(_BUILDER_PROG)
    exe = tmp_path / "pb"
    pkg = os.path.join(ROOT, "grid-rt_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", str(src), "-o", str(exe),
                    f"-L{pkg}", "-lgrt", f"-Wl,-rpath,{pkg}"], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.strip().splitlines()
    assert lines[0] == "9 3"
    # strip (0,1,2) and (3,2,1) repeat vertex 1 = 2 and are dropped; (2,3,4)
    # is kept, then the polygon's (5,6,7), (5,7,8)
    assert lines[1:4] == ["2 3 4", "5 6 7", "5 7 8"]
    vx, nx, ny, nz = (float.fromhex(x) for x in lines[4].split())
    assert vx == 3.0                                           # 2 * 1 + 1
    assert np.allclose([nx, ny, nz], np.array([0.5, 1, 0]) / math.sqrt(1.25), atol=1e-6)


# ---------------------------------------------------------------------------
# rtdb scene layouts (rtdb.cpp:9-57, 215-287)
# ---------------------------------------------------------------------------
UNIT_TRI = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)


def test_rtdb_triangle_and_logo():
    m = grt.Mesh()
    m.add_triangle()
    m.add_logo()
    t, n, _, _ = _tris(m)
    assert m.info() == (4, 0)
    np.testing.assert_array_equal(t[0], UNIT_TRI)
    logo = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[2, 0, 0], [3, 0, 0], [3, 1, 0]],
                     [[1.5, 0, 0], [2.5, 1, 0], [0.5, 1, 0]]], np.float32)
    np.testing.assert_array_equal(t[1:], logo)
    assert np.all(n.reshape(-1, 3) == [0, 0, 1])


def test_rtdb_even_spaced_post_multiplies_the_transform():
    """27 unit triangles at (1.5 x, 1.5 y, 1.5 z), x fastest; under a current
    transform M the layout's translation applies first: M(T(v))."""
    m = grt.Mesh()
    base = np.array([2, 0, 0, 10, 0, 2, 0, 0, 0, 0, 2, 0], np.float32)
    m.set_transform(base)
    m.add_even_spaced()
    t, n, _, _ = _tris(m)
    assert m.info() == (27, 0)
    k = 0
    for z in range(3):
        for y in range(3):
            for x in range(3):
                off = np.float32(1.5) * np.array([x, y, z], np.float32)
                want = 2 * (UNIT_TRI + off) + np.array([10, 0, 0], np.float32)
                np.testing.assert_array_equal(t[k], want)
                k += 1
    assert np.all(n.reshape(-1, 3) == [0, 0, 1])
    # the current transform is left as it was
    m.add_triangle()
    np.testing.assert_array_equal(_tris(m)[0][-1], 2 * UNIT_TRI + np.float32([10, 0, 0]))


def _model():
    """A stand-in for the teapot with no near-degenerate triangles (a sphere's
    pole triangles collapse once offset, and the filter then drops them): a
    cube and a fan."""
    s = grt.Mesh()
    s.set_color(0.2, 0.4, 0.6)
    s.set_transform([0.5, 0, 0, 0.25, 0, 0.25, 0, 0.5, 0, 0, 0.75, 0.125])
    s.add_cube()
    s.set_transform(None)
    s.primitive(grt.TRIANGLE_FAN, [[0, 0, 0], [1, 0, 0.25], [1, 1, 0.5], [0, 1, 0.25]])
    return s


def test_rtdb_mixed_scene():
    model = _model()
    mv, mn, _, _ = model.arrays()
    nm = mv.shape[0]
    m = grt.Mesh()
    m.add_mixed(model)
    t, n, kd, _ = _tris(m)
    assert m.info()[0] == 3 + 4 * 12 + 4 * nm
    # the logo at (1, 2, 2.5)
    np.testing.assert_array_equal(t[0], UNIT_TRI + np.float32([1, 2, 2.5]))
    # cubes centred at (0,5,0), (5,5,0), (5,5,5), (0,5,5): the translations accumulate
    for c, centre in enumerate([(0, 5, 0), (5, 5, 0), (5, 5, 5), (0, 5, 5)]):
        cube = t[3 + 12 * c: 3 + 12 * (c + 1)].reshape(-1, 3)
        np.testing.assert_array_equal(cube.min(axis=0), np.float32(centre) - 0.5)
        np.testing.assert_array_equal(cube.max(axis=0), np.float32(centre) + 0.5)
    # the model scaled by 5 after offsets (0,0,0), (5,0,0), (5,0,5), (0,0,5)
    first = 3 + 48
    for c, off in enumerate([(0, 0, 0), (5, 0, 0), (5, 0, 5), (0, 0, 5)]):
        copy = t[first + nm * c: first + nm * (c + 1)]
        want = np.float32(5) * (mv.reshape(-1, 3, 3) + np.float32(off))
        np.testing.assert_allclose(copy, want, rtol=0, atol=2e-6)
        np.testing.assert_allclose(n[first + nm * c: first + nm * (c + 1)],
                                   mn.reshape(-1, 3, 3), atol=1e-6)
        assert np.all(kd[first + nm * c: first + nm * (c + 1)] == np.float32([0.2, 0.4, 0.6]))


def test_rtdb_teapots125_from_an_obj_model(tmp_path):
    """125 copies of a model read from an OBJ (the teapot's place), each
    5 x (v + 2 (x, y, z)), x fastest."""
    src = _model()
    v, _, _, _ = src.arrays()
    lines = []
    for tri in v.reshape(-1, 3, 3):
        for p in tri:
            lines.append("v %.9g %.9g %.9g" % tuple(p))
    for k in range(v.shape[0]):
        lines.append("f %d %d %d" % (3 * k + 1, 3 * k + 2, 3 * k + 3))
    (tmp_path / "model.obj").write_text("\n".join(lines) + "\n")
    model = grt.Mesh()
    model.load_obj(tmp_path / "model.obj", drop_degenerate=True)
    mv = model.arrays()[0].reshape(-1, 3, 3)
    nm = mv.shape[0]
    assert nm == v.shape[0]
    m = grt.Mesh()
    m.add_teapots125(model)
    t = _tris(m)[0]
    assert m.info()[0] == 125 * nm
    k = 0
    for z in range(5):
        for y in range(5):
            for x in range(5):
                want = np.float32(5) * (mv + np.float32([2 * x, 2 * y, 2 * z]))
                np.testing.assert_allclose(t[nm * k: nm * (k + 1)], want, rtol=0, atol=1e-5)
                k += 1
    with pytest.raises(grt.GrtError):
        m.add_mesh(m)
