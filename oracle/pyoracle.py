"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as parity reference / CPU baseline.  The product
(libgrt.so and its Python bindings) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_F = ctypes.c_float


class Frame(ctypes.Structure):
    _fields_ = [("bmin", _F * 3), ("bmax", _F * 3), ("cell", _F * 3), ("inv_cell", _F * 3),
                ("n", ctypes.c_int32 * 3), ("m", _U32)]


class Camera(ctypes.Structure):
    _fields_ = [("pos", _F * 3), ("base", _F * 3), ("nu", _F * 3), ("nv", _F * 3),
                ("width", _U32), ("height", _U32)]


class Phong(ctypes.Structure):
    _fields_ = [("light", _F * 3), ("light_color", _F * 3), ("ambient", _F * 3),
                ("specular", _F * 3), ("spec_exp", _U32), ("shadow_coeff", _F),
                ("reflect_coeff", _F), ("background", _F * 3)]


# shading_FRAG.glsl:8-36: background 1, ambient 0, specular 0.5, exponent 50,
# shadow 0.2, reflection 0.4, white light at (1e10, 1e10, 1e10).
PHONG_DEFAULTS = {"light": (1e10, 1e10, 1e10), "light_color": (1.0, 1.0, 1.0),
                  "ambient": (0.0, 0.0, 0.0), "specular": (0.5, 0.5, 0.5), "spec_exp": 50,
                  "shadow_coeff": 0.2, "reflect_coeff": 0.4, "background": (1.0, 1.0, 1.0)}


def phong_struct(params: dict | None = None) -> Phong:
    d = dict(PHONG_DEFAULTS)
    d.update(params or {})
    p = Phong()
    for k in ("light", "light_color", "ambient", "specular", "background"):
        getattr(p, k)[:] = [float(x) for x in d[k]]
    p.spec_exp = int(d["spec_exp"])
    p.shadow_coeff = float(d["shadow_coeff"])
    p.reflect_coeff = float(d["reflect_coeff"])
    return p


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_cbrtf": (_F, [_F]),
            "orc_grid_frame": (ctypes.c_int, [_P, _U32, _F, ctypes.POINTER(Frame)]),
            "orc_grid_frame_from_box": (ctypes.c_int, [_F * 3, _F * 3, _U32, _F,
                                                       ctypes.POINTER(Frame)]),
            "orc_tri_cell_range": (None, [ctypes.POINTER(Frame), _P, _P, _P]),
            "orc_grid_count": (ctypes.c_uint64, [ctypes.POINTER(Frame), _P, _U32, _P]),
            "orc_grid_fill": (None, [ctypes.POINTER(Frame), _P, _U32, _P, _P, _P]),
            "orc_camera_lookat": (None, [_F * 3, _F * 3, _F * 3, _F, _F, _U32, _U32,
                                         ctypes.POINTER(Camera)]),
            "orc_primary_rays": (None, [ctypes.POINTER(Camera)] + [_P] * 8),
            "orc_trace_closest": (None, [ctypes.POINTER(Frame), _P, _P, _P, _U32] + [_P] * 8
                                  + [_P] * 4 + [_P, ctypes.c_int, ctypes.c_int]),
            "orc_trace_any": (None, [ctypes.POINTER(Frame), _P, _P, _P, _U32] + [_P] * 8
                              + [_P, _P, ctypes.c_int, ctypes.c_int]),
            "orc_grid2_plan": (ctypes.c_uint64, [ctypes.POINTER(Frame), _P, _U32, _F, _U32, _P]),
            "orc_grid2_count": (ctypes.c_uint64, [ctypes.POINTER(Frame), _P, _P, _P, _P, _P,
                                                  ctypes.c_uint64]),
            "orc_grid2_fill": (None, [ctypes.POINTER(Frame), _P, _P, _P, _P, _P,
                                      ctypes.c_uint64, _P, _P]),
            "orc_grid2_trace_closest": (None, [ctypes.POINTER(Frame)] + [_P] * 6 + [_U32]
                                        + [_P] * 8 + [_P] * 4 + [_P, ctypes.c_int, ctypes.c_int]),
            "orc_grid2_trace_any": (None, [ctypes.POINTER(Frame)] + [_P] * 6 + [_U32]
                                    + [_P] * 8 + [_P, _P, ctypes.c_int, ctypes.c_int]),
            "orc_kd_trace_closest": (None, [_P, _P, _P, _P, _U32] + [_P] * 8 + [_P] * 4
                                     + [_P, ctypes.c_int, ctypes.c_int]),
            "orc_kd_trace_any": (None, [_P, _P, _P, _P, _U32] + [_P] * 8
                                 + [_P, _P, ctypes.c_int, ctypes.c_int]),
            "orc_dda_cells": (_U32, [ctypes.POINTER(Frame), _F * 3, _F * 3, _F, _F, _P, _U32]),
            "orc_mt": (ctypes.c_int, [_P, _F * 3, _F * 3, _F, _F, ctypes.POINTER(_F),
                                      ctypes.POINTER(_F), ctypes.POINTER(_F)]),
            "orc_shading_normal": (None, [_P, _F, _F, _F * 3]),
            "orc_gen_shadow": (None, [_U32] + [_P] * 6 + [_P] * 4 + [_P, _F * 3, _F] + [_P] * 8),
            "orc_gen_reflection": (None, [_U32] + [_P] * 6 + [_P] * 4 + [_P, _F] + [_P] * 8),
            "orc_gen_ao": (None, [_U32, _U32, ctypes.c_uint64] + [_P] * 6 + [_P] * 4
                           + [_P, _F, _F] + [_P] * 8),
            "orc_sincos_turns": (None, [_F, ctypes.POINTER(_F), ctypes.POINTER(_F)]),
            "orc_uniform24": (_F, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
            "orc_transform": (None, [_U32, _U32, _U32, _P, _P, _P, _P, _P]),
            "orc_brute_closest_f64": (None, [_P, _U32, _U32] + [_P] * 8 + [_P] * 5
                                      + [ctypes.c_int]),
            "orc_brute_any_f64": (None, [_P, _U32, _U32] + [_P] * 8 + [_P, _P, ctypes.c_int]),
            "orc_headlight_rgba8": (_U32, [_F * 3, _P, ctypes.c_int32, _F, _F]),
            "orc_headlight_image": (None, [_U32, _P, _P, _P, _P, _P, _P, _P, _P]),
            "orc_pow_u32": (_F, [_F, _U32]),
            "orc_shade_phong": (None, [_U32] + [_P] * 6 + [_P] * 4 + [_P, _P, _P,
                                ctypes.POINTER(Phong), _U32, _P, _P]),
            "orc_now": (ctypes.c_double, []),
            "orc_max_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")


def frame_dict(f: Frame) -> dict:
    return {"bmin": np.array(f.bmin[:], np.float32), "bmax": np.array(f.bmax[:], np.float32),
            "cell": np.array(f.cell[:], np.float32), "inv_cell": np.array(f.inv_cell[:], np.float32),
            "n": np.array(f.n[:], np.int32), "m": int(f.m)}


def frame_struct(d: dict) -> Frame:
    f = Frame()
    for k in ("bmin", "bmax", "cell", "inv_cell"):
        getattr(f, k)[:] = [float(x) for x in d[k]]
    f.n[:] = [int(x) for x in d["n"]]
    f.m = int(np.prod(np.asarray(d["n"], np.int64)))
    return f


def cbrtf(x: float) -> float:
    return lib().orc_cbrtf(x)


def grid_frame(verts: np.ndarray, lam: float) -> dict:
    verts = _f32(verts).reshape(-1, 9)
    f = Frame()
    rc = lib().orc_grid_frame(_p(verts), verts.shape[0], lam, ctypes.byref(f))
    if rc != 0:
        raise ValueError(f"orc_grid_frame failed ({rc})")
    return frame_dict(f)


def grid_frame_from_box(bmin, bmax, n_tris: int, lam: float) -> dict:
    f = Frame()
    f3 = _F * 3
    rc = lib().orc_grid_frame_from_box(f3(*bmin), f3(*bmax), n_tris, lam, ctypes.byref(f))
    if rc != 0:
        raise ValueError(f"orc_grid_frame_from_box failed ({rc})")
    return frame_dict(f)


def tri_cell_range(frame: dict, tri9) -> tuple:
    f = frame_struct(frame)
    v = _f32(tri9).reshape(9)
    lo = np.zeros(3, np.int32)
    hi = np.zeros(3, np.int32)
    lib().orc_tri_cell_range(ctypes.byref(f), _p(v), _p(lo), _p(hi))
    return lo, hi


def grid_build(verts: np.ndarray, frame: dict):
    """(counts, offsets, refs) of the reference CPU builder."""
    verts = _f32(verts).reshape(-1, 9)
    f = frame_struct(frame)
    counts = np.empty(f.m, np.uint32)
    r = lib().orc_grid_count(ctypes.byref(f), _p(verts), verts.shape[0], _p(counts))
    offsets = np.empty(f.m + 1, np.uint32)
    refs = np.empty(max(int(r), 1), np.uint32)
    lib().orc_grid_fill(ctypes.byref(f), _p(verts), verts.shape[0], _p(counts), _p(offsets),
                        _p(refs))
    return counts, offsets, refs[:int(r)]


def camera_lookat(eye, center, up, fovy: float, width: int, height: int, znear: float = 1.0):
    c = Camera()
    f3 = _F * 3
    lib().orc_camera_lookat(f3(*eye), f3(*center), f3(*up), fovy, znear, width, height,
                            ctypes.byref(c))
    return {"pos": list(c.pos), "base": list(c.base), "nu": list(c.nu), "nv": list(c.nv),
            "width": c.width, "height": c.height}


def _cam_struct(cam: dict) -> Camera:
    c = Camera()
    for k in ("pos", "base", "nu", "nv"):
        getattr(c, k)[:] = [float(x) for x in cam[k]]
    c.width = int(cam["width"])
    c.height = int(cam["height"])
    return c


def primary_rays(cam: dict) -> dict:
    c = _cam_struct(cam)
    n = c.width * c.height
    out = {k: np.empty(n, np.float32) for k in RAY_FIELDS}
    lib().orc_primary_rays(ctypes.byref(c), *[_p(out[k]) for k in RAY_FIELDS])
    return out


def _rays(r: dict):
    return {k: _f32(r[k]) for k in RAY_FIELDS}


def trace_closest(frame: dict, offsets, refs, verts, rays: dict, threads: int = 0,
                  counters: bool = False, chunk: int = 64):
    f = frame_struct(frame)
    verts = _f32(verts).reshape(-1, 9)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    refs = np.ascontiguousarray(refs, np.uint32)
    r = _rays(rays)
    n = r["ox"].shape[0]
    hid = np.empty(n, np.int32)
    ht, hu, hv = (np.empty(n, np.float32) for _ in range(3))
    cnt = np.empty((n, 2), np.uint32) if counters else None
    threads = threads or lib().orc_max_threads()
    lib().orc_trace_closest(ctypes.byref(f), _p(offsets), _p(refs), _p(verts), n,
                            *[_p(r[k]) for k in RAY_FIELDS], _p(hid), _p(ht), _p(hu), _p(hv),
                            _p(cnt), threads, chunk)
    out = {"id": hid, "t": ht, "u": hu, "v": hv}
    if counters:
        out["cells"] = cnt[:, 0].copy()
        out["tests"] = cnt[:, 1].copy()
    return out


def trace_any(frame: dict, offsets, refs, verts, rays: dict, threads: int = 0,
              counters: bool = False, chunk: int = 64):
    f = frame_struct(frame)
    verts = _f32(verts).reshape(-1, 9)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    refs = np.ascontiguousarray(refs, np.uint32)
    r = _rays(rays)
    n = r["ox"].shape[0]
    occ = np.empty(n, np.uint8)
    cnt = np.empty((n, 2), np.uint32) if counters else None
    threads = threads or lib().orc_max_threads()
    lib().orc_trace_any(ctypes.byref(f), _p(offsets), _p(refs), _p(verts), n,
                        *[_p(r[k]) for k in RAY_FIELDS], _p(occ), _p(cnt), threads, chunk)
    out = {"occluded": occ}
    if counters:
        out["cells"] = cnt[:, 0].copy()
        out["tests"] = cnt[:, 1].copy()
    return out


def grid2_build(verts, frame: dict, offsets, refs, min_refs: int, lambda2: float, rmax: int):
    """Sub-grids of a built top grid: (cellsub [m, 2], sub_offsets, sub_refs)."""
    verts = _f32(verts).reshape(-1, 9)
    f = frame_struct(frame)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    refs = np.ascontiguousarray(refs, np.uint32)
    cellsub = np.empty((f.m, 2), np.uint32)
    n_sub = lib().orc_grid2_plan(ctypes.byref(f), _p(offsets), min_refs, lambda2, rmax,
                                 _p(cellsub))
    counts = np.empty(max(int(n_sub), 1), np.uint32)
    r2 = lib().orc_grid2_count(ctypes.byref(f), _p(offsets), _p(refs), _p(verts), _p(cellsub),
                               _p(counts), n_sub)
    sub_offsets = np.empty(int(n_sub) + 1, np.uint32)
    sub_refs = np.empty(max(int(r2), 1), np.uint32)
    lib().orc_grid2_fill(ctypes.byref(f), _p(offsets), _p(refs), _p(verts), _p(cellsub),
                         _p(counts), n_sub, _p(sub_offsets), _p(sub_refs))
    return cellsub, sub_offsets, sub_refs[:int(r2)]


def _grid2_call(name, frame, offsets, refs, g2, verts, rays, extra, threads, chunk):
    f = frame_struct(frame)
    cellsub, sub_offsets, sub_refs = g2
    arrs = [np.ascontiguousarray(a, np.uint32) for a in (offsets, refs, cellsub, sub_offsets,
                                                        sub_refs if len(sub_refs) else
                                                        np.zeros(1, np.uint32))]
    verts = _f32(verts).reshape(-1, 9)
    r = _rays(rays)
    n = r["ox"].shape[0]
    threads = threads or lib().orc_max_threads()
    getattr(lib(), name)(ctypes.byref(f), *[_p(a) for a in arrs], _p(verts), n,
                         *[_p(r[k]) for k in RAY_FIELDS], *[_p(a) for a in extra], threads, chunk)
    return n


def grid2_trace_closest(frame, offsets, refs, g2, verts, rays: dict, threads: int = 0,
                        counters: bool = False, chunk: int = 64):
    n = rays["ox"].shape[0]
    hid = np.empty(n, np.int32)
    ht, hu, hv = (np.empty(n, np.float32) for _ in range(3))
    cnt = np.empty((n, 2), np.uint32) if counters else None
    _grid2_call("orc_grid2_trace_closest", frame, offsets, refs, g2, verts, rays,
                [hid, ht, hu, hv, cnt], threads, chunk)
    out = {"id": hid, "t": ht, "u": hu, "v": hv}
    if counters:
        out["cells"] = cnt[:, 0].copy()
        out["tests"] = cnt[:, 1].copy()
    return out


def grid2_trace_any(frame, offsets, refs, g2, verts, rays: dict, threads: int = 0,
                    counters: bool = False, chunk: int = 64):
    n = rays["ox"].shape[0]
    occ = np.empty(n, np.uint8)
    cnt = np.empty((n, 2), np.uint32) if counters else None
    _grid2_call("orc_grid2_trace_any", frame, offsets, refs, g2, verts, rays, [occ, cnt],
                threads, chunk)
    out = {"occluded": occ}
    if counters:
        out["cells"] = cnt[:, 0].copy()
        out["tests"] = cnt[:, 1].copy()
    return out


def _kd_args(box, nodes, elements):
    box = np.ascontiguousarray(np.concatenate([np.asarray(box[0], np.float32),
                                               np.asarray(box[1], np.float32)]), np.float32)
    nodes = np.ascontiguousarray(nodes, np.uint32).reshape(-1)
    elements = np.ascontiguousarray(elements, np.uint32)
    if elements.size == 0:
        elements = np.zeros(1, np.uint32)
    return box, nodes, elements


def kd_trace_closest(box, nodes, elements, verts, rays: dict, threads: int = 0,
                     counters: bool = False, chunk: int = 64):
    """Closest hit through a flattened kd-tree; box = (bmin, bmax)."""
    box, nodes, elements = _kd_args(box, nodes, elements)
    verts = _f32(verts).reshape(-1, 9)
    r = _rays(rays)
    n = r["ox"].shape[0]
    hid = np.empty(n, np.int32)
    ht, hu, hv = (np.empty(n, np.float32) for _ in range(3))
    cnt = np.empty((n, 2), np.uint32) if counters else None
    threads = threads or lib().orc_max_threads()
    lib().orc_kd_trace_closest(_p(box), _p(nodes), _p(elements), _p(verts), n,
                               *[_p(r[k]) for k in RAY_FIELDS], _p(hid), _p(ht), _p(hu), _p(hv),
                               _p(cnt), threads, chunk)
    out = {"id": hid, "t": ht, "u": hu, "v": hv}
    if counters:
        out["leaves"] = cnt[:, 0].copy()
        out["tests"] = cnt[:, 1].copy()
    return out


def kd_trace_any(box, nodes, elements, verts, rays: dict, threads: int = 0,
                 counters: bool = False, chunk: int = 64):
    box, nodes, elements = _kd_args(box, nodes, elements)
    verts = _f32(verts).reshape(-1, 9)
    r = _rays(rays)
    n = r["ox"].shape[0]
    occ = np.empty(n, np.uint8)
    cnt = np.empty((n, 2), np.uint32) if counters else None
    threads = threads or lib().orc_max_threads()
    lib().orc_kd_trace_any(_p(box), _p(nodes), _p(elements), _p(verts), n,
                           *[_p(r[k]) for k in RAY_FIELDS], _p(occ), _p(cnt), threads, chunk)
    out = {"occluded": occ}
    if counters:
        out["leaves"] = cnt[:, 0].copy()
        out["tests"] = cnt[:, 1].copy()
    return out


def dda_cells(frame: dict, o, d, tmin=0.0, tmax=np.inf, max_cells: int = 1 << 16):
    f = frame_struct(frame)
    out = np.empty(max_cells, np.uint32)
    f3 = _F * 3
    k = lib().orc_dda_cells(ctypes.byref(f), f3(*o), f3(*d), tmin, tmax, _p(out), max_cells)
    return out[:min(k, max_cells)]


def mt(tri9, o, d, tmin=0.0, tmax=np.inf):
    v = _f32(tri9).reshape(9)
    f3 = _F * 3
    t, u, w = _F(), _F(), _F()
    hit = lib().orc_mt(_p(v), f3(*o), f3(*d), tmin, tmax, ctypes.byref(t), ctypes.byref(u),
                       ctypes.byref(w))
    return bool(hit), t.value, u.value, w.value


def shading_normal(n9, u, v):
    n9 = _f32(n9).reshape(9)
    out = (_F * 3)()
    lib().orc_shading_normal(_p(n9), u, v, out)
    return np.array(out[:], np.float32)


def _hits(h: dict):
    return (np.ascontiguousarray(h["id"], np.int32), _f32(h["t"]), _f32(h["u"]), _f32(h["v"]))


def gen_shadow(rays: dict, hits: dict, normals, light, offset: float) -> dict:
    r = _rays(rays)
    n = r["ox"].shape[0]
    hid, ht, hu, hv = _hits(hits)
    normals = _f32(normals).reshape(-1, 9)
    out = {k: np.empty(n, np.float32) for k in RAY_FIELDS}
    lib().orc_gen_shadow(n, *[_p(r[k]) for k in RAY_FIELDS[:6]], _p(hid), _p(ht), _p(hu), _p(hv),
                         _p(normals), (_F * 3)(*light), offset, *[_p(out[k]) for k in RAY_FIELDS])
    return out


def gen_reflection(rays: dict, hits: dict, normals, offset: float) -> dict:
    r = _rays(rays)
    n = r["ox"].shape[0]
    hid, ht, hu, hv = _hits(hits)
    normals = _f32(normals).reshape(-1, 9)
    out = {k: np.empty(n, np.float32) for k in RAY_FIELDS}
    lib().orc_gen_reflection(n, *[_p(r[k]) for k in RAY_FIELDS[:6]], _p(hid), _p(ht), _p(hu),
                             _p(hv), _p(normals), offset, *[_p(out[k]) for k in RAY_FIELDS])
    return out


def gen_ao(rays: dict, hits: dict, normals, spp: int, seed: int, offset: float,
           ao_tmax: float) -> dict:
    r = _rays(rays)
    n = r["ox"].shape[0]
    hid, ht, hu, hv = _hits(hits)
    normals = _f32(normals).reshape(-1, 9)
    out = {k: np.empty(n * spp, np.float32) for k in RAY_FIELDS}
    lib().orc_gen_ao(n, spp, seed, *[_p(r[k]) for k in RAY_FIELDS[:6]], _p(hid), _p(ht), _p(hu),
                     _p(hv), _p(normals), offset, ao_tmax, *[_p(out[k]) for k in RAY_FIELDS])
    return out


def sincos_turns(x: float):
    s, c = _F(), _F()
    lib().orc_sincos_turns(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def uniform24(seed: int, a: int, b: int) -> float:
    return lib().orc_uniform24(seed, a, b)


def transform(rest_verts, rest_normals, xf, tris_per_cluster: int, n_cluster_tris: int):
    rv = _f32(rest_verts).reshape(-1, 9)
    rn = _f32(rest_normals).reshape(-1, 9)
    xf = _f32(xf).reshape(-1, 12)
    v = np.empty_like(rv)
    n = np.empty_like(rn)
    lib().orc_transform(rv.shape[0], tris_per_cluster, n_cluster_tris, _p(xf), _p(rv), _p(rn),
                        _p(v), _p(n))
    return v, n


def brute_closest_f64(verts, rays: dict, threads: int = 0) -> dict:
    verts = _f32(verts).reshape(-1, 9)
    r = _rays(rays)
    n = r["ox"].shape[0]
    hid = np.empty(n, np.int32)
    t1, t2, u, v = (np.empty(n, np.float64) for _ in range(4))
    threads = threads or lib().orc_max_threads()
    lib().orc_brute_closest_f64(_p(verts), verts.shape[0], n, *[_p(r[k]) for k in RAY_FIELDS],
                                _p(hid), _p(t1), _p(t2), _p(u), _p(v), threads)
    return {"id": hid, "t1": t1, "t2": t2, "u": u, "v": v}


def brute_any_f64(verts, rays: dict, threads: int = 0) -> dict:
    verts = _f32(verts).reshape(-1, 9)
    r = _rays(rays)
    n = r["ox"].shape[0]
    occ = np.empty(n, np.uint8)
    tn = np.empty(n, np.float64)
    threads = threads or lib().orc_max_threads()
    lib().orc_brute_any_f64(_p(verts), verts.shape[0], n, *[_p(r[k]) for k in RAY_FIELDS],
                            _p(occ), _p(tn), threads)
    return {"occluded": occ, "t_near": tn}


def headlight(d, n9, hid: int, u: float, v: float) -> int:
    return lib().orc_headlight_rgba8((_F * 3)(*d), _p(_f32(n9).reshape(9)), hid, u, v)


def headlight_image(rays: dict, hits: dict, normals) -> np.ndarray:
    """RGBA8 headlight image (uint32 per ray) of closest-hit records."""
    n = rays["dx"].shape[0]
    out = np.empty(n, np.uint32)
    nrm = _f32(normals).reshape(-1)
    ids = np.ascontiguousarray(hits["id"], np.int32)
    d = [_f32(rays[k]) for k in ("dx", "dy", "dz")]
    uv = [_f32(hits[k]) for k in ("u", "v")]
    lib().orc_headlight_image(n, _p(d[0]), _p(d[1]), _p(d[2]), _p(nrm), _p(ids), _p(uv[0]),
                              _p(uv[1]), _p(out))
    return out


def pow_u32(x: float, e: int) -> float:
    return lib().orc_pow_u32(x, e)


def shade_phong(rays: dict, hits: dict, normals, occluded=None, params: dict | None = None,
                depth: int = 0, rgb=None, diffuse=None):
    """Phong shading of closest-hit records: returns (rgb float32 [n, 3], rgba8
    uint32 [n]).  depth 0 writes rgb; depth d > 0 adds reflect_coeff^d x colour
    for hits into a copy of `rgb`."""
    n = rays["dx"].shape[0]
    out = np.zeros((n, 3), np.float32) if rgb is None else np.array(rgb, np.float32).reshape(n, 3)
    img = np.empty(n, np.uint32)
    r = [_f32(rays[k]) for k in ("ox", "oy", "oz", "dx", "dy", "dz")]
    h = [np.ascontiguousarray(hits["id"], np.int32)] + [_f32(hits[k]) for k in ("t", "u", "v")]
    occ = None if occluded is None else np.ascontiguousarray(occluded, np.uint8)
    kd = None if diffuse is None else _f32(diffuse).reshape(-1)
    p = phong_struct(params)
    lib().orc_shade_phong(n, *[_p(a) for a in r], *[_p(a) for a in h], _p(occ),
                          _p(_f32(normals).reshape(-1)), _p(kd), ctypes.byref(p), depth,
                          _p(out), _p(img))
    return out, img


def now() -> float:
    return lib().orc_now()


def max_threads() -> int:
    return lib().orc_max_threads()
