"""grt -- Python bindings (ctypes) over libgrt.so, the gfx950 uniform-grid tracer.

The library's C ABI (include/grt/grt.h) takes plain device pointers; here they
come from torch tensors on ``cuda:0`` (torch is only plumbing for device memory
and the stream).  Importing this module never needs a GPU; creating a
:class:`Context` does, and fails loudly (GrtError) when the HIP library or the
device is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "libgrt.so")

GRT_OK = 0
GRT_NO_MEMORY = 1
GRT_INVALID_PARAM = 3
GRT_DEVICE_ERROR = 4
GRT_NOT_READY = 5
GRT_CAPACITY = 6

RAY_EPSILON = np.float32(2e-4)
RAY_FIELDS = ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")
HIT_FIELDS = ("id", "t", "u", "v")


class GrtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"grt error {code}: {msg}")
        self.code = code


class Rays(ctypes.Structure):
    _fields_ = [(f, ctypes.c_void_p) for f in RAY_FIELDS]


class Hits(ctypes.Structure):
    _fields_ = [(f, ctypes.c_void_p) for f in HIT_FIELDS]


class Camera(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("base", ctypes.c_float * 3),
                ("nu", ctypes.c_float * 3), ("nv", ctypes.c_float * 3),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32)]


class GridInfo(ctypes.Structure):
    _fields_ = [("bmin", ctypes.c_float * 3), ("bmax", ctypes.c_float * 3),
                ("cell", ctypes.c_float * 3), ("inv_cell", ctypes.c_float * 3),
                ("n", ctypes.c_int32 * 3), ("m", ctypes.c_uint32), ("r", ctypes.c_uint32)]


class SubgridInfo(ctypes.Structure):
    _fields_ = [("lambda2", ctypes.c_float), ("min_refs", ctypes.c_uint32),
                ("built", ctypes.c_uint32), ("cells", ctypes.c_uint32),
                ("refs", ctypes.c_uint32), ("status", ctypes.c_uint32)]


class KdTreeInfo(ctypes.Structure):
    _fields_ = [("bmin", ctypes.c_float * 3), ("bmax", ctypes.c_float * 3),
                ("nodes", ctypes.c_uint32), ("leaves", ctypes.c_uint32),
                ("empty_leaves", ctypes.c_uint32), ("elements", ctypes.c_uint32),
                ("depth", ctypes.c_uint32), ("build_ms", ctypes.c_double)]


class Instance(ctypes.Structure):
    """grt_instance: a scene and its row-major 3x4 object-to-world matrix."""
    _fields_ = [("scene", ctypes.c_void_p), ("xf", ctypes.c_float * 12)]


def affine_inverse(xf) -> np.ndarray:
    """The binary32 inverse grt_trace_instances takes rays through."""
    a = np.ascontiguousarray(xf, np.float32).reshape(12)
    out = np.empty(12, np.float32)
    check(lib().grt_affine_inverse(_ptr(a), _ptr(out)))
    return out


class Phong(ctypes.Structure):
    """grt_phong: Phong shading parameters (shading_FRAG.glsl:8-36)."""
    _fields_ = [("light", ctypes.c_float * 3), ("light_color", ctypes.c_float * 3),
                ("ambient", ctypes.c_float * 3), ("specular", ctypes.c_float * 3),
                ("spec_exp", ctypes.c_uint32), ("shadow_coeff", ctypes.c_float),
                ("reflect_coeff", ctypes.c_float), ("background", ctypes.c_float * 3)]


_lib = None

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_F = ctypes.c_float
_SIGS = {
    "grt_last_error": (ctypes.c_char_p, []),
    "grt_grid_set_subgrid": (ctypes.c_int, [_P, _F, _U32]),
    "grt_subgrid_info_get": (ctypes.c_int, [_P, ctypes.POINTER(SubgridInfo)]),
    "grt_subgrid_download": (ctypes.c_int, [_P, _P, _P, _P]),
    "grt_kdtree_build": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "grt_kdtree_build_host": (ctypes.c_int, [_P, _U32, ctypes.POINTER(_P)]),
    "grt_kdtree_destroy": (ctypes.c_int, [_P]),
    "grt_kdtree_info_get": (ctypes.c_int, [_P, ctypes.POINTER(KdTreeInfo)]),
    "grt_kdtree_download": (ctypes.c_int, [_P, _P, _P]),
    "grt_kdtree_trace_closest": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits),
                                                _U32, _U32, _P]),
    "grt_kdtree_trace_any": (ctypes.c_int, [_P, ctypes.POINTER(Rays), _P, _U32, _U32, _P]),
    "grt_init": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "grt_shutdown": (ctypes.c_int, [_P]),
    "grt_set_stream": (ctypes.c_int, [_P, _P]),
    "grt_synchronize": (ctypes.c_int, [_P]),
    "grt_memory_used": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "grt_set_block_size": (ctypes.c_int, [_P, _U32, _U32]),
    "grt_set_trace_mode": (ctypes.c_int, [_P, ctypes.c_int, _U32]),
    "grt_set_coop": (ctypes.c_int, [_P, _U32]),
    "grt_scene_create": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "grt_scene_destroy": (ctypes.c_int, [_P]),
    "grt_scene_upload_triangles": (ctypes.c_int, [_P, _P, _U32]),
    "grt_scene_upload_normals": (ctypes.c_int, [_P, _P, _U32]),
    "grt_scene_set_triangles_device": (ctypes.c_int, [_P, _P, _P, _U32]),
    "grt_scene_triangle_count": (ctypes.c_int, [_P, ctypes.POINTER(_U32)]),
    "grt_scene_device_pointers": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "grt_scene_transform_clusters": (ctypes.c_int, [_P, _P, _U32, _U32, _U32]),
    "grt_scene_download_geometry": (ctypes.c_int, [_P, _P, _P]),
    "grt_grid_build": (ctypes.c_int, [_P, _F]),
    "grt_grid_rebuild_async": (ctypes.c_int, [_P, _F]),
    "grt_grid_reserve": (ctypes.c_int, [_P, _U32, _U32]),
    "grt_grid_check": (ctypes.c_int, [_P]),
    "grt_grid_upload": (ctypes.c_int, [_P, ctypes.POINTER(GridInfo), _P, _P]),
    "grt_grid_info_get": (ctypes.c_int, [_P, ctypes.POINTER(GridInfo)]),
    "grt_grid_download": (ctypes.c_int, [_P, _P, _P]),
    "grt_trace_closest": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _U32,
                                         _U32, _P]),
    "grt_trace_any": (ctypes.c_int, [_P, ctypes.POINTER(Rays), _P, _U32, _U32, _P]),
    "grt_camera_lookat": (ctypes.c_int, [_F * 3, _F * 3, _F * 3, _F, _F, _U32, _U32,
                                         ctypes.POINTER(Camera)]),
    "grt_gen_primary": (ctypes.c_int, [_P, ctypes.POINTER(Camera), ctypes.POINTER(Rays)]),
    "grt_gen_shadow": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _U32,
                                      _F * 3, _F, ctypes.POINTER(Rays)]),
    "grt_gen_reflection": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _U32,
                                          _F, ctypes.POINTER(Rays)]),
    "grt_gen_ao": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _U32, _U32,
                                  ctypes.c_uint64, _F, _F, ctypes.POINTER(Rays)]),
    "grt_shade_headlight": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _U32,
                                           _P]),
    "grt_phong_defaults": (None, [ctypes.POINTER(Phong)]),
    "grt_mesh_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "grt_mesh_destroy": (ctypes.c_int, [_P]),
    "grt_mesh_clear": (ctypes.c_int, [_P]),
    "grt_mesh_set_transform": (ctypes.c_int, [_P, _P]),
    "grt_mesh_set_color": (ctypes.c_int, [_P, _F, _F, _F]),
    "grt_mesh_begin": (ctypes.c_int, [_P, ctypes.c_int]),
    "grt_mesh_normal": (ctypes.c_int, [_P, _F, _F, _F]),
    "grt_mesh_vertex": (ctypes.c_int, [_P, _F, _F, _F]),
    "grt_mesh_end": (ctypes.c_int, [_P]),
    "grt_mesh_load_obj": (ctypes.c_int, [_P, ctypes.c_char_p, _U32]),
    "grt_mesh_add_sphere": (ctypes.c_int, [_P, _U32, _U32]),
    "grt_mesh_add_cube": (ctypes.c_int, [_P]),
    "grt_mesh_add_triangle": (ctypes.c_int, [_P]),
    "grt_mesh_add_logo": (ctypes.c_int, [_P]),
    "grt_mesh_add_mesh": (ctypes.c_int, [_P, _P]),
    "grt_mesh_add_even_spaced": (ctypes.c_int, [_P]),
    "grt_mesh_add_mixed": (ctypes.c_int, [_P, _P]),
    "grt_mesh_add_teapots125": (ctypes.c_int, [_P, _P]),
    "grt_mesh_info": (ctypes.c_int, [_P, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "grt_mesh_copy": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "grt_scene_upload_mesh": (ctypes.c_int, [_P, _P]),
    "grt_triangle_valid": (ctypes.c_int, [_P]),
    "grt_affine_inverse": (ctypes.c_int, [_P, _P]),
    "grt_compact_rays": (ctypes.c_int, [_P, ctypes.POINTER(Rays), _U32, _U32, _P, _P, _P,
                                        ctypes.POINTER(Hits)]),
    "grt_compact_rays_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(Rays), _U32, _U32, _P, _P,
                                                _P, ctypes.POINTER(Hits)]),
    "grt_trace_closest_list": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _P,
                                              _P, _U32, _P]),
    "grt_trace_any_list": (ctypes.c_int, [_P, ctypes.POINTER(Rays), _P, _P, _P, _U32, _P]),
    "grt_trace_instances": (ctypes.c_int, [_P, _P, _U32, ctypes.POINTER(Rays), ctypes.POINTER(Hits),
                                           _P, _U32, _U32]),
    "grt_shade_phong": (ctypes.c_int, [_P, ctypes.POINTER(Rays), ctypes.POINTER(Hits), _P, _U32,
                                       _P, ctypes.POINTER(Phong), _U32, _P, _P]),
    "grt_debug_scan": (ctypes.c_int, [_P, _P, _P, _U32]),
    "grt_batch_create": (ctypes.c_int, [_P, _U32, ctypes.POINTER(_P)]),
    "grt_batch_set": (ctypes.c_int, [_P, _P, _U32]),
    "grt_batch_trace": (ctypes.c_int, [_P]),
    "grt_batch_slots": (ctypes.c_int, [_P, ctypes.POINTER(_U32)]),
    "grt_batch_destroy": (ctypes.c_int, [_P]),
    "grt_graph_begin": (ctypes.c_int, [_P]),
    "grt_graph_end": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "grt_graph_launch": (ctypes.c_int, [_P]),
    "grt_graph_destroy": (ctypes.c_int, [_P]),
    "grt_stream_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "grt_stream_destroy": (ctypes.c_int, [_P]),
    # scene generators (host only)
    "grt_gen_spheres": (_U32, [ctypes.c_uint64, _U32, _F, _F, _P, _P]),
    "grt_gen_sphere_clusters": (_U32, [ctypes.c_uint64, _U32, _U32, _F, _F, _F, _P, _P]),
    "grt_gen_stadium": (_U32, [ctypes.c_uint64, _U32, _U32, _U32, _P, _P, _P]),
    "grt_gen_cluster_transforms": (None, [ctypes.c_uint64, _U32, _U32, _P, _F, _P]),
    "grt_gen_uniform": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64]),
    # the reference tracer's flat C ABI (include/grt/rt_cabi.h)
    "rtInitContext": (ctypes.c_bool, []),
    "rtShutdownContext": (ctypes.c_int, []),
    "grtLastError": (ctypes.c_char_p, []),
    "grtRegisterBuffer": (ctypes.c_int, [_P]),
    "grtUnregisterBuffer": (ctypes.c_int, [_P]),
    "grtSetScreenSize": (ctypes.c_int, [_U32, _U32]),
    "grtSetCameraParameters": (ctypes.c_int, [_F] * 12),
    "grtTransferTriangleVertices": (ctypes.c_int, [_P, _U32]),
    "grtTransferTriangleNormals": (ctypes.c_int, [_P, _U32]),
    "grtTransferCellPointers": (ctypes.c_int, [_P, _U32]),
    "grtTransferCellTriangleIds": (ctypes.c_int, [_P, _U32]),
    "grtSetBoxMin": (ctypes.c_int, [_F, _F, _F]),
    "grtSetBoxMax": (ctypes.c_int, [_F, _F, _F]),
    "grtSetGridParameters": (ctypes.c_int, [_F, _F, _F, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32]),
    "grtSetBlockSize": (ctypes.c_int, [_U32, _U32]),
    "grtRayTrace": (ctypes.c_int, []),
    "grtGetHitRecords": (ctypes.c_int, [_P, _P, _P, _P]),
}


def lib():
    """Load libgrt.so (built in-tree by ``make -C grid-rt_amd``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GrtError(GRT_NOT_READY, f"{LIB_PATH} is missing: build it with "
                           "`make -C grid-rt_amd` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != GRT_OK:
        raise GrtError(rc, lib().grt_last_error().decode(errors="replace"))


def _call(name: str, *args) -> int:
    """Call a libgrt entry point with its declared signature."""
    fn = getattr(lib(), name)
    fn.restype, fn.argtypes = _SIGS[name]
    return fn(*args)


def _ptr(a) -> int:
    """Data pointer of a numpy array or torch tensor (0 for None)."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


# ---------------------------------------------------------------------------
# Scenes (host generators)
# ---------------------------------------------------------------------------
def gen_spheres(seed: int, n_spheres: int, rmin: float = 0.02, rmax: float = 0.08):
    L = lib()
    n = L.grt_gen_spheres(seed, n_spheres, rmin, rmax, None, None)
    v = np.empty((n, 9), np.float32)
    nn = np.empty((n, 9), np.float32)
    L.grt_gen_spheres(seed, n_spheres, rmin, rmax, _ptr(v), _ptr(nn))
    return v, nn


def gen_sphere_clusters(seed: int, n_clusters: int, per_cluster: int, sigma: float,
                        rmin: float, rmax: float):
    L = lib()
    n = L.grt_gen_sphere_clusters(seed, n_clusters, per_cluster, sigma, rmin, rmax, None, None)
    v = np.empty((n, 9), np.float32)
    nn = np.empty((n, 9), np.float32)
    L.grt_gen_sphere_clusters(seed, n_clusters, per_cluster, sigma, rmin, rmax, _ptr(v), _ptr(nn))
    return v, nn


def gen_stadium(seed: int, n_clusters: int, tris_per_cluster: int, n_cluster_tris: int):
    L = lib()
    n = L.grt_gen_stadium(seed, n_clusters, tris_per_cluster, n_cluster_tris, None, None, None)
    if n == 0:
        raise GrtError(GRT_INVALID_PARAM, "gen_stadium: bad cluster sizes")
    v = np.empty((n, 9), np.float32)
    nn = np.empty((n, 9), np.float32)
    c = np.empty((n_clusters, 3), np.float32)
    L.grt_gen_stadium(seed, n_clusters, tris_per_cluster, n_cluster_tris, _ptr(v), _ptr(nn), _ptr(c))
    return v, nn, c


def gen_cluster_transforms(seed: int, frame: int, centers: np.ndarray, max_translation: float):
    centers = np.ascontiguousarray(centers, np.float32)
    xf = np.empty((centers.shape[0], 12), np.float32)
    lib().grt_gen_cluster_transforms(seed, frame, centers.shape[0], _ptr(centers),
                                     max_translation, _ptr(xf))
    return xf


def camera_lookat(eye, center, up, fovy_deg: float, width: int, height: int,
                  znear: float = 1.0) -> Camera:
    cam = Camera()
    f3 = ctypes.c_float * 3
    check(lib().grt_camera_lookat(f3(*eye), f3(*center), f3(*up), fovy_deg, znear, width, height,
                                  ctypes.byref(cam)))
    return cam


def camera_dict(cam: Camera) -> dict:
    return {"pos": list(cam.pos), "base": list(cam.base), "nu": list(cam.nu), "nv": list(cam.nv),
            "width": cam.width, "height": cam.height}


# ---------------------------------------------------------------------------
# Device batches
# ---------------------------------------------------------------------------
def _torch():
    import torch  # noqa: deferred so CPU-only users never need a GPU
    return torch


class RayBatch:
    """Structure-of-arrays ray buffer on the device (8 float32 tensors)."""

    def __init__(self, fields: dict):
        self.f = fields
        self.n = int(fields["ox"].numel())

    @classmethod
    def empty(cls, n: int, device="cuda:0"):
        t = _torch()
        return cls({k: t.empty(n, dtype=t.float32, device=device) for k in RAY_FIELDS})

    @classmethod
    def from_numpy(cls, d: dict, device="cuda:0"):
        t = _torch()
        return cls({k: t.from_numpy(np.ascontiguousarray(d[k], np.float32)).to(device)
                    for k in RAY_FIELDS})

    def numpy(self) -> dict:
        return {k: v.cpu().numpy() for k, v in self.f.items()}

    def c(self, offset: int = 0) -> Rays:
        r = Rays()
        for k in RAY_FIELDS:
            setattr(r, k, self.f[k].data_ptr() + 4 * offset)
        return r


class HitBatch:
    def __init__(self, fields: dict):
        self.f = fields
        self.n = int(fields["id"].numel())

    @classmethod
    def empty(cls, n: int, device="cuda:0"):
        t = _torch()
        return cls({"id": t.empty(n, dtype=t.int32, device=device),
                    "t": t.empty(n, dtype=t.float32, device=device),
                    "u": t.empty(n, dtype=t.float32, device=device),
                    "v": t.empty(n, dtype=t.float32, device=device)})

    @classmethod
    def from_numpy(cls, d: dict, device="cuda:0"):
        t = _torch()
        return cls({"id": t.from_numpy(np.ascontiguousarray(d["id"], np.int32)).to(device),
                    **{k: t.from_numpy(np.ascontiguousarray(d[k], np.float32)).to(device)
                       for k in ("t", "u", "v")}})

    def numpy(self) -> dict:
        return {k: v.cpu().numpy() for k, v in self.f.items()}

    def c(self, offset: int = 0) -> Hits:
        h = Hits()
        for k in HIT_FIELDS:
            setattr(h, k, self.f[k].data_ptr() + 4 * offset)
        return h


# ---------------------------------------------------------------------------
# Context / scene handles
# ---------------------------------------------------------------------------
def phong(params: dict | None = None) -> Phong:
    """The shader's defaults (grt_phong_defaults), with `params` overriding
    fields by name."""
    p = Phong()
    lib().grt_phong_defaults(ctypes.byref(p))
    for k, v in (params or {}).items():
        if k in ("light", "light_color", "ambient", "specular", "background"):
            getattr(p, k)[:] = [float(x) for x in v]
        elif k == "spec_exp":
            p.spec_exp = int(v)
        elif k in ("shadow_coeff", "reflect_coeff"):
            setattr(p, k, float(v))
        else:
            raise KeyError(k)
    return p


class Context:
    """One device context; by default enqueues on torch's current stream."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        L = lib()
        h = _P()
        check(L.grt_init(device, ctypes.byref(h)))
        self.h = h
        self.device = device
        if use_torch_stream:
            t = _torch()
            check(L.grt_set_stream(self.h, t.cuda.current_stream(device).cuda_stream))

    def set_stream(self, stream_handle: int):
        check(lib().grt_set_stream(self.h, stream_handle))

    def synchronize(self):
        check(lib().grt_synchronize(self.h))

    def memory_used(self) -> int:
        v = ctypes.c_uint64()
        check(lib().grt_memory_used(self.h, ctypes.byref(v)))
        return v.value

    def set_block_size(self, closest: int, any_hit: int):
        check(lib().grt_set_block_size(self.h, closest, any_hit))

    TRACE_MODES = {"simple": 0, "persistent": 1, "pooled": 2, "split": 3}

    def set_trace_mode(self, mode="simple", param: int = 16):
        """Traversal schedule: "simple" (one ray per lane), "pooled" (one ray
        per lane, the wave's triangle tests pooled), "persistent" (refilled
        persistent waves, one test or cell per iteration; param = refill
        threshold 1..64), "split" (param = 2, 4, 8 or 16 lanes per ray, each
        lane visiting every param-th cell of the ray); a bool selects
        persistent (True) or simple."""
        if isinstance(mode, bool):
            mode = "persistent" if mode else "simple"
        check(lib().grt_set_trace_mode(self.h, self.TRACE_MODES[mode], param))

    def set_coop(self, min_refs: int = 16):
        """Cooperative cell tests in drained persistent waves (0: off)."""
        check(lib().grt_set_coop(self.h, min_refs))

    def trace_instances(self, instances, rays: RayBatch, hits: HitBatch, instance_out,
                        image_width: int = 0):
        """Closest hit over [(Scene, 3x4 object-to-world matrix), ...];
        instance_out (int32 per ray) receives the instance index or -1."""
        arr = (Instance * max(len(instances), 1))()
        for k, (sc, xf) in enumerate(instances):
            arr[k].scene = sc.h
            arr[k].xf[:] = [float(x) for x in np.asarray(xf, np.float32).reshape(12)]
        check(lib().grt_trace_instances(self.h, arr, len(instances), ctypes.byref(rays.c()),
                                        ctypes.byref(hits.c()), _ptr(instance_out), rays.n,
                                        image_width))

    def compact_rays(self, rays: RayBatch, image_width: int, ray_list, count, occluded=None,
                     miss_hits: "HitBatch | None" = None):
        """List the active rays of a batch (grt_compact_rays): ray_list a device
        int32 tensor of rays.n entries, count a one-element device tensor."""
        check(lib().grt_compact_rays(self.h, ctypes.byref(rays.c()), rays.n, image_width,
                                     _ptr(ray_list), _ptr(count),
                                     _ptr(occluded) if occluded is not None else None,
                                     ctypes.byref(miss_hits.c()) if miss_hits is not None
                                     else None))

    def compact_rays_noalloc(self, rays: RayBatch, image_width: int, ray_list, count,
                             occluded=None, miss_hits: "HitBatch | None" = None,
                             n: int | None = None):
        """compact_rays without device allocation (grt_compact_rays_noalloc):
        the tile sums live in a fixed per-GPU array, so at most 67,100,672 ray
        slots per call."""
        check(lib().grt_compact_rays_noalloc(self.h, ctypes.byref(rays.c()),
                                             rays.n if n is None else n, image_width,
                                             _ptr(ray_list), _ptr(count),
                                             _ptr(occluded) if occluded is not None else None,
                                             ctypes.byref(miss_hits.c()) if miss_hits is not None
                                             else None))

    def gen_primary(self, cam: Camera, out: RayBatch):
        check(lib().grt_gen_primary(self.h, ctypes.byref(cam), ctypes.byref(out.c())))

    def debug_scan(self, x, out):
        check(lib().grt_debug_scan(self.h, _ptr(x), _ptr(out), int(x.numel())))

    def close(self):
        if getattr(self, "h", None):
            lib().grt_shutdown(self.h)
            self.h = None


TRIANGLES, TRIANGLE_STRIP, TRIANGLE_FAN, POLYGON = 0x1001, 0x1002, 0x1003, 0x1004
MESH_DROP_DEGENERATE = 1


class Mesh:
    """Host-side triangle soup (grt_mesh_*): immediate-mode primitives with the
    reference's assembly and degenerate filter, OBJ/MTL loading, procedural
    sphere and cube.  Needs no GPU."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().grt_mesh_create(ctypes.byref(h)))
        self.h = h

    def set_transform(self, xf=None):
        a = None if xf is None else np.ascontiguousarray(xf, np.float32).reshape(12)
        check(lib().grt_mesh_set_transform(self.h, _ptr(a) if a is not None else None))
        self._xf = a   # keep alive

    def set_color(self, r, g, b):
        check(lib().grt_mesh_set_color(self.h, r, g, b))

    def begin(self, primitive):
        check(lib().grt_mesh_begin(self.h, primitive))

    def normal(self, x, y, z):
        check(lib().grt_mesh_normal(self.h, x, y, z))

    def vertex(self, x, y, z):
        check(lib().grt_mesh_vertex(self.h, x, y, z))

    def end(self):
        check(lib().grt_mesh_end(self.h))

    def primitive(self, primitive, vertices, normals=None):
        """One primitive from arrays of vertices (k x 3) and optional normals."""
        self.begin(primitive)
        for k, p in enumerate(np.asarray(vertices, np.float32).reshape(-1, 3)):
            if normals is not None:
                self.normal(*map(float, np.asarray(normals, np.float32).reshape(-1, 3)[k]))
            self.vertex(*map(float, p))
        self.end()

    def load_obj(self, path, drop_degenerate: bool = False):
        check(lib().grt_mesh_load_obj(self.h, str(path).encode(),
                                      MESH_DROP_DEGENERATE if drop_degenerate else 0))

    def add_sphere(self, slices: int, stacks: int):
        check(lib().grt_mesh_add_sphere(self.h, slices, stacks))

    def add_cube(self):
        check(lib().grt_mesh_add_cube(self.h))

    # rtdb's layouts (rtdb.cpp:9-57,215-287); `model` stands for the teapot.
    def add_triangle(self):
        check(lib().grt_mesh_add_triangle(self.h))

    def add_logo(self):
        check(lib().grt_mesh_add_logo(self.h))

    def add_mesh(self, model: "Mesh"):
        check(lib().grt_mesh_add_mesh(self.h, model.h))

    def add_even_spaced(self):
        check(lib().grt_mesh_add_even_spaced(self.h))

    def add_mixed(self, model: "Mesh"):
        check(lib().grt_mesh_add_mixed(self.h, model.h))

    def add_teapots125(self, model: "Mesh"):
        check(lib().grt_mesh_add_teapots125(self.h, model.h))

    def info(self):
        n, d = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().grt_mesh_info(self.h, ctypes.byref(n), ctypes.byref(d)))
        return n.value, d.value

    def arrays(self):
        """(vertices [n, 9], normals [n, 9], diffuse [n, 3], source index [n])."""
        n, _ = self.info()
        v = np.empty((n, 9), np.float32)
        nr = np.empty((n, 9), np.float32)
        kd = np.empty((n, 3), np.float32)
        src = np.empty(n, np.uint32)
        check(lib().grt_mesh_copy(self.h, _ptr(v), _ptr(nr), _ptr(kd), _ptr(src)))
        return v, nr, kd, src

    def close(self):
        if getattr(self, "h", None):
            lib().grt_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = _P()
        check(lib().grt_scene_create(ctx.h, ctypes.byref(h)))
        self.h = h
        self.n_tris = 0

    def upload(self, verts: np.ndarray, normals: np.ndarray | None = None):
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
        self.n_tris = verts.shape[0]
        check(lib().grt_scene_upload_triangles(self.h, _ptr(verts), self.n_tris))
        if normals is not None:
            normals = np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
            check(lib().grt_scene_upload_normals(self.h, _ptr(normals), normals.shape[0]))

    def upload_mesh(self, mesh: "Mesh"):
        check(lib().grt_scene_upload_mesh(self.h, mesh.h))
        self.n_tris = mesh.info()[0]

    def transform_clusters(self, xf: np.ndarray, tris_per_cluster: int, n_cluster_tris: int):
        xf = np.ascontiguousarray(xf, np.float32).reshape(-1, 12)
        check(lib().grt_scene_transform_clusters(self.h, _ptr(xf), xf.shape[0], tris_per_cluster,
                                                 n_cluster_tris))

    def download_geometry(self):
        v = np.empty((self.n_tris, 9), np.float32)
        n = np.empty((self.n_tris, 9), np.float32)
        check(lib().grt_scene_download_geometry(self.h, _ptr(v), _ptr(n)))
        return v, n

    def device_geometry(self):
        """(verts, normals) device pointers of the current pose."""
        v, n = _P(), _P()
        check(lib().grt_scene_device_pointers(self.h, ctypes.byref(v), ctypes.byref(n)))
        return v.value, n.value

    def build_grid(self, lam: float = 5.0):
        check(lib().grt_grid_build(self.h, lam))

    def rebuild_grid_async(self, lam: float = 5.0):
        check(lib().grt_grid_rebuild_async(self.h, lam))

    def reserve_grid(self, cells: int, refs: int):
        check(lib().grt_grid_reserve(self.h, cells, refs))

    def check_grid(self):
        check(lib().grt_grid_check(self.h))

    def set_subgrid(self, lambda2: float, min_refs: int = 8):
        """Second grid level from the next build on (lambda2 = 0: off)."""
        check(lib().grt_grid_set_subgrid(self.h, lambda2, min_refs))

    def subgrid_info(self) -> dict:
        i = SubgridInfo()
        check(lib().grt_subgrid_info_get(self.h, ctypes.byref(i)))
        return {"lambda2": float(i.lambda2), "min_refs": int(i.min_refs), "built": bool(i.built),
                "cells": int(i.cells), "refs": int(i.refs), "status": int(i.status)}

    def download_subgrid(self):
        """(cellsub [m, 2], sub_offsets, sub_refs) of the second level."""
        si = self.subgrid_info()
        m = self.grid_info()["m"]
        cellsub = np.empty((m, 2), np.uint32)
        offs = np.empty(si["cells"] + 1, np.uint32)
        refs = np.empty(max(si["refs"], 1), np.uint32)
        check(lib().grt_subgrid_download(self.h, _ptr(cellsub), _ptr(offs), _ptr(refs)))
        return cellsub, offs, refs[:si["refs"]]

    def upload_grid(self, frame: dict, offsets: np.ndarray, refs: np.ndarray):
        gi = GridInfo()
        for k in ("bmin", "bmax", "cell", "inv_cell"):
            getattr(gi, k)[:] = [float(x) for x in frame[k]]
        gi.n[:] = [int(x) for x in frame["n"]]
        offsets = np.ascontiguousarray(offsets, np.uint32)
        refs = np.ascontiguousarray(refs, np.uint32)
        check(lib().grt_grid_upload(self.h, ctypes.byref(gi), _ptr(offsets), _ptr(refs)))

    def grid_info(self) -> dict:
        gi = GridInfo()
        check(lib().grt_grid_info_get(self.h, ctypes.byref(gi)))
        return {"bmin": np.array(gi.bmin[:], np.float32), "bmax": np.array(gi.bmax[:], np.float32),
                "cell": np.array(gi.cell[:], np.float32),
                "inv_cell": np.array(gi.inv_cell[:], np.float32),
                "n": np.array(gi.n[:], np.int32), "m": int(gi.m), "r": int(gi.r)}

    def download_grid(self):
        info = self.grid_info()
        offsets = np.empty(info["m"] + 1, np.uint32)
        refs = np.empty(max(info["r"], 1), np.uint32)
        check(lib().grt_grid_download(self.h, _ptr(offsets), _ptr(refs)))
        return info, offsets, refs[:info["r"]]

    def trace_closest(self, rays: RayBatch, hits: HitBatch, n: int | None = None,
                      image_width: int = 0, counters=None, offset: int = 0):
        n = rays.n if n is None else n
        check(lib().grt_trace_closest(self.h, ctypes.byref(rays.c(offset)),
                                      ctypes.byref(hits.c(offset)), n, image_width,
                                      _ptr(counters)))

    def trace_any(self, rays: RayBatch, occluded, n: int | None = None, image_width: int = 0,
                  counters=None):
        n = rays.n if n is None else n
        check(lib().grt_trace_any(self.h, ctypes.byref(rays.c()), _ptr(occluded), n, image_width,
                                  _ptr(counters)))

    def gen_shadow(self, rays: RayBatch, hits: HitBatch, light, offset: float, out: RayBatch):
        f3 = ctypes.c_float * 3
        check(lib().grt_gen_shadow(self.h, ctypes.byref(rays.c()), ctypes.byref(hits.c()),
                                   rays.n, f3(*light), offset, ctypes.byref(out.c())))

    def gen_reflection(self, rays: RayBatch, hits: HitBatch, offset: float, out: RayBatch):
        check(lib().grt_gen_reflection(self.h, ctypes.byref(rays.c()), ctypes.byref(hits.c()),
                                       rays.n, offset, ctypes.byref(out.c())))

    # Ray lists (grt_compact_rays / grt_trace_*_list): `ray_list` is a device
    # int32/uint32 tensor of n entries, `count` a one-element device tensor.
    def trace_closest_list(self, rays: RayBatch, hits: HitBatch, ray_list, count, counters=None):
        check(lib().grt_trace_closest_list(self.h, ctypes.byref(rays.c()), ctypes.byref(hits.c()),
                                           _ptr(ray_list), _ptr(count), rays.n, _ptr(counters)))

    def trace_any_list(self, rays: RayBatch, occluded, ray_list, count, counters=None):
        check(lib().grt_trace_any_list(self.h, ctypes.byref(rays.c()), _ptr(occluded),
                                       _ptr(ray_list), _ptr(count), rays.n, _ptr(counters)))

    def gen_ao(self, rays: RayBatch, hits: HitBatch, spp: int, seed: int, offset: float,
               ao_tmax: float, out: RayBatch):
        check(lib().grt_gen_ao(self.h, ctypes.byref(rays.c()), ctypes.byref(hits.c()), rays.n,
                               spp, seed, offset, ao_tmax, ctypes.byref(out.c())))

    def shade_headlight(self, rays: RayBatch, hits: HitBatch, rgba):
        check(lib().grt_shade_headlight(self.h, ctypes.byref(rays.c()), ctypes.byref(hits.c()),
                                        rays.n, _ptr(rgba)))

    def shade_phong(self, rays: RayBatch, hits: HitBatch, rgb, occluded=None, depth: int = 0,
                    params: "Phong | dict | None" = None, diffuse=None, rgba=None):
        """Phong shading into rgb (device float32, 3 per ray): depth 0 writes,
        depth d > 0 adds reflect_coeff^d x colour for hits.  occluded: the
        hits' shadow-ray flags; diffuse: 3 floats per triangle; rgba: optional
        RGBA8 output (int32/uint32 per ray)."""
        p = params if isinstance(params, Phong) else phong(params)
        check(lib().grt_shade_phong(self.h, ctypes.byref(rays.c()), ctypes.byref(hits.c()),
                                    _ptr(occluded) if occluded is not None else None, rays.n,
                                    _ptr(diffuse) if diffuse is not None else None,
                                    ctypes.byref(p), depth, _ptr(rgb),
                                    _ptr(rgba) if rgba is not None else None))

    def close(self):
        if getattr(self, "h", None):
            lib().grt_scene_destroy(self.h)
            self.h = None


class KdTree:
    """The kd-tree comparison accelerator (grt_kdtree_*): a surface-area-
    heuristic kd-tree over a scene's current triangles, built on the host and
    traced on the device."""

    def __init__(self, scene: Scene):
        self.scene = scene
        h = _P()
        check(lib().grt_kdtree_build(scene.h, ctypes.byref(h)))
        self.h = h

    @classmethod
    def from_host(cls, verts: np.ndarray) -> "KdTree":
        """The same tree built from host vertices without a device (no traces)."""
        verts = np.ascontiguousarray(verts, np.float32).reshape(-1, 9)
        self = cls.__new__(cls)
        self.scene = None
        h = _P()
        check(lib().grt_kdtree_build_host(_ptr(verts), verts.shape[0], ctypes.byref(h)))
        self.h = h
        return self

    def info(self) -> dict:
        i = KdTreeInfo()
        check(lib().grt_kdtree_info_get(self.h, ctypes.byref(i)))
        return {"bmin": np.array(i.bmin[:], np.float32), "bmax": np.array(i.bmax[:], np.float32),
                "nodes": int(i.nodes), "leaves": int(i.leaves),
                "empty_leaves": int(i.empty_leaves), "elements": int(i.elements),
                "depth": int(i.depth), "build_ms": float(i.build_ms)}

    def download(self):
        """(nodes [n, 2] uint32, elements [m] uint32) as built."""
        inf = self.info()
        nodes = np.empty((inf["nodes"], 2), np.uint32)
        elements = np.empty(max(inf["elements"], 1), np.uint32)
        check(lib().grt_kdtree_download(self.h, _ptr(nodes), _ptr(elements)))
        return nodes, elements[:inf["elements"]]

    def trace_closest(self, rays: RayBatch, hits: HitBatch, image_width: int = 0,
                      counters=None):
        check(lib().grt_kdtree_trace_closest(self.h, ctypes.byref(rays.c()),
                                             ctypes.byref(hits.c()), rays.n, image_width,
                                             _ptr(counters)))

    def trace_any(self, rays: RayBatch, occluded, image_width: int = 0, counters=None):
        check(lib().grt_kdtree_trace_any(self.h, ctypes.byref(rays.c()), _ptr(occluded), rays.n,
                                         image_width, _ptr(counters)))

    def close(self):
        if getattr(self, "h", None):
            lib().grt_kdtree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Batched multi-scene tracing and launch graphs
# ---------------------------------------------------------------------------
class BatchItem(ctypes.Structure):
    _fields_ = [("scene", _P), ("rays", Rays), ("hits", Hits), ("occluded", _P),
                ("n", _U32), ("image_width", _U32), ("any", _U32)]


class Batch:
    """Many (scene, ray batch) items traced in one persistent launch."""

    def __init__(self, ctx: Context, max_items: int):
        self.ctx = ctx
        h = _P()
        check(lib().grt_batch_create(ctx.h, max_items, ctypes.byref(h)))
        self.h = h
        self._keep = None

    def set(self, items):
        """items: iterable of (scene, rays, hits_or_None, occluded_or_None, image_width)."""
        arr = (BatchItem * max(1, len(items)))()
        for i, (scene, rays, hits, occ, width) in enumerate(items):
            it = arr[i]
            it.scene = scene.h
            it.rays = rays.c()
            if hits is not None:
                it.hits = hits.c()
                it.any = 0
            else:
                it.occluded = _ptr(occ)
                it.any = 1
            it.n = rays.n
            it.image_width = width
        self._keep = (arr, items)
        check(lib().grt_batch_set(self.h, ctypes.cast(arr, _P), len(items)))

    def slots(self) -> int:
        v = _U32()
        check(lib().grt_batch_slots(self.h, ctypes.byref(v)))
        return v.value

    def trace(self):
        check(lib().grt_batch_trace(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().grt_batch_destroy(self.h)
            self.h = None


class Graph:
    """hipGraph of the launches enqueued on a context's stream while capturing."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.h = None

    def begin(self):
        check(lib().grt_graph_begin(self.ctx.h))

    def end(self):
        h = _P()
        check(lib().grt_graph_end(self.ctx.h, ctypes.byref(h)))
        self.h = h

    def launch(self):
        check(lib().grt_graph_launch(self.h))

    def close(self):
        if self.h:
            lib().grt_graph_destroy(self.h)
            self.h = None


def stream_create() -> int:
    h = _P()
    check(lib().grt_stream_create(ctypes.byref(h)))
    return h.value


def stream_destroy(handle: int):
    check(lib().grt_stream_destroy(handle))


# ---------------------------------------------------------------------------
# Serving
# ---------------------------------------------------------------------------
class Request(ctypes.Structure):
    _fields_ = [("scene", _P), ("camera", Camera), ("depth", _U32), ("arrival_s", ctypes.c_double),
                ("out_hits", Hits), ("out_occluded", _P)]


class RequestResult(ctypes.Structure):
    _fields_ = [("start_s", ctypes.c_double), ("finish_s", ctypes.c_double),
                ("rays", ctypes.c_uint64)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("max_active", _U32), ("batched", _U32),
                ("light", _F * 3), ("offset", _F)]


_SIGS.update({
    "grt_serve_run": (ctypes.c_int, [_P, _P, _U32, ctypes.POINTER(ServeConfig), _P,
                                     ctypes.POINTER(ctypes.c_double)]),
})


def serve_run(ctx: Context, requests, batched: bool, memory_budget: int, max_active: int = 64,
              light=(1000.0, 1000.0, 1000.0), offset: float = 1e-4):
    """requests: list of (scene, camera, depth, arrival_s), sorted by arrival.
    Returns (arrival, start, finish, rays) numpy arrays and the wall time."""
    L = lib()
    fn = L.grt_serve_run
    fn.restype, fn.argtypes = _SIGS["grt_serve_run"]
    n = len(requests)
    arr = (Request * max(1, n))()
    for i, req in enumerate(requests):
        scene, cam, depth, t = req[:4]
        arr[i].scene = scene.h
        arr[i].camera = cam
        arr[i].depth = depth
        arr[i].arrival_s = t
        if len(req) > 4 and req[4] is not None:   # (HitBatch, occluded tensor) outputs
            hb, occ = req[4]
            arr[i].out_hits = hb.c()
            arr[i].out_occluded = _ptr(occ)
    res = (RequestResult * max(1, n))()
    cfg = ServeConfig()
    cfg.memory_budget = memory_budget
    cfg.max_active = max_active
    cfg.batched = 1 if batched else 0
    cfg.light[:] = list(light)
    cfg.offset = offset
    wall = ctypes.c_double()
    check(fn(ctx.h, ctypes.cast(arr, _P), n, ctypes.byref(cfg), ctypes.cast(res, _P),
             ctypes.byref(wall)))
    arrival = np.array([r[3] for r in requests], np.float64)
    start = np.array([res[i].start_s for i in range(n)], np.float64)
    finish = np.array([res[i].finish_s for i in range(n)], np.float64)
    rays = np.array([res[i].rays for i in range(n)], np.uint64)
    return arrival, start, finish, rays, wall.value


_SIGS.update({
    "grt_server_create": (ctypes.c_int, [_P, ctypes.POINTER(ServeConfig), ctypes.POINTER(_P)]),
    "grt_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Request),
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "grt_server_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(RequestResult)]),
    "grt_server_query": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(RequestResult)]),
    "grt_server_destroy": (ctypes.c_int, [_P]),
})


class Server:
    """Requests submitted as they come, from any thread (grt_server_*): a
    thread of the library owns the device loop on the context's stream, which
    is reserved to the server until close()."""

    def __init__(self, ctx: Context, memory_budget: int, max_active: int = 64,
                 batched: bool = True, light=(1000.0, 1000.0, 1000.0), offset: float = 1e-4):
        self.ctx = ctx
        cfg = ServeConfig()
        cfg.memory_budget = memory_budget
        cfg.max_active = max_active
        cfg.batched = 1 if batched else 0
        cfg.light[:] = list(light)
        cfg.offset = offset
        h = _P()
        check(_call("grt_server_create", ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}   # ticket -> output buffers, alive until the result is taken

    def submit(self, scene, camera, depth: int, out_hits=None, out_occluded=None) -> int:
        """Queue a request; returns its ticket at once.  out_hits (HitBatch) and
        out_occluded (uint8 tensor) receive the primary hits and shadow flags."""
        r = Request()
        r.scene = scene.h
        r.camera = camera
        r.depth = depth
        r.arrival_s = 0.0
        if out_hits is not None:
            r.out_hits = out_hits.c()
        if out_occluded is not None:
            r.out_occluded = _ptr(out_occluded)
        t = ctypes.c_uint64()
        check(_call("grt_server_submit", self.h, ctypes.byref(r), ctypes.byref(t)))
        self._keep[t.value] = (scene, out_hits, out_occluded)
        return t.value

    @staticmethod
    def _res(r):
        return {"start_s": r.start_s, "finish_s": r.finish_s, "rays": int(r.rays)}

    def wait(self, ticket: int) -> dict:
        r = RequestResult()
        try:
            check(_call("grt_server_wait", self.h, ticket, ctypes.byref(r)))
        finally:
            self._keep.pop(ticket, None)
        return self._res(r)

    def query(self, ticket: int):
        """The result, or None while the request runs."""
        r = RequestResult()
        rc = _call("grt_server_query", self.h, ticket, ctypes.byref(r))
        if rc == GRT_NOT_READY:
            return None
        self._keep.pop(ticket, None)
        check(rc)
        return self._res(r)

    def close(self):
        if getattr(self, "h", None):
            h, self.h = self.h, None
            check(_call("grt_server_destroy", h))
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
