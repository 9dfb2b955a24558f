"""ctypes view of the C ABI declared in include/hptracing.h.

Two in-tree shared libraries are loaded from ``hptracing_amd/lib``:
``libhpt_host.so`` (CPU: OBJ/MTL, textures, light table, KD build) and
``libhptracing.so`` (GPU: HIP kernels for gfx950).  Nothing here falls back
to another implementation: a missing library raises ``NativeLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HOST_LIB = os.path.join(LIB_DIR, "libhpt_host.so")
DEVICE_LIB = os.path.join(LIB_DIR, "libhptracing.so")

HPT_OK = 0
HPT_ERR_INVALID = -1
HPT_ERR_IO = -2
HPT_ERR_SCENE = -3
HPT_ERR_HIP = -4
HPT_ERR_NOMEM = -5
HPT_ERR_BUDGET = -6
HPT_ERR_STATE = -7
HPT_ERR_TIMEOUT = -8

HPT_TEXTURE_SIZE = 512
HPT_LIGHT_TABLE_SIZE = 256
HPT_MAX_DEPTH = 64


class NativeLibraryMissing(RuntimeError):
    """The compiled library is absent: run __graft_entry__.build()."""


class HptError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"hpt error {code}: {message}")
        self.code = code


class Material(C.Structure):
    _fields_ = [
        ("emission", C.c_float * 4),
        ("ambient", C.c_float * 4),
        ("diffuse", C.c_float * 4),
        ("specular", C.c_float * 4),
        ("optical_density", C.c_float),
        ("dissolve", C.c_float),
        ("shininess", C.c_float),
        ("diffuse_texture_id", C.c_int32),
        ("ambient_texture_id", C.c_int32),
        ("specular_possibility", C.c_float),
        ("refract_possibility", C.c_float),
        ("diffuse_possibility", C.c_float),
    ]


class KDNode(C.Structure):
    _fields_ = [
        ("box_start", C.c_float * 4),
        ("box_end", C.c_float * 4),
        ("child", C.c_int32),
        ("parent", C.c_int32),
        ("sibling", C.c_int32),
        ("data", C.c_int32),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("points", C.c_void_p), ("n_points", C.c_int64),
        ("normals", C.c_void_p), ("n_normals", C.c_int64),
        ("texcoords", C.c_void_p), ("n_texcoords", C.c_int64),
        ("mesh", C.c_void_p), ("n_triangles", C.c_int64),
        ("materials", C.c_void_p), ("n_materials", C.c_int64),
        ("lights", C.c_void_p), ("n_lights", C.c_int64),
        ("nodes", C.c_void_p), ("n_nodes", C.c_int64),
        ("leaf_data", C.c_void_p), ("n_leaf_data", C.c_int64),
        ("textures", C.c_void_p), ("n_textures", C.c_int64),
    ]


class KdCounts(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("n_leaves", C.c_int64), ("n_leaf_entries", C.c_int64),
        ("n_leaf_ints", C.c_int64), ("max_leaf", C.c_int64), ("max_depth", C.c_int32),
        ("reserved", C.c_int32), ("seconds", C.c_double),
    ]


class SceneStats(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("n_leaf_entries", C.c_int64),
        ("n_triangles", C.c_int64), ("n_textures", C.c_int64),
        ("device_bytes", C.c_int64),
    ]


class View(C.Structure):
    _fields_ = [
        ("view", C.c_float * 3),
        ("up", C.c_float * 3),
        ("right", C.c_float * 3),
        ("width", C.c_float),
        ("height", C.c_float),
        ("sample_x", C.c_int32), ("sample_y", C.c_int32),
        ("supersample_x", C.c_int32), ("supersample_y", C.c_int32),
        ("background", C.c_float * 3),
        ("sample", C.c_int32),
        ("depth", C.c_int32),
        ("disable_diffuse", C.c_int32),
        ("brightness", C.c_float),
        ("seed", C.c_uint64),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("primary_rays", C.c_int64),
        ("rays_cast", C.c_int64),
        ("live_per_depth", C.c_int64 * HPT_MAX_DEPTH),
        ("seconds", C.c_double),
        ("cast_ms", C.c_double), ("shade_ms", C.c_double),
        ("cast_launches", C.c_int64), ("shade_launches", C.c_int64),
        ("node_visits", C.c_int64), ("leaf_visits", C.c_int64), ("tri_tests", C.c_int64),
    ]


HPT_RENDER_TIME_KERNELS = 1
HPT_RENDER_COUNT_WORK = 2
HPT_RENDER_NO_GRAPH = 4
HPT_RENDER_PASS_AT_A_TIME = 8
HPT_CAST_COUNTS = 5
HPT_NOALLOC_MAX_RAYS = 1 << 20
HPT_SERVE_PROGRESSIVE = 1


class Tuning(C.Structure):
    _fields_ = [("fast_division", C.c_int32), ("cast_blocks_per_cu", C.c_int32),
                ("shade_blocks_per_cu", C.c_int32), ("traversal", C.c_int32),
                ("refill_threshold", C.c_int32), ("node_steps", C.c_int32),
                ("fetch", C.c_int32), ("walk_min", C.c_int32), ("sort_rays", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("device_budget_bytes", C.c_int64)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("queue_ms", C.c_double),
                ("latency_ms", C.c_double), ("camera_rays", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int64), ("active", C.c_int64), ("completed", C.c_int64),
                ("macro_passes", C.c_int64), ("rays_cast", C.c_int64),
                ("busy_seconds", C.c_double), ("capacity_rays", C.c_int64)]


_host = None
_device = None


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    return C.CDLL(path)


def host() -> C.CDLL:
    """libhpt_host.so with argument types declared."""
    global _host
    if _host is None:
        lib = _load(HOST_LIB)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        lib.hpt_host_last_error.restype = C.c_char_p
        lib.hpt_host_load_obj.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(vp)]
        lib.hpt_host_from_arrays.argtypes = [vp, i64, vp, i64, vp, i64, C.c_int, C.POINTER(vp)]
        lib.hpt_host_describe.argtypes = [vp, C.POINTER(SceneDesc)]
        lib.hpt_host_kd_count.argtypes = [vp, i64, vp, i64, C.c_int, C.POINTER(KdCounts)]
        lib.hpt_host_read_image.argtypes = [C.c_char_p, C.POINTER(i32), C.POINTER(i32),
                                            C.POINTER(i32), vp, i64]
        lib.hpt_host_material_from_mtl.argtypes = [vp, vp, vp, vp, C.c_float, C.c_float,
                                                   C.c_float, C.POINTER(Material)]
        lib.hpt_host_free.argtypes = [vp]
        lib.hpt_host_free.restype = None
        lib.hpt_host_build_seconds.argtypes = [vp]
        lib.hpt_host_build_seconds.restype = C.c_double
        _host = lib
    return _host


def _init_torch_runtime_first() -> None:
    """torch carries its own copy of the HIP runtime.  That copy reports no
    GPU when another HIP runtime in the same process has opened the device
    first, while the reverse order works; so when torch is importable its
    runtime is initialised before libhptracing.so makes its first call."""
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def device() -> C.CDLL:
    """libhptracing.so with argument types declared."""
    global _device
    if _device is None:
        _init_torch_runtime_first()
        lib = _load(DEVICE_LIB)
        vp, i64 = C.c_void_p, C.c_int64
        lib.hpt_last_error.restype = C.c_char_p
        lib.hpt_scene_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(vp)]
        lib.hpt_scene_destroy.argtypes = [vp]
        lib.hpt_scene_stats_get.argtypes = [vp, C.POINTER(SceneStats)]
        lib.hpt_cast.argtypes = [vp, vp, vp, i64, vp, vp, vp]
        lib.hpt_cast_counted.argtypes = [vp, vp, vp, i64, vp, vp, vp, vp]
        lib.hpt_cast_noalloc.argtypes = [vp, vp, vp, i64, vp, vp, vp]
        lib.hpt_render.argtypes = [vp, C.POINTER(View), vp, vp, C.POINTER(RenderStats)]
        lib.hpt_render_ex.argtypes = [vp, C.POINTER(View), vp, vp, C.POINTER(RenderStats),
                                      C.c_uint]
        lib.hpt_batch_render.argtypes = [vp, C.POINTER(View), C.c_int, C.POINTER(vp),
                                         C.POINTER(vp), C.POINTER(RenderStats), C.c_uint]
        lib.hpt_scene_set_arithmetic.argtypes = [vp, C.c_int]
        lib.hpt_autotune.argtypes = [vp, C.POINTER(View), C.c_int, C.POINTER(Tuning)]
        lib.hpt_autotune_cast.argtypes = [vp, vp, vp, i64, i64, C.POINTER(Tuning)]
        lib.hpt_scene_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        lib.hpt_scene_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        lib.hpt_server_create.argtypes = [vp, C.POINTER(ServerOpts), C.POINTER(vp)]
        lib.hpt_server_submit.argtypes = [vp, C.POINTER(View), vp, vp, C.POINTER(C.c_int64)]
        lib.hpt_server_submit_ex.argtypes = [vp, C.POINTER(View), vp, vp, C.c_uint,
                                             C.POINTER(C.c_int64)]
        lib.hpt_server_peek.argtypes = [vp, C.c_int64, C.c_int32, C.c_double, vp,
                                        C.POINTER(C.c_int32)]
        lib.hpt_server_wait.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(RequestInfo)]
        lib.hpt_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
        lib.hpt_server_destroy.argtypes = [vp]
        _device = lib
    return _device


def check_host(rc: int) -> None:
    if rc != HPT_OK:
        raise HptError(rc, host().hpt_host_last_error().decode(errors="replace"))


def check_device(rc: int) -> None:
    if rc != HPT_OK:
        raise HptError(rc, device().hpt_last_error().decode(errors="replace"))
