"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
import this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OrcScene(C.Structure):
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


class OrcView(C.Structure):
    _fields_ = [
        ("view", C.c_float * 3), ("up", C.c_float * 3), ("right", C.c_float * 3),
        ("width", C.c_float), ("height", C.c_float),
        ("sample_x", C.c_int32), ("sample_y", C.c_int32),
        ("supersample_x", C.c_int32), ("supersample_y", C.c_int32),
        ("background", C.c_float * 3),
        ("sample", C.c_int32), ("depth", C.c_int32), ("disable_diffuse", C.c_int32),
        ("brightness", C.c_float), ("seed", C.c_uint64),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, i64, f = C.c_void_p, C.c_int64, C.c_float
        L.orc_box_intersect.argtypes = [vp, vp, vp, vp]
        L.orc_single_intersect.argtypes = [vp, vp, vp, vp, vp]
        L.orc_single_intersect.restype = f
        L.orc_rand.argtypes = [C.POINTER(C.c_int64)]
        L.orc_rand.restype = C.c_int64
        L.orc_randf.argtypes = [C.POINTER(C.c_int64)]
        L.orc_randf.restype = f
        L.orc_randf3.argtypes = [C.POINTER(C.c_int64), vp]
        L.orc_kd_cast.argtypes = [C.POINTER(OrcScene), vp, vp, i64, vp, vp, vp, C.c_int]
        L.orc_naive_cast.argtypes = [C.POINTER(OrcScene), vp, vp, i64, vp, vp, C.c_int]
        L.orc_set_viewdirs.argtypes = [vp, vp, vp, f, f, C.c_int, C.c_int, vp]
        L.orc_set_viewdirs_vr.argtypes = [C.c_int] * 5 + [vp]
        L.orc_get_image.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f, vp]
        L.orc_read_texture.argtypes = [C.POINTER(OrcScene), C.c_int, f, f, vp]
        L.orc_render.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcView), C.c_int, C.c_int,
                                 vp, vp, vp]
        L.orc_render_wavefront.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcView), C.c_int,
                                           vp, vp, vp]
        L.orc_render_partial.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcView), i64, i64,
                                         C.c_int, vp]
        L.orc_render_partial.restype = i64
        L.orc_render_rays.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcView), vp, i64, C.c_int,
                                      vp]
        L.orc_render_rays.restype = i64
        L.orc_path_seed.argtypes = [C.c_uint64, i64, C.c_int32]
        L.orc_path_seed.restype = C.c_uint64
        L.orc_set_arithmetic.argtypes = [C.c_int]
        L.orc_set_arithmetic.restype = None
        L.orc_dot_cross.argtypes = [C.c_int, vp, vp, i64, vp]
        L.orc_dot_cross.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data if a is not None and a.size else 0


class Scene:
    """Oracle view of a hptracing_amd.scene.KDTree (arrays are kept alive)."""

    def __init__(self, kd):
        self.kd = kd
        self._keep = [kd.points, kd.normals, kd.texcoords, kd.geometries, kd.materials,
                      kd.lights, kd.nodes, kd.leaf_data, kd.texture_data]
        s = OrcScene()
        s.points, s.n_points = _p(kd.points), len(kd.points)
        s.normals, s.n_normals = _p(kd.normals), len(kd.normals)
        s.texcoords, s.n_texcoords = _p(kd.texcoords), len(kd.texcoords)
        s.mesh, s.n_triangles = _p(kd.geometries), len(kd.geometries)
        s.materials, s.n_materials = _p(kd.materials), len(kd.materials)
        s.lights, s.n_lights = _p(kd.lights), len(kd.lights)
        s.nodes, s.n_nodes = _p(kd.nodes), len(kd.nodes)
        s.leaf_data, s.n_leaf_data = _p(kd.leaf_data), len(kd.leaf_data)
        s.textures, s.n_textures = _p(kd.texture_data), kd.n_textures
        self.c = s

    def kd_cast(self, org, dir, counts=False, n_threads=0):
        org = np.ascontiguousarray(org, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        n = len(org)
        ids = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        cnt = np.empty((n, 5), np.int32) if counts else None
        lib().orc_kd_cast(C.byref(self.c), _p(org), _p(dir), n, _p(ids), _p(t),
                          _p(cnt) if counts else None, n_threads)
        return (ids, t, cnt) if counts else (ids, t)

    def naive_cast(self, org, dir, n_threads=0):
        org = np.ascontiguousarray(org, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        n = len(org)
        ids = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        lib().orc_naive_cast(C.byref(self.c), _p(org), _p(dir), n, _p(ids), _p(t), n_threads)
        return ids, t

    def read_texture(self, layer, u, v):
        out = np.empty(3, np.float32)
        lib().orc_read_texture(C.byref(self.c), layer, u, v, _p(out))
        return out

    def render(self, view: OrcView, seed_mode=0, n_threads=0, wavefront=False):
        R = view.sample_x * view.supersample_x * view.sample_y * view.supersample_y
        accum = np.zeros(R * 3, np.float32)
        rgba = np.zeros(view.sample_x * view.sample_y * 4, np.uint8)
        live = np.zeros(max(view.depth, 1), np.int64)
        if wavefront:
            rc = lib().orc_render_wavefront(C.byref(self.c), C.byref(view), seed_mode,
                                            _p(accum), _p(rgba), _p(live))
        else:
            rc = lib().orc_render(C.byref(self.c), C.byref(view), seed_mode, n_threads,
                                  _p(accum), _p(rgba), _p(live))
        if rc != 0:
            raise RuntimeError("orc_render failed")
        return accum.reshape(-1, 3), rgba.reshape(view.sample_y, view.sample_x, 4), live

    def render_partial(self, view: OrcView, begin: int, end: int, n_threads=0):
        R = view.sample_x * view.supersample_x * view.sample_y * view.supersample_y
        accum = np.zeros(R * 3, np.float32)
        n = lib().orc_render_partial(C.byref(self.c), C.byref(view), begin, end, n_threads,
                                     _p(accum))
        return accum.reshape(-1, 3), int(n)


    def render_rays(self, view: OrcView, rays, n_threads=0, accum=None):
        """The render of the listed camera rays only; accum (3 floats per
        camera ray of the view) is allocated unless given.  Returns (accum,
        rays cast)."""
        R = view.sample_x * view.supersample_x * view.sample_y * view.supersample_y
        rays = np.ascontiguousarray(rays, np.int64)
        if accum is None:
            accum = np.zeros(R * 3, np.float32)
        n = lib().orc_render_rays(C.byref(self.c), C.byref(view), _p(rays), len(rays), n_threads,
                                  _p(accum))
        if n < 0:
            raise RuntimeError("orc_render_rays failed")
        return accum.reshape(-1, 3), int(n)


def make_view(**kw) -> OrcView:
    v = OrcView()
    defaults = dict(view=(0, 0, -500), up=(0, 1, 0), right=(1, 0, 0), width=1.0, height=1.0,
                    sample_x=64, sample_y=64, supersample_x=1, supersample_y=1,
                    background=(0, 0, 0), sample=4, depth=5, disable_diffuse=0,
                    brightness=1.0, seed=1)
    defaults.update(kw)
    for k, val in defaults.items():
        if isinstance(val, (tuple, list)):
            getattr(v, k)[:] = [float(x) for x in val]
        else:
            setattr(v, k, val)
    return v


def box_intersect(bs, be, sp, d) -> bool:
    a = [np.asarray(x, np.float32) for x in (bs, be, sp, d)]
    return bool(lib().orc_box_intersect(*[_p(x) for x in a]))


def single_intersect(sp, d, pa, pb, pc) -> float:
    a = [np.asarray(x, np.float32) for x in (sp, d, pa, pb, pc)]
    return float(lib().orc_single_intersect(*[_p(x) for x in a]))


def rand_sequence(seed: int, n: int, kind: str = "rand"):
    s = C.c_int64(seed)
    out = []
    for _ in range(n):
        if kind == "rand":
            out.append(int(lib().orc_rand(C.byref(s))))
        elif kind == "randf":
            out.append(float(lib().orc_randf(C.byref(s))))
        else:
            v = np.empty(3, np.float32)
            lib().orc_randf3(C.byref(s), _p(v))
            out.append(v)
    return out, s.value


def viewdirs(view, up, right, width, height, sx, sy):
    d = np.empty(sx * sy * 3, np.float32)
    f3 = lambda x: np.asarray(x, np.float32)
    a, b, c = f3(view), f3(up), f3(right)
    lib().orc_set_viewdirs(_p(a), _p(b), _p(c), width, height, sx, sy, _p(d))
    return d.reshape(-1, 3)


def viewdirs_vr(x_shift, sample_x, sample_y, grid_x, grid_y):
    d = np.empty(grid_x * grid_y * 3, np.float32)
    lib().orc_set_viewdirs_vr(x_shift, sample_x, sample_y, grid_x, grid_y, _p(d))
    return d.reshape(-1, 3)


def get_image(accum, width, height, ssx, ssy, spp, brightness):
    accum = np.ascontiguousarray(accum, np.float32)
    out = np.empty(width * height * 4, np.uint8)
    lib().orc_get_image(_p(accum), width, height, ssx, ssy, spp, brightness, _p(out))
    return out.reshape(height, width, 4)


def path_seed(seed: int, ray: int, pass_: int) -> int:
    return int(lib().orc_path_seed(seed, ray, pass_))


ARITH_LITERAL = 0
ARITH_ROCM_OPENCL = 1


def set_arithmetic(mode: int) -> None:
    """dot/cross evaluation of the intersection test (process-wide;
    ARITH_ROCM_OPENCL by default, as in the library)."""
    lib().orc_set_arithmetic(mode)


_lib64 = None


def render64(scene: "Scene", view: OrcView, n_threads: int = 0):
    """Float64 restatement (liboracle64.so): returns (sum, sum of squares)
    of the per-pass radiance per camera ray, each (R, 3) float64."""
    global _lib64
    if _lib64 is None:
        path = os.path.join(HERE, "liboracle64.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc64_render.argtypes = [C.POINTER(OrcScene), C.POINTER(OrcView), C.c_int,
                                   C.c_void_p, C.c_void_p]
        _lib64 = L
    R = view.sample_x * view.supersample_x * view.sample_y * view.supersample_y
    s = np.zeros(R * 3, np.float64)
    s2 = np.zeros(R * 3, np.float64)
    if _lib64.orc64_render(C.byref(scene.c), C.byref(view), n_threads, _p(s), _p(s2)) != 0:
        raise RuntimeError("orc64_render failed")
    return s.reshape(-1, 3), s2.reshape(-1, 3)


def dot_cross(mode: int, a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.empty((len(a), 4), np.float32)
    lib().orc_dot_cross(mode, _p(a), _p(b), len(a), _p(out))
    return out[:, 0], out[:, 1:]
