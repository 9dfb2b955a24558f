"""ctypes declarations of oracle/ref_harness.hip's extern "C" functions, the
launcher of the reference's own kernels.  Every test that calls the harness
loads it through `load()`, so the argument lists live in one place, and
tests/test_ref_abi.py checks them against the C prototypes on the CPU: a
list one pointer short makes the C side read that pointer from whatever the
stack holds, and the harness then writes through it."""
from __future__ import annotations

import ctypes as C
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
LIB = os.path.join(REF_DIR, "libref_harness.so")
SOURCE = os.path.join(ROOT, "oracle", "ref_harness.hip")

vp, i64, i32, f32, f64 = C.c_void_p, C.c_int64, C.c_int32, C.c_float, C.c_double

# name -> (restype, argtypes), in the prototypes' order.
SIGNATURES = {
    "ref_last_error": (C.c_char_p, []),
    "ref_kdtree_intersect_timed": (C.c_int, [
        C.c_char_p,            # co_path
        vp, i64,               # points
        vp, i64,               # mesh
        vp, i64,               # nodes
        vp, i64,               # leaf data
        vp, vp, i32,           # org, dir, n
        vp, vp,                # geometry, t
        vp,                    # kernel_ms
    ]),
    "ref_kdtree_intersect": (C.c_int, [
        C.c_char_p, vp, i64, vp, i64, vp, i64, vp, i64, vp, vp, i32, vp, vp,
    ]),
    "ref_render": (C.c_int, [
        C.c_char_p,            # co_path
        vp, i64,               # points
        vp, i64,               # normals
        vp, i64,               # texcoords
        vp, i64,               # mesh
        vp, i64,               # materials
        vp, i64,               # lights
        vp, i64,               # nodes
        vp, i64,               # leaf data
        vp, vp, vp,            # view, up, right
        f32, f32,              # width, height
        i32, i32, i32, i32,    # sample_x, sample_y, ssx, ssy
        vp,                    # background
        i32, i32, i32,         # sample, depth, disable_diffuse
        vp,                    # seeds
        vp, vp, vp, vp,        # results, s0_sizes, loop_seconds, primary_geometry
    ]),
}


def available() -> bool:
    return os.path.exists(LIB)


def load() -> C.CDLL:
    """The harness with every function's types declared.  torch's HIP
    runtime must be the first in the process (it reports no GPU after
    another copy has opened the device), so the library's own loader runs
    first."""
    from hptracing_amd import _native as N
    N.device()
    lib = C.CDLL(LIB)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
