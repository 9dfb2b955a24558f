"""Pins of tests/f32ops.py, the vectorised float32 intersection test used
where the oracle's scalar walk is too slow (millions of rays)."""
from __future__ import annotations

import ctypes
import ctypes.util
import os

import numpy as np
import oracle as O
from conftest import GOLDEN
from f32ops import dot_cross, fma32, moller_trumbore

F32 = np.float32


def _libm_fmaf(a, b, c):
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.fmaf.restype = ctypes.c_float
    libm.fmaf.argtypes = [ctypes.c_float] * 3
    return np.array([libm.fmaf(float(x), float(y), float(z)) for x, y, z in zip(a, b, c)], F32)


def test_fma32_double_rounding_cases():
    """Sums that lie exactly halfway between two float32 values once rounded
    to float64: (1 + 2^-23) * 2^-24 (1 - 2^-23) + c is 2^-24 - 2^-70 above
    c.  With c = 1 + 2^-23 (odd last bit) rounding twice goes up to
    1 + 2^-22; the one correct rounding stays at 1 + 2^-23."""
    a = F32(1 + 2.0 ** -23)
    b = F32(2.0 ** -24 * (1 - 2.0 ** -23))
    for c, want in [(1 + 2.0 ** -23, 1 + 2.0 ** -23), (1.0, 1.0)]:
        got = fma32(np.array([a]), np.array([b]), np.array([F32(c)]))[0]
        assert got == F32(want), (c, got)
        assert _libm_fmaf([a], [b], [F32(c)])[0] == F32(want)
        # the naive float64 route errs on the first case
    naive = F32(np.float64(a) * np.float64(b) + np.float64(F32(1 + 2.0 ** -23)))
    assert naive == F32(1 + 2.0 ** -22)


def test_fma32_matches_libm_fmaf():
    """Random operands over wide exponent ranges, cancellation and scaled
    halfway constructions, against glibc's correctly rounded fmaf."""
    rng = np.random.default_rng(3)
    n = 40000
    a = (rng.standard_normal(n) * 2.0 ** rng.integers(-30, 30, n)).astype(F32)
    b = (rng.standard_normal(n) * 2.0 ** rng.integers(-30, 30, n)).astype(F32)
    c = (rng.standard_normal(n) * 2.0 ** rng.integers(-60, 60, n)).astype(F32)
    c[: n // 4] = -(a[: n // 4] * b[: n // 4]).astype(F32)          # cancellation
    k = rng.integers(-20, 20, n // 4).astype(np.float64)
    s = rng.choice([-1.0, 1.0], n // 4)
    a[n // 4: n // 2] = F32(1 + 2.0 ** -23)
    b[n // 4: n // 2] = (s * 2.0 ** (k - 24) * (1 - 2.0 ** -23)).astype(F32)
    c[n // 4: n // 2] = (s * 2.0 ** k * (1 + 2.0 ** -23 * rng.integers(0, 2, n // 4))).astype(F32)
    got = fma32(a, b, c)
    want = _libm_fmaf(a, b, c)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_fused_dot_cross_match_rocm_library_outputs():
    """The fused dot/cross equal the outputs recorded from ROCm's OpenCL
    device library on gfx950 (tests/golden/rocm_ocl_builtins) bit for bit."""
    z = np.load(os.path.join(GOLDEN, "rocm_ocl_builtins", "samples.npz"))
    d, c = dot_cross(z["a"], z["b"], fused=True)
    assert np.array_equal(d.view(np.int32), z["dot"].view(np.int32))
    assert np.array_equal(c.view(np.int32), z["cross"].view(np.int32))
    d0, c0 = dot_cross(z["a"], z["b"], fused=False)
    d1, c1 = O.dot_cross(O.ARITH_LITERAL, z["a"], z["b"])
    assert np.array_equal(d0.view(np.int32), d1.view(np.int32))
    assert np.array_equal(c0.view(np.int32), c1.view(np.int32))


def test_moller_trumbore_matches_oracle_both_forms():
    """Per ray, t equals the oracle's _single_intersect in both arithmetic
    modes, bit for bit (hits, misses, slivers)."""
    rng = np.random.default_rng(8)
    n = 3000
    pa = rng.uniform(0, 1, (n, 3)).astype(F32)
    e1 = (rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-3, -1, (n, 1))).astype(F32)
    e2 = (rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-3, -1, (n, 1))).astype(F32)
    pb, pc = (pa + e1).astype(F32), (pa + e2).astype(F32)
    target = (pa + 0.3 * e1 + 0.3 * e2).astype(F32)
    org = rng.uniform(-1, 2, (n, 3)).astype(F32)
    d = (target - org + rng.standard_normal((n, 3)).astype(F32) * 1e-3).astype(F32)
    for mode, fused in [(O.ARITH_ROCM_OPENCL, True), (O.ARITH_LITERAL, False)]:
        O.set_arithmetic(mode)
        try:
            want = np.array([O.single_intersect(org[i], d[i], pa[i], pb[i], pc[i])
                             for i in range(n)], F32)
        finally:
            O.set_arithmetic(O.ARITH_ROCM_OPENCL)
        got, _, _ = moller_trumbore(org, d, pa, pb, pc, fused=fused)
        assert (got > 0).mean() > 0.3
        assert np.array_equal(got.view(np.int32), want.view(np.int32)), mode
