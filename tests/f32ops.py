"""Float32 arithmetic of the intersection test, vectorised in numpy, for
checks over millions of rays where the oracle's scalar walk is too slow.

`fma32` is a correctly rounded float32 fused multiply-add.  numpy has none,
and rounding a*b + c first to float64 and then to float32 can round twice
wrongly, so the float64 sum's error is carried exactly (Knuth's TwoSum; the
product of two float32 values is exact in float64) and decides the one case
where double rounding errs: a float64 sum lying exactly halfway between two
float32 neighbours.

`moller_trumbore` restates `_single_intersect` (kernel.cl:85-109) with
dot/cross either literal (every product and sum rounded) or fused as ROCm's
OpenCL device library evaluates them:
  dot(a,b)   = fma(a.z, b.z, fma(a.y, b.y, a.x*b.x))
  cross(a,b) = (fma(a.y, b.z, -(a.z*b.y)), fma(a.z, b.x, -(a.x*b.z)),
                fma(a.x, b.y, -(a.y*b.x)))
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
F64 = np.float64


def fma32(a, b, c):
    """round_to_float32(a*b + c) with one rounding (IEEE fmaf)."""
    a, b, c = (np.asarray(x, F32).astype(F64) for x in (a, b, c))
    with np.errstate(all="ignore"):
        p = a * b                      # exact: 24 + 24 significant bits
        s = p + c                      # rounded to float64
        bb = s - p                     # TwoSum: s + e == p + c exactly
        e = (p - (s - bb)) + (c - bb)
        r = s.astype(F32)              # round to nearest even
        diff = s - r.astype(F64)
        toward = np.where(diff > 0, F32(np.inf), F32(-np.inf))
        other = np.nextafter(r, toward)
        mid = (r.astype(F64) + other.astype(F64)) * 0.5
        tie = (diff != 0) & (s == mid) & (e != 0) & np.isfinite(s)
        # On an exact halfway float64 sum the true value lies on the side
        # of the error: past the midpoint (same sign as diff) -> other.
        away = tie & (np.sign(e) == np.sign(diff))
        return np.where(away, other, r).astype(F32)


def _dot(a, b, fused):
    if fused:
        return fma32(a[:, 2], b[:, 2], fma32(a[:, 1], b[:, 1], (a[:, 0] * b[:, 0]).astype(F32)))
    return ((a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]).astype(F32)


def _cross(a, b, fused):
    if fused:
        return np.stack([fma32(a[:, 1], b[:, 2], -(a[:, 2] * b[:, 1]).astype(F32)),
                         fma32(a[:, 2], b[:, 0], -(a[:, 0] * b[:, 2]).astype(F32)),
                         fma32(a[:, 0], b[:, 1], -(a[:, 1] * b[:, 0]).astype(F32))], 1)
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1).astype(F32)


def dot_cross(a, b, fused):
    """(dot, cross) of float32 (n, 3) arrays."""
    a = np.asarray(a, F32)
    b = np.asarray(b, F32)
    with np.errstate(all="ignore"):
        return _dot(a, b, fused), _cross(a, b, fused)


def moller_trumbore(org, d, pa, pb, pc, fused=True):
    """t per ray (-1 on a miss) and (u, v), float32, kernel.cl's order."""
    org, d, pa, pb, pc = (np.asarray(x, F32) for x in (org, d, pa, pb, pc))
    with np.errstate(all="ignore"):
        e1, e2 = (pb - pa).astype(F32), (pc - pa).astype(F32)
        P = _cross(d, e2, fused)
        inv = (F32(1) / _dot(e1, P, fused)).astype(F32)
        T = (org - pa).astype(F32)
        u = (_dot(T, P, fused) * inv).astype(F32)
        Q = _cross(T, e1, fused)
        v = (_dot(d, Q, fused) * inv).astype(F32)
        t = (_dot(e2, Q, fused) * inv).astype(F32)
        ok = ~((u < 0) | (u > 1) | (v < 0) | (u + v > 1)) & (t > 0)
    return np.where(ok, t, F32(-1)), u, v
