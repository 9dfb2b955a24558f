"""The certified reciprocal slab test (box_hit_rcp in
hptracing_amd/csrc/device/dmath.hpp) as a method, on the CPU: quotients
formed as (p - s) * y with y within 1 ulp of 1/d (here the correctly rounded
reciprocal moved by -1, 0 or +1 ulp at random, the accuracy of v_rcp_f32),
entry/exit compared with the 2^-20 relative margin, no start-inside test,
and the division form where the margin does not certify, where the exit is
0, and for rays with a zero direction component.  Every decision must equal the division
form's (the slab test of kernel.cl:51-83, restated by the oracle) on
adversarial rays aimed at box corners, edges and faces, and the fallback
must be rare on random rays.  The GPU kernels themselves are checked against the oracle by
test_gpu_cast.py."""
from __future__ import annotations

import numpy as np

F = np.float32


def _slab(bs, be, sp, q0, q1):
    entry = np.fmax(np.fmax(np.fmin(q0[:, 0], q1[:, 0]), np.fmin(q0[:, 1], q1[:, 1])),
                    np.fmin(q0[:, 2], q1[:, 2]))
    exit_ = np.fmin(np.fmin(np.fmax(q0[:, 0], q1[:, 0]), np.fmax(q0[:, 1], q1[:, 1])),
                    np.fmax(q0[:, 2], q1[:, 2]))
    inside = np.all((sp >= bs) & (sp <= be), axis=1)
    return inside, entry, exit_


def box_div(bs, be, sp, d):
    with np.errstate(all="ignore"):
        inside, entry, exit_ = _slab(bs, be, sp, (bs - sp) / d, (be - sp) / d)
    return inside | ((exit_ > 0) & (entry <= exit_))


MARGIN = F(2.0 ** -20)


def _certified(a, b, margin):
    """|b - a| is at least `margin` of |a| + |b| (as the kernel computes it)."""
    return np.abs(b - a) >= (np.abs(a) + np.abs(b)) * margin


def box_rcp(bs, be, sp, d, y, margin=MARGIN):
    """The certified form: no start-inside test (a start point inside the
    box makes entry <= 0 <= exit, so only exit == 0 needs it), and rays with
    a zero direction component take the division form throughout, as
    fast_box_ray sends them there.  Returns the decisions and which tests
    the division form decided."""
    with np.errstate(all="ignore"):
        _, entry, exit_ = _slab(bs, be, sp, (bs - sp) * y, (be - sp) * y)
        sure = _certified(entry, exit_, margin)
    ahead = exit_ > 0
    fast = ahead & (entry <= exit_)
    need = (exit_ == 0) | (ahead & ~sure) | np.any(d == 0, axis=1)
    out = fast.copy()
    out[need] = box_div(bs[need], be[need], sp[need], d[need])
    return out, need


def _approx_rcp(d, rng):
    with np.errstate(divide="ignore"):
        y = (F(1) / d).astype(F)          # +-inf for d = +-0, as v_rcp_f32
    step = np.where(d == 0, 0, rng.integers(-1, 2, y.shape))
    y = np.where(step > 0, np.nextafter(y, F(np.inf)), y)
    return np.where(step < 0, np.nextafter(y, F(-np.inf)), y).astype(F)


def _rays(n, rng):
    lo = rng.uniform(-10, 10, (n, 3)).astype(F)
    hi = (lo + rng.uniform(0.001, 5, (n, 3))).astype(F)
    lo[: n // 8] = np.round(lo[: n // 8])          # faces on integer planes
    hi[: n // 8] = lo[: n // 8] + np.round(rng.uniform(1, 4, (n // 8, 3))).astype(F)
    sel = rng.integers(0, 2, (n, 3)) == 1
    corner = np.where(sel, hi, lo)
    ax = rng.integers(0, 3, n)
    r = np.arange(n)
    edge = corner.copy()
    edge[r, ax] = (lo + rng.uniform(0, 1, (n, 3)).astype(F) * (hi - lo))[r, ax]
    kind = rng.integers(0, 4, n)
    target = np.where((kind == 0)[:, None], corner, edge)
    org = (target + rng.normal(size=(n, 3)).astype(F) * 10).astype(F)
    d = (target - org).astype(F)
    # rays starting on a face of the box, some parallel to an axis
    on = kind == 3
    org[on] = edge[on]
    d[on] = rng.normal(size=(on.sum(), 3)).astype(F)
    d[on & (r % 2 == 0), 1] = 0
    # axis-aligned rays through edges
    aa = kind == 2
    d[aa & (r % 3 == 0)] = np.eye(3, dtype=F)[ax[aa & (r % 3 == 0)]]
    org[aa & (r % 3 == 0)] = (edge[aa & (r % 3 == 0)] - 7 * d[aa & (r % 3 == 0)]).astype(F)
    return lo, hi, org, d


def test_certified_slab_test_matches_division():
    rng = np.random.default_rng(7)
    need_total = 0
    for _ in range(4):
        lo, hi, org, d = _rays(500_000, rng)
        y = _approx_rcp(d, rng)
        ref = box_div(lo, hi, org, d)
        got, need = box_rcp(lo, hi, org, d, y)
        bad = np.nonzero(got != ref)[0]
        assert len(bad) == 0, (lo[bad[:3]], hi[bad[:3]], org[bad[:3]], d[bad[:3]])
        need_total += int(need.sum())
    # the rays are built to be near-ties, so the division form decides many
    assert need_total > 100_000


def test_certified_slab_test_random_rays_rarely_fall_back():
    rng = np.random.default_rng(8)
    n = 1_000_000
    lo = rng.uniform(-10, 10, (n, 3)).astype(F)
    hi = (lo + rng.uniform(0.01, 5, (n, 3))).astype(F)
    org = rng.uniform(-20, 20, (n, 3)).astype(F)
    d = rng.normal(size=(n, 3)).astype(F)
    y = _approx_rcp(d, rng)
    ref = box_div(lo, hi, org, d)
    got, need = box_rcp(lo, hi, org, d, y)
    assert np.array_equal(got, ref)
    assert need.mean() < 1e-4, need.mean()


def test_margin_is_needed():
    """With no margin the reciprocal form does decide some near-ties
    differently from the division form, so the certification is not vacuous."""
    rng = np.random.default_rng(9)
    lo, hi, org, d = _rays(500_000, rng)
    y = _approx_rcp(d, rng)
    with np.errstate(all="ignore"):
        inside, entry, exit_ = _slab(lo, hi, org, (lo - org) * y, (hi - org) * y)
    naive = inside | ((exit_ > 0) & (entry <= exit_))
    assert (naive != box_div(lo, hi, org, d)).sum() > 0


def test_start_inside_without_inside_test():
    """Start points inside the closed box, on its faces, edges and corners,
    with directions that have no zero component: the certified form, which
    has no start-inside test, decides as the division form does; and the
    exit == 0 fallback is what keeps it exact (without it, start points on
    a face the ray leaves through are reported as misses)."""
    rng = np.random.default_rng(10)
    n = 400_000
    lo = rng.uniform(-10, 10, (n, 3)).astype(F)
    hi = (lo + rng.uniform(0.001, 5, (n, 3))).astype(F)
    u = rng.uniform(0, 1, (n, 3)).astype(F)
    snap = rng.integers(0, 3, (n, 3))          # 0: inside, 1: on lo, 2: on hi
    org = np.where(snap == 1, lo, np.where(snap == 2, hi, (lo + u * (hi - lo)).astype(F)))
    org = np.clip(org, lo, hi).astype(F)
    d = rng.normal(size=(n, 3)).astype(F)
    d[d == 0] = F(1)
    y = _approx_rcp(d, rng)
    ref = box_div(lo, hi, org, d)
    assert ref.all()                           # every start point is in the closed box
    got, need = box_rcp(lo, hi, org, d, y)
    assert np.array_equal(got, ref)
    with np.errstate(all="ignore"):
        _, entry, exit_ = _slab(lo, hi, org, (lo - org) * y, (hi - org) * y)
    no_fallback = (exit_ > 0) & (entry <= exit_)
    assert (no_fallback != ref).sum() > 0


def test_margin_is_four_times_the_failing_one():
    """On the adversarial near-ties the model's quotients (reciprocal within
    1.5 ulp, then rounded) stay exact down to a margin of 2^-22; at 2^-23,
    8 times tighter than the kernel's 2^-20, decisions differ from the
    division form.  So the margin sits a factor 4 above where it fails, as
    the error bound (2^-21.9 per quotient) says."""
    rng = np.random.default_rng(12)
    lo, hi, org, d = _rays(500_000, rng)
    y = _approx_rcp(d, rng)
    ref = box_div(lo, hi, org, d)
    got, _ = box_rcp(lo, hi, org, d, y, margin=F(2.0 ** -22))
    assert np.array_equal(got, ref)
    got, _ = box_rcp(lo, hi, org, d, y, margin=F(2.0 ** -23))
    assert (got != ref).sum() > 1000
