"""Pins of the CPU oracle (oracle/) against the reference's known answers,
closed forms, brute force and published constants."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O
from hptracing_amd import scenes


# ------------------------------------------------------------------- slab
def test_box_intersect_known_answers():
    """SceneCLTest.box_intersect (tests/src/test_scene_cl.cpp:91-116)."""
    bs, be = (0, 0, 0), (100, 100, 100)
    n = lambda v: (np.asarray(v, np.float64) / np.linalg.norm(v)).astype(np.float32)
    assert O.box_intersect(bs, be, (50, 50, -100), (0, 0, 1))
    assert not O.box_intersect(bs, be, (50, 50, -100), (0, 0, -1))
    assert O.box_intersect(bs, be, (-100, -100, -100), n((1, 1, 1)))
    assert not O.box_intersect(bs, be, (-100, -100, -100), n((-1, 1, 1)))
    assert not O.box_intersect(bs, be, (-100, -100, -100), n((-1, 0, 0)))
    assert not O.box_intersect(bs, be, (-100, -100, -100), n((-1, -2, -3)))
    assert O.box_intersect(bs, be, (-100, 50, -100), n((1, 0, 1)))


def test_box_intersect_edge_cases():
    bs, be = (0, 0, 0), (1, 1, 1)
    assert O.box_intersect(bs, be, (1, 1, 1), (0, 0, 0))          # start on the box: inside
    assert not O.box_intersect(bs, be, (2, 0.5, 0.5), (1, 0, 0))   # exit behind the start
    assert O.box_intersect(bs, be, (2, 0.5, 0.5), (-1, 0, 0))
    # a zero direction component outside the slab: -inf..inf handled by fmin/fmax
    assert not O.box_intersect(bs, be, (-1, 2, 0.5), (1, 0, 0))
    # the start on the slab plane with a zero component gives 0/0 = NaN, which
    # fmin/fmax drop: both bounds become the other quotient (-inf), so a ray
    # running along the top face is rejected
    assert not O.box_intersect(bs, be, (-1, 1, 0.5), (1, 0, 0))
    assert O.box_intersect(bs, be, (-1, 0.5, 0.5), (1, 0, 0))


# ---------------------------------------------------------- Moller-Trumbore
def test_single_intersect_against_float64():
    rng = np.random.default_rng(1)
    agree = 0
    for _ in range(3000):
        pa, pb, pc = rng.normal(size=(3, 3)).astype(np.float32)
        sp = (rng.normal(size=3) * 3).astype(np.float32)
        target = (pa + rng.uniform(0, 1) * (pb - pa) * 0.5 + rng.uniform(0, 1) * (pc - pa) * 0.5)
        d = (target - sp + rng.normal(size=3) * 0.3).astype(np.float32)
        t = O.single_intersect(sp, d, pa, pb, pc)
        # float64 Cramer solve of sp + t d = pa + u e1 + v e2
        e1, e2 = (pb - pa).astype(np.float64), (pc - pa).astype(np.float64)
        A = np.stack([e1, e2, -d.astype(np.float64)], 1)
        u, v, tt = np.linalg.solve(A, (sp - pa).astype(np.float64))
        inside = u >= 0 and v >= 0 and u + v <= 1 and tt > 0
        margin = min(abs(u), abs(v), abs(1 - u - v), abs(u - 1))
        if margin < 1e-4:
            continue
        assert (t > 0) == inside
        if inside:
            assert abs(t - tt) <= 1e-4 * max(1.0, abs(tt))
        agree += 1
    assert agree > 2500


def test_single_intersect_degenerate_guard_inert():
    """`det > EPSILON && det < EPSILON` never holds (kernel.cl:93): a ray in
    the triangle's plane is decided by NaN/inf propagation, never by it."""
    pa, pb, pc = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    assert O.single_intersect((0.2, 0.2, 1), (0, 0, -1), pa, pb, pc) == pytest.approx(1.0)
    assert O.single_intersect((-1, 0.2, 0), (1, 0, 0), pa, pb, pc) == -1
    assert O.single_intersect((0.2, 0.2, 1), (0, 0, 1), pa, pb, pc) == -1


# ------------------------------------------------------------- traversal
def _rays(n, lo, hi, seed):
    return scenes.random_rays(n, lo, hi, seed)


@pytest.mark.parametrize("which", ["cornell", "room"])
def test_kd_cast_equals_brute_force(which, request):
    """The reference's invariant (test_scene_cl.cpp:118-232): the KD walk
    finds the brute-force nearest hit (naive_intersect, kernel.cl:210-245)."""
    kd = request.getfixturevalue(which)
    s = O.Scene(kd)
    pts = kd.points[:, :3]
    lo, hi = pts.min(0) - 1, pts.max(0) + 1
    org, d = _rays(4000, lo, hi, 21)
    ids, t, cnt = s.kd_cast(org, d, counts=True)
    nid, nt = s.naive_cast(org, d)
    capped = cnt[:, 3] > 0
    ok = ~capped
    # identical t; ids equal unless two triangles give the same t (shared edges)
    assert np.array_equal(t[ok].view(np.int32), nt[ok].view(np.int32))
    same = ids[ok] == nid[ok]
    assert same.mean() > 0.999
    assert (ids >= 0).sum() > 1000


def test_kd_cast_counts_consistent(room):
    s = O.Scene(room)
    org, d = _rays(3000, (0, 0, -2), (10, 10, 10), 5)
    _, _, c = s.kd_cast(org, d, counts=True)
    assert (c[:, 0] <= c[:, 4]).all()
    assert (c[:, 1] <= 512).all()
    assert ((c[:, 3] > 0) <= (c[:, 1] == 512)).all()
    assert (c[:, 2] >= c[:, 1]).all()


def test_kd_cast_soup_matches_naive(soup_small):
    s = O.Scene(soup_small)
    org, d = _rays(2000, (0, 0, 0), (1, 1, 1), 8)
    ids, t, cnt = s.kd_cast(org, d, counts=True)
    nid, nt = s.naive_cast(org, d)
    ok = cnt[:, 3] == 0
    assert ok.mean() > 0.5
    assert np.array_equal(t[ok].view(np.int32), nt[ok].view(np.int32))


def test_rocm_opencl_builtin_forms():
    """The oracle's ORC_ARITH_ROCM_OPENCL dot/cross reproduce the outputs
    recorded from ROCm's OpenCL device library on gfx950 bit for bit; the
    literal forms do not, so the fixture tells the two apart."""
    import os
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, "rocm_ocl_builtins", "samples.npz"))
    d1, c1 = O.dot_cross(O.ARITH_ROCM_OPENCL, z["a"], z["b"])
    assert np.array_equal(d1.view(np.int32), z["dot"].view(np.int32))
    assert np.array_equal(c1.view(np.int32), z["cross"].view(np.int32))
    d0, c0 = O.dot_cross(O.ARITH_LITERAL, z["a"], z["b"])
    assert (d0.view(np.int32) != z["dot"].view(np.int32)).mean() > 0.1
    assert (c0.view(np.int32) != z["cross"].view(np.int32)).mean() > 0.1
    # the library's division is correctly rounded
    q = (z["a"][:, 0] / z["b"][:, 0]).astype(np.float32)
    assert np.array_equal(q.view(np.int32), z["div"].view(np.int32))


# -------------------------------------------------------------------- RNG
def test_lcg_matches_java_random():
    """The 48-bit LCG is java.util.Random's (same multiplier, increment and
    output bits).  new Random(42).nextInt() == -1170105035, published widely;
    java scrambles the seed with the multiplier first."""
    seed = (42 ^ 0x5DEECE66D) & ((1 << 48) - 1)
    vals, _ = O.rand_sequence(seed, 2, "rand")
    as_int = [(v ^ 0x80000000) - 0x80000000 for v in vals]
    assert as_int[0] == -1170105035
    assert as_int[1] == 234785527


def test_randf_and_randf3():
    vals, _ = O.rand_sequence(7, 2000, "randf")
    a = np.array(vals)
    assert (a >= 0).all() and (a <= 1).all() and 0.45 < a.mean() < 0.55
    # float(0xFFFFFFFF) rounds to 2^32, so the largest output maps to 1.0f
    v3, _ = O.rand_sequence(11, 1000, "randf3")
    v3 = np.array(v3)
    assert np.allclose(np.linalg.norm(v3, axis=1), 1, atol=1e-6)
    assert (v3[:, 2] >= 0).all()


def test_path_seed_is_48_bit_and_distinct():
    s = {O.path_seed(1, r, p) for r in range(64) for p in range(8)}
    assert len(s) == 512 and max(s) < (1 << 48)


# ----------------------------------------------------------------- camera
def test_viewdirs_closed_form():
    up, right = (0, 1, 0), (-1, 0, 0)
    d = O.viewdirs((0, 0, 0), up, right, 1.5, 1.0, 6, 4)
    fwd = np.cross(up, right)
    for gy in range(4):
        for gx in range(6):
            ref = fwd + np.array(up) * (gy / 4 * 1.0 - 0.5) + np.array(right) * (gx / 6 * 1.5 - 0.75)
            ref = ref / np.linalg.norm(ref)
            assert np.allclose(d[6 * gy + gx], ref, atol=1e-6)


def test_viewdirs_vr_closed_form():
    d = O.viewdirs_vr(0, 8, 4, 8, 4)
    for gy in range(4):
        for gx in range(8):
            lon = gx / 8 * 2 * math.pi - math.pi
            lat = gy / 4 * math.pi - math.pi / 2
            ref = (math.cos(lat) * math.cos(lon), math.sin(lat), math.cos(lat) * math.sin(lon))
            assert np.allclose(d[8 * gy + gx], ref, atol=2e-6)


# ---------------------------------------------------------------- resolve
def test_get_image_closed_form():
    W, H, ss, spp = 3, 2, 2, 4
    acc = np.zeros((W * ss * H * ss, 3), np.float32)
    vals = np.arange(W * ss * H * ss, dtype=np.float32) * 0.1
    acc[:, 0] = vals
    acc[:, 1] = 1e9
    acc[:, 2] = np.nan
    img = O.get_image(acc, W, H, ss, ss, spp, 1.5)
    for y in range(H):
        for x in range(W):
            idx = [(y * ss + j) * W * ss + x * ss + i for i in range(ss) for j in range(ss)]
            v = vals[idx].sum() / 4 / spp * 1.5
            assert abs(int(img[y, x, 0]) - round((1 - math.exp(-v)) * 255)) <= 1
            assert img[y, x, 1] == 255 and img[y, x, 2] == 0 and img[y, x, 3] == 0


# ---------------------------------------------------------------- texture
def test_texture_bilinear_repeat(tmp_path):
    from hptracing_amd.scene import KDTree
    rng = np.random.default_rng(3)
    src = rng.integers(0, 256, (512, 512, 3)).astype(np.uint8)
    src[0, 0] = (0, 0, 0)
    src[511, 511] = (255, 255, 255)
    with open(tmp_path / "t.ppm", "wb") as fh:
        fh.write(b"P6\n512 512\n255\n" + src.tobytes())
    (tmp_path / "m.mtl").write_text("newmtl t\nKd 1 1 1\nmap_Kd t.ppm\n")
    (tmp_path / "s.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl t\nf 1 2 3\n")
    kd = KDTree(str(tmp_path / "s.obj"), str(tmp_path) + "/")
    s = O.Scene(kd)
    tex = kd.texture_data.reshape(512, 512, 4)[:, :, :3].astype(np.float64) / 255
    # texel centres return the texel
    for (i, j) in [(0, 0), (10, 300), (511, 511)]:
        got = s.read_texture(0, (i + 0.5) / 512, (j + 0.5) / 512)
        assert np.allclose(got, tex[j, i], atol=1e-6)
    # half-way between two texels averages them; REPEAT wraps at the edge
    got = s.read_texture(0, 1.0 / 512, 0.5 / 512)
    assert np.allclose(got, (tex[0, 0] + tex[0, 1]) / 2, atol=1e-6)
    got = s.read_texture(0, 0.0, 0.5 / 512)
    assert np.allclose(got, (tex[0, 511] + tex[0, 0]) / 2, atol=1e-6)
    assert np.allclose(s.read_texture(0, 1.25, -0.75), s.read_texture(0, 0.25, 0.25), atol=1e-6)


# ------------------------------------------------------------------ render
def _cornell_view(**kw):
    v = dict(view=(278, 273, -800), up=(0, 1, 0), right=(-1, 0, 0), width=0.7, height=0.7,
             sample_x=24, sample_y=24, sample=8, depth=5, seed=5)
    v.update(kw)
    return O.make_view(**v)


def test_wavefront_equals_depth_first(cornell):
    """The parallel depth-first engine reorders the literal wavefront of
    TraceRunner::run without changing a bit (per-path seeds)."""
    s = O.Scene(cornell)
    v = _cornell_view()
    a, img, live = s.render(v)
    b, img2, live2 = s.render(v, wavefront=True)
    assert np.array_equal(a.view(np.int32), b.view(np.int32))
    assert np.array_equal(live, live2)
    assert live[0] == 24 * 24


def test_float_oracle_matches_float64_reference(room):
    """The float32 oracle against the independent float64 restatement's
    16384-spp render at 256 spp: >= 99% of pixel means within 4 standard
    errors (plus float32 rounding on constant-radiance pixels)."""
    import os
    from conftest import GOLDEN, within_4se
    ref = np.load(os.path.join(GOLDEN, "room_reference", "reference.npz"))
    spp = 256
    v = O.make_view(view=scenes.ROOM_VIEW["view_p"], up=scenes.ROOM_VIEW["top_dir"],
                    right=scenes.ROOM_VIEW["right_dir"], width=1.0, height=1.0,
                    sample_x=int(ref["width"]), sample_y=int(ref["height"]), sample=spp,
                    depth=int(ref["depth"]), seed=4242)
    acc, _, _ = O.Scene(room).render(v)
    frac, rmse = within_4se(acc.astype(np.float64) / spp, spp, ref)
    assert frac >= 0.99, (frac, rmse)


def test_seed_schemes_statistically_equivalent(cornell):
    """Per-path LCG streams (the GPU scheme) against the reference's slot
    seeds (v_seed[slot] + slot): per-pixel means agree within 4 standard
    errors for >= 99% of pixels."""
    s = O.Scene(cornell)
    n_batches = 8
    outs = {0: [], 1: []}
    for mode in (0, 1):
        for b in range(n_batches):
            v = _cornell_view(sample=16, seed=100 + b)
            acc, _, _ = s.render(v, seed_mode=mode, wavefront=(mode == 1))
            outs[mode].append(acc / 16)
    m0, m1 = np.mean(outs[0], 0), np.mean(outs[1], 0)
    se = np.sqrt(np.var(outs[0], 0, ddof=1) / n_batches + np.var(outs[1], 0, ddof=1) / n_batches)
    ok = np.abs(m0 - m1) <= 4 * se + 1e-6
    assert ok.all(axis=1).mean() >= 0.99


def test_render_rays_equals_the_full_render(room):
    """orc_render_rays (the CPU baseline's stratified sample) computes each
    listed camera ray's direction alone and traces its paths: the result on
    those rays equals the full render's bit for bit, pinhole and
    equirectangular, and unlisted rays stay zero."""
    s = O.Scene(room)
    for kw in [dict(view=(5, 5, -6), up=(0, 1, 0), right=(-1, 0, 0), width=1.0, height=1.0,
                    sample_x=24, sample_y=20, supersample_x=2, supersample_y=2),
               dict(view=(5, 5, 5), up=(0, 1, 0), right=(3, 0, 0), width=1.0, height=-1.0,
                    sample_x=24, sample_y=12, supersample_x=2, supersample_y=2)]:
        v = O.make_view(**kw, sample=3, depth=5, seed=17, background=(0.1, 0.2, 0.3))
        full, _, _ = s.render(v)
        R = len(full)
        pick = np.random.default_rng(1).choice(R, R // 3, replace=False)
        part, n = s.render_rays(v, pick)
        assert n >= len(pick)
        assert np.array_equal(part[pick].view(np.int32), full[pick].view(np.int32))
        rest = np.setdiff1d(np.arange(R), pick)
        assert (part[rest] == 0).all()


def test_randf3_distribution():
    """randf3 (kernel.cl:41-47) with alpha, beta = randf() * pi:
    (cos b cos a, cos b sin a, sin b).  Its observable distribution -- all
    the render can depend on -- is pinned here: z = sin(beta) has
    P(z <= s) = (2 / pi) asin(s) on [0, 1]; the azimuth atan2(y, x) is
    uniform on (-pi, pi] (cos b is symmetric about 0, so the upper
    semicircle of (cos a, sin a) fills the circle) and independent of the
    radius |cos b| = sqrt(1 - z^2).  Which of alpha and beta is drawn first,
    and whether x takes cos a or sin a, leave this distribution unchanged
    (alpha and beta are independent and identically distributed), so no
    image can tell them apart.  Kolmogorov-Smirnov at the 1e-4 level on
    200 000 draws; a z of cos(beta) or of a plain randf(), or an azimuth
    confined to a half plane, fails."""
    from scipy import stats
    v, _ = O.rand_sequence(2024, 200_000, "randf3")
    v = np.asarray(v, np.float64)
    z = v[:, 2]
    assert stats.kstest(z, lambda s: 2 / np.pi * np.arcsin(np.clip(s, 0, 1))).pvalue > 1e-4
    phi = np.arctan2(v[:, 1], v[:, 0])
    assert stats.kstest(phi, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 1e-4
    # azimuth independent of the radius: same law in the inner and outer halves
    r = np.hypot(v[:, 0], v[:, 1])
    inner = r < np.median(r)
    assert stats.ks_2samp(phi[inner], phi[~inner]).pvalue > 1e-4
    # the tests have power: plausible mistakes fail them
    u = np.random.default_rng(0).uniform(0, np.pi, (2, 200_000))
    assert stats.kstest(np.cos(u[1]), lambda s: 2 / np.pi * np.arcsin(np.clip(s, 0, 1))).pvalue < 1e-4
    assert stats.kstest(u[1] / np.pi, lambda s: 2 / np.pi * np.arcsin(np.clip(s, 0, 1))).pvalue < 1e-4
    half = np.arctan2(np.sin(u[0]), np.cos(u[0]))      # |cos b| instead of cos b: a half plane
    assert stats.kstest(half, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue < 1e-4
