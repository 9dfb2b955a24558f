"""Whole renders by the reference's OWN kernels against this library's render.

oracle/build_ref.sh compiles kernel.cl unmodified for gfx950 and
oracle/ref_harness.hip's ref_render drives its set_viewdirs[_vr],
kdtree_intersect, s1_run and s2_*_run kernels through TraceRunner::run's host
protocol (trace_runner.cpp:116-327), so the shading, bounce and compaction
arithmetic is the reference's own, not a restatement.  The two renderers use
different random streams (the reference seeds its per-slot generators with
rand(), as two of its own runs would differ), so images are judged
statistically at equal spp: at least 99 % of pixel-channel means within 4
combined standard errors, as the float64 and float32 oracle images are.
Skipped where the reference build is absent (it needs the reference sources
at build time).  HIP cannot pass an OpenCL image object, so the harness runs
the build whose s1_run takes its texture argument as an int
(ref_kernels_notex.co) and the scenes are untextured (the room), where s1_run
never reaches a texture read."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN, frac_within_4se, within_4se
import ref_abi
from hptracing_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
CO = os.path.join(REF_DIR, "ref_kernels_notex.co")
REF64 = os.path.join(GOLDEN, "room_reference", "reference.npz")

needs_ref = pytest.mark.skipif(not (os.path.exists(CO) and ref_abi.available()),
                               reason="reference kernels not built (oracle/build_ref.sh)")


def ref_render(kd, view, spp, depth, seed, timed=False, primary=False, co=None):
    """results_mem of the reference's kernels for `view` (make_view's
    arguments), the S0 count each level's intersect ran over, with `timed`
    the host wall time of the render loop in seconds, and with `primary`
    the int4 of each camera ray's primary triangle (-1s: a miss)."""
    lib = ref_abi.load()
    sx, sy = view["sample_x"], view["sample_y"]
    ssx, ssy = view.get("supersample_x", 1), view.get("supersample_y", 1)
    R = sx * ssx * sy * ssy
    seeds = np.random.default_rng(seed).integers(0, 2 ** 31 - 1, R).astype(np.int64)
    acc = np.zeros((R, 3), np.float32)
    s0 = np.zeros((spp, depth), np.int32)
    secs = C.c_double(0.0)
    geo = np.zeros((R, 4), np.int32)
    f3 = lambda v: np.ascontiguousarray(v, np.float32)  # noqa: E731
    eye, up, right = f3(view["view_p"]), f3(view["top_dir"]), f3(view["right_dir"])
    bg = f3(view.get("background_color", (0.0, 0.0, 0.0)))
    arr = lambda a: (a.ctypes.data if len(a) else None, len(a))  # noqa: E731
    rc = lib.ref_render((co or CO).encode(), *arr(kd.points), *arr(kd.normals), *arr(kd.texcoords),
                        *arr(kd.geometries), *arr(kd.materials), *arr(kd.lights),
                        *arr(kd.nodes), *arr(kd.leaf_data), eye.ctypes.data, up.ctypes.data,
                        right.ctypes.data, view["width"], view["height"], sx, sy, ssx, ssy,
                        bg.ctypes.data, spp, depth, view.get("disable_diffuse", 0),
                        seeds.ctypes.data, acc.ctypes.data, s0.ctypes.data, C.byref(secs),
                        geo.ctypes.data if primary else None)
    assert rc == 0, lib.ref_last_error()
    out = (acc, s0) + ((secs.value,) if timed else ()) + ((geo,) if primary else ())
    return out


def primary_geometry(kd, view):
    """This library's primary triangle (int4, -1s: a miss) per camera ray of
    a pinhole view: k_camera's float32 arithmetic restated in numpy (each
    operation rounded once, no contraction), cast by hpt_cast."""
    import torch
    from hptracing_amd.runner import TraceRunner
    f = np.float32
    sx = view["sample_x"] * view.get("supersample_x", 1)
    sy = view["sample_y"] * view.get("supersample_y", 1)
    up, right = np.array(view["top_dir"], f), np.array(view["right_dir"], f)
    w, h = f(view["width"]), f(view["height"])
    nd0 = np.array([up[1] * right[2] - up[2] * right[1], up[2] * right[0] - up[0] * right[2],
                    up[0] * right[1] - up[1] * right[0]], f)
    gy, gx = np.meshgrid(np.arange(sy), np.arange(sx), indexing="ij")
    ty = gy.astype(f) / f(sy) * h - h * f(0.5)
    tx = gx.astype(f) / f(sx) * w - w * f(0.5)
    nd = nd0 + ty[..., None] * up
    nd = (nd + tx[..., None] * right).reshape(-1, 3).astype(f)
    length = np.sqrt((nd[:, 0] * nd[:, 0] + nd[:, 1] * nd[:, 1]) + nd[:, 2] * nd[:, 2])
    d = (nd / length[:, None]).astype(f)
    o = np.tile(np.array(view["view_p"], f), (len(d), 1))
    ids, _ = TraceRunner(kd).cast(o, d)
    torch.cuda.synchronize()
    ids = ids.cpu().numpy()
    return np.where(ids[:, None] >= 0, kd.geometries[np.maximum(ids, 0)], -1)


def live_within_4_sigma(live, ref_live, n_paths):
    """Per bounce level, whether two independent counts of surviving paths
    out of n_paths differ by at most 4 standard deviations of the
    difference (binomial, p = the reference's fraction)."""
    p = ref_live / n_paths
    sd = np.sqrt(2.0 * n_paths * p * (1.0 - p))
    return np.abs(live - ref_live) <= 4.0 * sd + 1.0


def gpu_render(kd, view, spp, depth, seed, rocm_opencl=None):
    """This library's accumulator and live paths per level, in the default
    arithmetic (ROCm OpenCL dot/cross in the cast) unless `rocm_opencl` is
    given."""
    from hptracing_amd.runner import TraceRunner, make_view
    r = TraceRunner(kd)
    if rocm_opencl is not None:
        r.set_arithmetic(rocm_opencl)
    _, acc = r.render(make_view(**view, sample=spp, depth=depth, seed=seed), return_accum=True)
    return acc, np.array(r.last_stats.live_per_depth[:depth], np.int64)


ROOM64 = dict(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=64, sample_y=64)


@needs_ref
def test_reference_kernels_image_matches_gpu(room):
    """64x64 room view, depth 5, 1024 spp each: the reference kernels' image
    against this library's (standard errors from the per-pixel sample
    deviations of the float64 16384-spp render), both against that render,
    and the live paths per bounce level.

    The live counts depend on the last bit of t: a light-branch path that
    starts on the emitter samples a point in the emitter's own plane and is
    kept when the direction's side matches (kernel.cl:527-548), which for a
    hit point exactly in that plane is decided by rounding.  With ROCm's
    OpenCL dot/cross in the cast (HPT_ARITH_ROCM_OPENCL, the arithmetic of
    the reference build) every t, and so every such decision, is the
    reference's, and the counts of the default arithmetic agree within 4
    binomial standard deviations; with the literal forms fewer emitter hit
    points lie off the plane and 0.3 % more paths continue."""
    ref64 = np.load(REF64)
    spp, depth = 1024, int(ref64["depth"])
    acc_ref, s0 = ref_render(room, ROOM64, spp, depth, seed=11)
    acc_gpu, live = gpu_render(room, ROOM64, spp, depth, seed=12)
    mean_ref = acc_ref.astype(np.float64) / spp
    mean_gpu = acc_gpu.astype(np.float64) / spp
    se = np.sqrt(2.0) * ref64["std"] / np.sqrt(spp)
    frac, rmse = frac_within_4se(mean_gpu, mean_ref, se)
    frac_r64, rmse_r64 = within_4se(mean_ref, spp, ref64)
    frac_g64, rmse_g64 = within_4se(mean_gpu, spp, ref64)
    ref_live = s0.sum(axis=0).astype(np.int64)
    acc_lit, live_lit = gpu_render(room, ROOM64, spp, depth, seed=13, rocm_opencl=False)
    frac_lit, _ = frac_within_4se(acc_lit.astype(np.float64) / spp, mean_ref, se)
    print(f"\n[ref-render] {spp} spp: GPU vs reference kernels {frac:.4f} within 4 SE "
          f"(RMSE {rmse:.5f}; with the literal arithmetic {frac_lit:.4f}); against float64: "
          f"reference kernels {frac_r64:.4f} (RMSE {rmse_r64:.5f}), GPU {frac_g64:.4f} "
          f"(RMSE {rmse_g64:.5f}); live paths per level reference {ref_live[1:].tolist()} "
          f"GPU {live[1:].tolist()}, with the literal arithmetic {live_lit[1:].tolist()}")
    assert frac >= 0.99 and frac_lit >= 0.99
    # Against float64 both float32 renderers miss on the same pixels (camera
    # rays meeting a wall-corner edge pick the other wall in float32).
    assert abs(frac_r64 - frac_g64) < 0.005
    n_paths = 64 * 64 * spp
    assert live_within_4_sigma(live[1:], ref_live[1:], n_paths).all()
    rel = np.abs(live_lit[1:] - ref_live[1:]) / ref_live[1:]
    assert (rel < 0.005).all(), rel


@needs_ref
def test_reference_kernels_equirectangular_supersampled(room):
    """The equirectangular camera with 2x2 supersampling (set_viewdirs_vr,
    shift = int(right.x) * ssx): standard errors from the spread of sixteen
    independent GPU renders."""
    view = dict(view_p=(5.0, 5.0, 5.0), top_dir=(0.0, 1.0, 0.0), right_dir=(3.0, 0.0, 0.0),
                width=1.0, height=-1.0, sample_x=32, sample_y=16, supersample_x=2,
                supersample_y=2, background_color=(0.1, 0.2, 0.3))
    depth, spp, k = 5, 512, 16
    acc_ref, _ = ref_render(room, view, spp, depth, seed=21)
    chunks = np.stack([gpu_render(room, view, spp // k, depth, seed=100 + j)[0]
                       for j in range(k)]).astype(np.float64) / (spp // k)
    mean_gpu = chunks.mean(axis=0)
    se_gpu = chunks.std(axis=0, ddof=1) / np.sqrt(k)
    frac, rmse = frac_within_4se(mean_gpu, acc_ref.astype(np.float64) / spp,
                                 np.sqrt(2.0) * se_gpu)
    print(f"\n[ref-render] equirectangular 2x2, {spp} spp: GPU vs reference kernels {frac:.4f} "
          f"within 4 SE (RMSE {rmse:.5f})")
    assert frac >= 0.99


@needs_ref
def test_reference_kernels_no_diffuse_background(room):
    """--no-diffuse (disable_diffuse: the diffuse share goes to light
    sampling, kernel.cl:396-407) with a coloured background, 512 spp.
    Standard errors come from sixteen independent GPU renders; light-sampled
    paths that reach the emitter make pixels heavy-tailed, so the bar is the
    same test between two independent GPU images (the reference kernels'
    image within 0.5 % of it).  Direct light sampling makes pixel means
    nearly noise-free, so the pixels whose camera ray meets a different
    triangle in the two renderers -- rays on a triangle edge, where the
    reference's OpenCL `normalize` in set_viewdirs and this library's
    literal float32 camera round differently (wall corners pick the other wall, quad
    diagonals crack) -- stand out; those whose material or hit differs are
    counted (under 2 %) and left out.  Most paths here sample the light, and those that land on the
    emitter sample again in its own plane, where the side test is decided by
    rounding, so past the first level the live paths agree within 4
    binomial standard deviations plus 1 % (the first level within 4
    standard deviations)."""
    view = dict(ROOM64, background_color=(0.2, 0.1, 0.05), disable_diffuse=1)
    depth, spp, k = 5, 512, 16
    acc_ref, s0, geo_ref = ref_render(room, view, spp, depth, seed=51, primary=True)
    geo_ours = primary_geometry(room, view)
    same_tri = (geo_ref == geo_ours).all(axis=1)
    # a different triangle of the same material and plane (a quad's other
    # half) shades the same; a different material or a miss does not
    same = geo_ref[:, 3] == geo_ours[:, 3]
    chunks = np.stack([gpu_render(room, view, spp // k, depth, seed=400 + j)[0]
                       for j in range(k)]).astype(np.float64) / (spp // k)
    mean_gpu = chunks.mean(axis=0)
    se = np.sqrt(2.0) * chunks.std(axis=0, ddof=1) / np.sqrt(k)
    frac, rmse = frac_within_4se(mean_gpu[same], acc_ref[same].astype(np.float64) / spp, se[same])
    acc_b, _ = gpu_render(room, view, spp, depth, seed=499)
    frac_null, rmse_null = frac_within_4se(mean_gpu[same], acc_b[same].astype(np.float64) / spp,
                                           se[same])
    _, live_ocl = gpu_render(room, view, spp, depth, seed=500)
    ref_live = s0.sum(axis=0).astype(np.int64)
    print(f"\n[ref-render] no-diffuse, background, {spp} spp: primary triangle differs on "
          f"{int((~same_tri).sum())} of {len(same)} camera rays, its material (or hit) on "
          f"{int((~same).sum())}; on the rest GPU vs reference kernels "
          f"{frac:.4f} within 4 SE (RMSE {rmse:.5f}), GPU vs an independent GPU image "
          f"{frac_null:.4f} (RMSE {rmse_null:.5f}); live per level reference "
          f"{ref_live[1:].tolist()} GPU {live_ocl[1:].tolist()}")
    assert (~same).mean() < 0.02
    assert frac >= frac_null - 0.005
    n_paths = 64 * 64 * spp
    assert live_within_4_sigma(live_ocl[1:2], ref_live[1:2], n_paths).all()
    p = ref_live[2:] / n_paths
    allowed = 4.0 * np.sqrt(2.0 * n_paths * p * (1.0 - p)) + 0.01 * ref_live[2:]
    assert (np.abs(live_ocl[2:] - ref_live[2:]) <= allowed).all()


@pytest.fixture(scope="module")
def full_plain():
    """The 1080p workload's scene (721 316 triangles) with the vase
    untextured: the harness passes no texture."""
    import tempfile
    from hptracing_amd.scene import KDTree
    path = scenes.cached(os.path.join(tempfile.gettempdir(), "hpt_scenes"), scenes.write_full, 0,
                         water_side=500, textured=False)
    return KDTree(path, os.path.dirname(path) + "/")


FULL_VIEW = dict(**scenes.ROOM_VIEW, width=1920 / 1080, height=1.0)


@needs_ref
def test_reference_kernels_image_full_scene(full_plain):
    """The 1080p view of the full scene (water refraction, the mirror
    sphere's specular bounces, diffuse and light) at 128x72 with 2x2
    supersampling, depth 8, 256 spp.  Standard errors come from the spread
    of sixteen independent 16-spp GPU renders; specular chains that reach
    the emitter make some pixels heavy-tailed, so the bar is the same test
    between two independent GPU images: the reference kernels' image must
    score as well (within 0.5 %).  Live paths per level within 4 binomial
    standard deviations (the default arithmetic: ROCm's OpenCL dot/cross in
    the cast)."""
    view = dict(**FULL_VIEW, sample_x=128, sample_y=72, supersample_x=2, supersample_y=2)
    depth, spp, k = 8, 256, 16
    acc_ref, s0 = ref_render(full_plain, view, spp, depth, seed=41)
    chunks = np.stack([gpu_render(full_plain, view, spp // k, depth, seed=200 + j)[0]
                       for j in range(k)]).astype(np.float64) / (spp // k)
    mean_gpu = chunks.mean(axis=0)
    se = np.sqrt(2.0) * chunks.std(axis=0, ddof=1) / np.sqrt(k)
    frac, rmse = frac_within_4se(mean_gpu, acc_ref.astype(np.float64) / spp, se)
    acc_b, _ = gpu_render(full_plain, view, spp, depth, seed=299)
    frac_null, rmse_null = frac_within_4se(mean_gpu, acc_b.astype(np.float64) / spp, se)
    _, live_ocl = gpu_render(full_plain, view, spp, depth, seed=300)
    ref_live = s0.sum(axis=0).astype(np.int64)
    print(f"\n[ref-render] full scene 128x72 x4, depth 8, {spp} spp: GPU vs reference kernels "
          f"{frac:.4f} within 4 SE (RMSE {rmse:.5f}); GPU vs an independent GPU image "
          f"{frac_null:.4f} (RMSE {rmse_null:.5f}); live per level reference "
          f"{ref_live[1:].tolist()} GPU {live_ocl[1:].tolist()}")
    assert frac >= frac_null - 0.005
    assert live_within_4_sigma(live_ocl[1:], ref_live[1:], 128 * 72 * 4 * spp).all()


@needs_ref
def test_reference_kernels_speed_1080p(full_plain):
    """The 1080p workload (721 316 triangles, 1920x1080 with 2x2
    supersampling, depth 8) with the vase untextured, 8 spp: the reference's
    kernels under TraceRunner::run's host loop against this library's render
    on the same GPU, rays counted as the reference's intersect launches count
    them (primaries once, then every level's S0).  Prints both rates; the
    live paths per level agree within 1 %."""
    import time
    from hptracing_amd.runner import TraceRunner, make_view
    kd = full_plain
    view = dict(**FULL_VIEW, sample_x=1920, sample_y=1080, supersample_x=2, supersample_y=2)
    spp, depth = 8, 8
    _, s0, ref_s = ref_render(kd, view, spp, depth, seed=31, timed=True)
    R = 1920 * 1080 * 4
    ref_rays = R + int(s0.sum())
    r = TraceRunner(kd)
    r.render(make_view(**view, sample=1, depth=depth, seed=5))  # warm
    t0 = time.perf_counter()
    r.render(make_view(**view, sample=spp, depth=depth, seed=32))
    ours_s = time.perf_counter() - t0
    ours_rays = int(r.last_stats.rays_cast)
    live = np.array(r.last_stats.live_per_depth[:depth], np.int64)
    ref_live = s0.sum(axis=0).astype(np.int64)
    print(f"\n[ref-render] 1080p x4, depth 8, {spp} spp, untextured vase: reference kernels "
          f"{ref_rays} rays in {ref_s:.3f}s = {ref_rays / ref_s / 1e6:.1f} Mrays/s; this library "
          f"{ours_rays} rays in {ours_s:.3f}s = {ours_rays / ours_s / 1e6:.1f} Mrays/s "
          f"({(ours_rays / ours_s) / (ref_rays / ref_s):.1f}x); live per level reference "
          f"{ref_live[1:].tolist()} ours {live[1:].tolist()}")
    rel = np.abs(live[1:] - ref_live[1:]) / ref_live[1:]
    assert (rel < 0.01).all(), rel


CO_CONTRACT = os.path.join(REF_DIR, "ref_kernels_notex_contract.co")


@needs_ref
@pytest.mark.skipif(not os.path.exists(CO_CONTRACT), reason="contracted reference build absent")
def test_reference_kernels_speckle_on_the_simple_cornell_box(cornell):
    """The reference's simple Cornell box (tests/golden/cornell_box, a copy of
    scene/simple-cornell-box) is 556 units wide: near its far walls one
    float32 ulp is 6e-5, above the 1e-5 step that moves a continuation ray
    off its surface (kernel.cl:411, 447-553), so a hit point that rounds to
    the far side of the (slanted) red wall starts its bounces behind the
    wall, and camera rays are the same in every pass: a fixed fraction of
    the wall's camera rays stays dark.  The reference's own kernels and this
    library show the same fraction (the parity claim); the reference's
    kernels built with fused multiply-adds (-ffp-contract=fast) show it as
    well, so contraction is not what makes the published 5000-spp image of
    this scene smooth.  (In a scratch build of the oracle with the step at
    1e-3, camera rays of zero radiance drop from 12.5 % to 2.5 %.)"""
    view = dict(view_p=(278.0, 273.0, -800.0), top_dir=(0.0, 1.0, 0.0), right_dir=(-1.0, 0.0, 0.0),
                width=0.7, height=0.7, sample_x=64, sample_y=64, supersample_x=2,
                supersample_y=2)
    depth, spp = 10, 128
    kd = cornell
    red = [i for i, m in enumerate(kd.materials)
           if m["diffuse"][0] > 0.4 and m["diffuse"][1] < 0.1]
    assert red, "no red material"

    def dark_fraction(acc, geo):
        on = np.isin(geo[:, 3], red)
        lum = acc[on].astype(np.float64).sum(axis=1) / spp
        return float((lum < 0.25 * np.median(lum)).mean()), int(on.sum())

    acc_off, _, geo_off = ref_render(kd, view, spp, depth, seed=61, primary=True)
    acc_fma, _, geo_fma = ref_render(kd, view, spp, depth, seed=62, primary=True, co=CO_CONTRACT)
    acc_gpu, _ = gpu_render(kd, view, spp, depth, seed=63)
    geo_gpu = primary_geometry(kd, view)
    off, n_off = dark_fraction(acc_off, geo_off)
    fma, n_fma = dark_fraction(acc_fma, geo_fma)
    gpu, n_gpu = dark_fraction(acc_gpu, geo_gpu)
    print(f"\n[ref-render] simple Cornell box, {spp} spp: dark red-wall camera rays -- reference "
          f"kernels without contraction {off:.4f} of {n_off}, with fused multiply-adds "
          f"{fma:.4f} of {n_fma}, this library {gpu:.4f} of {n_gpu}")
    sd = np.sqrt(2 * off * (1 - off) / n_off)
    assert off > 0.2
    assert abs(gpu - off) <= 4 * sd + 0.01
    assert abs(fma - off) <= 0.1
