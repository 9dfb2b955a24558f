"""Full renders on the GPU against the oracle's TraceRunner::run restatement
(same per-path seeding), plus image-level statistics."""
from __future__ import annotations

import numpy as np
import pytest

from hptracing_amd import scenes

pytestmark = pytest.mark.gpu


def _views(oracle_mod, **kw):
    from hptracing_amd.runner import make_view
    v = dict(view_p=scenes.ROOM_VIEW["view_p"], top_dir=scenes.ROOM_VIEW["top_dir"],
             right_dir=scenes.ROOM_VIEW["right_dir"], width=1.0, height=1.0, sample_x=64,
             sample_y=64, supersample_x=1, supersample_y=1, background_color=(0.1, 0.2, 0.3),
             sample=4, depth=5, disable_diffuse=0, brightness=1.0, seed=7)
    v.update(kw)
    gv = make_view(**v)
    ov = oracle_mod.make_view(view=v["view_p"], up=v["top_dir"], right=v["right_dir"],
                              width=v["width"], height=v["height"], sample_x=v["sample_x"],
                              sample_y=v["sample_y"], supersample_x=v["supersample_x"],
                              supersample_y=v["supersample_y"], background=v["background_color"],
                              sample=v["sample"], depth=v["depth"],
                              disable_diffuse=v["disable_diffuse"], brightness=v["brightness"],
                              seed=v["seed"])
    return gv, ov


def _compare(room, oracle_mod, **kw):
    from hptracing_amd.runner import TraceRunner
    gv, ov = _views(oracle_mod, **kw)
    r = TraceRunner(room)
    rgba, acc = r.render(gv, return_accum=True)
    ref_acc, ref_rgba, ref_live = oracle_mod.Scene(room).render(ov)
    close = np.isclose(acc, ref_acc, rtol=1e-3, atol=1e-4).all(axis=1)
    frac = close.mean()
    live = np.array(r.last_stats.live_per_depth[: gv.depth])
    return frac, rgba, ref_rgba, live, ref_live, acc, ref_acc


def test_render_matches_oracle_paths(room, oracle_mod):
    frac, rgba, ref_rgba, live, ref_live, acc, ref_acc = _compare(room, oracle_mod)
    # Same per-path random streams: only transcendental-function rounding
    # (device vs host libm) separates the two, so almost every sample agrees.
    assert frac > 0.99, frac
    assert np.abs(live - ref_live).sum() <= 0.01 * ref_live.sum()
    assert (np.abs(rgba.astype(int) - ref_rgba.astype(int)) <= 1).mean() > 0.99


def test_render_supersample_vr(room, oracle_mod):
    frac, *_ = _compare(room, oracle_mod, supersample_x=2, supersample_y=2, sample_x=32,
                        sample_y=32, sample=2)
    assert frac > 0.99, frac
    frac, *_ = _compare(room, oracle_mod, height=-1.0, width=1.0, sample_x=64, sample_y=32,
                        right_dir=(0.0, 0.0, 0.0), view_p=(5.0, 5.0, 5.0), sample=2)
    assert frac > 0.99, frac


@pytest.fixture(scope="module")
def full_small(scene_dir):
    """The 1080p scene's parts at a reduced triangle count: the room, the
    water grid (refraction, IOR 1.33), the textured surface of revolution
    (texture fetch) and the mirror sphere (specular)."""
    import os
    from hptracing_amd.scene import KDTree
    p = scenes.write_full(os.path.join(scene_dir, "full_small"), 0, water_side=60, rev=(60, 40),
                          sphere=(24, 24))
    return KDTree(p, os.path.dirname(p) + "/")


@pytest.mark.parametrize("kw", [
    dict(depth=1),
    dict(sample_x=1, sample_y=1, sample=1),
    dict(sample_x=3, sample_y=5, sample=3, depth=2),
    dict(sample_x=17, sample_y=9, sample=33, depth=3),
    dict(right_dir=(1.0, 0.0, 0.0), sample_x=16, sample_y=16, sample=2),
], ids=["depth-1", "one-pixel", "odd-3x5-depth-2", "33-spp-over-two-launch-groups", "all-miss"])
def test_render_edge_cases(room, oracle_mod, kw):
    """Shapes at the edges of the pass loop: no bounce level, one pixel and
    one sample, odd sizes, a sample count that leaves a one-pass remainder
    after two sixteen-pass launch groups, and a camera facing away from the
    room (every path ends at the first cast; each pixel accumulates the
    background once per sample).  Paths equal the oracle's."""
    frac, rgba, ref_rgba, live, ref_live, acc, ref_acc = _compare(room, oracle_mod, **kw)
    assert frac >= 0.99, frac
    assert np.array_equal(live, np.asarray(ref_live)[: len(live)])
    if "right_dir" in kw:
        assert live[0] == 256 and not live[1:].any()
        bg = np.float32([0.1, 0.2, 0.3]) * np.float32(2)
        assert np.array_equal(acc, np.broadcast_to(bg, acc.shape))


def test_render_full_scene_matches_oracle_paths(full_small, oracle_mod):
    """Every material kind of the 1080p scene (texture fetch, specular,
    refraction, diffuse, light) in the 1080p view with -x and depth 8,
    sample for sample against the oracle."""
    from hptracing_amd.runner import TraceRunner
    kd = full_small
    # The view sees textured, mirror and refracting surfaces.
    W, H = 64, 36
    ys, xs = np.meshgrid((np.arange(H) + 0.5) / H - 0.5, (np.arange(W) + 0.5) / W - 0.5,
                         indexing="ij")
    d = np.stack([-xs.ravel() * (W / H), ys.ravel(), np.ones(W * H)], 1).astype(np.float32)
    o = np.tile(np.array(scenes.ROOM_VIEW["view_p"], np.float32), (W * H, 1))
    ids, _ = TraceRunner(kd).cast(o, d)
    ids = ids.cpu().numpy()
    m = kd.materials[kd.geometries[ids[ids >= 0], 3]]
    assert (m["diffuse_texture_id"] >= 0).mean() > 0.005
    assert (m["specular_possibility"] > 0.5).mean() > 0.01
    assert (m["refract_possibility"] > 0.5).mean() > 0.01
    frac, rgba, ref_rgba, live, ref_live, *_ = _compare(
        kd, oracle_mod, width=W / H, sample_x=W, sample_y=H, supersample_x=2, supersample_y=2,
        sample=8, depth=8, background_color=(0.0, 0.0, 0.0))
    assert frac > 0.99, frac
    assert np.abs(live - ref_live).sum() <= 0.01 * ref_live.sum()
    assert (np.abs(rgba.astype(int) - ref_rgba.astype(int)) <= 1).mean() > 0.99


def test_render_deterministic(room):
    from hptracing_amd.runner import TraceRunner, make_view
    r = TraceRunner(room)
    v = make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=48, sample_y=48,
                  sample=3, depth=5, seed=3)
    a = r.render(v, return_accum=True)[1]
    b = r.render(v, return_accum=True)[1]
    assert np.array_equal(a, b)


def test_autotune_keeps_results(room):
    """The autotuner changes launch forms only: the image is the same bits
    before and after, and the chosen tuning is the one left on the scene."""
    from hptracing_amd.runner import TraceRunner, make_view
    r = TraceRunner(room)
    v = make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=64, sample_y=64, sample=4,
                  depth=5, seed=3)
    _, before = r.render(v, return_accum=True)
    chosen = r.autotune(v, max_rays=4096)
    assert chosen == r.get_tuning()
    _, after = r.render(v, return_accum=True)
    assert np.array_equal(before.view(np.int32), after.view(np.int32))


def test_pass_at_a_time_same_bits(room):
    """Sixteen passes per launch sequence or one at a time (the reference's
    launch structure): the accumulator slices make the sums identical."""
    from hptracing_amd.runner import TraceRunner, make_view
    r = TraceRunner(room)
    v = make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=48, sample_y=40, sample=19,
                  depth=5, seed=11)
    img_a, acc_a = r.render(v, return_accum=True)
    img_b, acc_b = r.render(v, return_accum=True, pass_at_a_time=True)
    assert np.array_equal(acc_a.view(np.int32), acc_b.view(np.int32))
    assert np.array_equal(img_a, img_b)
