"""Parity at the batching and serving workloads' own shapes.

Batching: the 64 uneven views of the 1080p render's scene (721 316
triangles; 128..1024 x 128..768 pixels, 4..64 spp, depth 8) in one
hpt_batch_render, each view's accumulator and image equal to its solo
hpt_render bit for bit -- the per-view result that must match is
TraceRunner::run's (trace_runner.cpp:116-356) -- and three of the views at
reduced pixel counts, path by path against the oracle.

Serving: the 40-request mix (previews, 1280x720 views, equirectangular
panoramas) on the 2.2M-triangle scene, submitted with staggered arrivals so
that requests join while others run; every result equal to its solo
hpt_render bit for bit."""
from __future__ import annotations

import os
import tempfile
import threading
import time

import numpy as np
import pytest

from hptracing_amd import scenes, serving

pytestmark = pytest.mark.gpu

SCENE_DIR = os.path.join(tempfile.gettempdir(), "hpt_scenes")


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))


@pytest.fixture(scope="module")
def full_1080p():
    from hptracing_amd.scene import KDTree
    path = scenes.cached(SCENE_DIR, scenes.write_full, 0, water_side=500)
    return KDTree(path, os.path.dirname(path) + "/")


@pytest.fixture(scope="module")
def serving_scene():
    from hptracing_amd.scene import KDTree
    path = scenes.cached(SCENE_DIR, scenes.write_full, 0, water_side=1000)
    return KDTree(path, os.path.dirname(path) + "/")


def test_batching_workload_each_view_equals_its_solo_render(full_1080p):
    from hptracing_amd.runner import TraceRunner
    r = TraceRunner(full_1080p)
    views = serving.batch_views(64, seed=7, depth=8)
    assert len({(v.sample_x, v.sample_y, v.sample) for v in views}) == 64
    print(f"\n[workload] batch of {len(views)} views, "
          f"{sum(serving.camera_samples(v) for v in views)} camera samples", flush=True)
    imgs, accs = r.render_batch(views, return_accum=True)
    batch_rays = int(r.last_stats.rays_cast)
    solo_rays = 0
    for k, v in enumerate(views):
        img, acc = r.render(v, return_accum=True)
        solo_rays += int(r.last_stats.rays_cast)
        assert _same(acc, accs[k]), f"view {k} ({v.sample_x}x{v.sample_y}, {v.sample} spp)"
        assert np.array_equal(img, imgs[k]), f"view {k}"
    # the same paths were traced: the batch's ray count is the views' sum
    assert batch_rays == solo_rays


def test_batching_workload_views_reduced_match_oracle_paths(full_1080p, oracle_mod):
    """Three of the workload's views (their cameras, spp capped at 4) at
    48x36 pixels, batched: sample for sample against the oracle's
    TraceRunner::run restatement with the same per-path seeds."""
    from hptracing_amd.runner import TraceRunner, make_view
    r = TraceRunner(full_1080p)
    src = serving.batch_views(64, seed=7, depth=8)
    picks = [src[0], src[21], src[42]]
    views, oviews = [], []
    for v in picks:
        kw = dict(view_p=tuple(v.view), top_dir=tuple(v.up), right_dir=tuple(v.right),
                  width=48 / 36, height=1.0, sample_x=48, sample_y=36, supersample_x=1,
                  supersample_y=1, background_color=(0.0, 0.0, 0.0), sample=min(4, v.sample),
                  depth=8, disable_diffuse=0, brightness=1.0, seed=int(v.seed))
        views.append(make_view(**kw))
        oviews.append(oracle_mod.make_view(
            view=kw["view_p"], up=kw["top_dir"], right=kw["right_dir"], width=kw["width"],
            height=1.0, sample_x=48, sample_y=36, background=(0.0, 0.0, 0.0),
            sample=kw["sample"], depth=8, seed=kw["seed"]))
    _, accs = r.render_batch(views, return_accum=True)
    s = oracle_mod.Scene(full_1080p)
    for k, ov in enumerate(oviews):
        ref_acc, _, _ = s.render(ov)
        close = np.isclose(accs[k], ref_acc, rtol=1e-3, atol=1e-4).all(axis=1)
        print(f"[workload] reduced view {k}: {close.mean():.4f} of samples agree", flush=True)
        assert close.mean() > 0.99


def test_serving_workload_each_request_equals_its_solo_render(serving_scene):
    from hptracing_amd.runner import RenderServer, TraceRunner
    r = TraceRunner(serving_scene)
    views = serving.request_views(40, seed=3, depth=8)
    kinds = {(v.sample_x, v.sample_y, v.height < 0) for v in views}
    assert any(vr for *_, vr in kinds) and len(kinds) >= 3
    solo = []
    for v in views:
        solo.append(r.render(v, return_accum=True))
    srv = RenderServer(r, device_budget_bytes=48 << 30)
    ids = [0] * len(views)
    gaps = np.random.default_rng(5).exponential(1.0 / 40.0, len(views))

    def producer():
        for i, v in enumerate(views):
            time.sleep(gaps[i])
            ids[i] = srv.submit(v, return_accum=True)

    th = threading.Thread(target=producer)
    th.start()
    results = []
    try:
        for i in range(len(views)):
            while ids[i] == 0:
                time.sleep(0.001)
            results.append(srv.wait(ids[i], 300.0))
        th.join()
        st = srv.stats()
    finally:
        th.join()
        srv.close()
    print(f"\n[workload] served {st['completed']} requests in {st['macro_passes']} macro passes",
          flush=True)
    for k, ((img0, acc0), (img, acc, _)) in enumerate(zip(solo, results)):
        assert _same(acc0, acc), f"request {k}"
        assert np.array_equal(img0, img), f"request {k}"
    # requests shared macro passes
    assert st["macro_passes"] < sum((v.sample + 7) // 8 for v in views)
