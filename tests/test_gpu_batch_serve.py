"""Multi-view batching and the continuous-batching server on the GPU: every
view's result must equal its stand-alone render bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from hptracing_amd import _native as N
from hptracing_amd import scenes

pytestmark = pytest.mark.gpu


def _view(**kw):
    from hptracing_amd.runner import make_view
    v = dict(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=40, sample_y=30,
             sample=3, depth=5, seed=1, background_color=(0.05, 0.1, 0.2))
    v.update(kw)
    return make_view(**v)


def _views():
    return [
        _view(sample_x=48, sample_y=32, sample=4, seed=11),
        _view(sample_x=20, sample_y=20, sample=1, seed=12, depth=3),
        _view(sample_x=32, sample_y=16, sample=2, seed=13, height=-1.0,
              right_dir=(3.0, 0.0, 0.0), view_p=(5.0, 5.0, 5.0)),
        _view(sample_x=16, sample_y=16, sample=5, seed=14, supersample_x=2, supersample_y=2),
    ]


@pytest.fixture(scope="module")
def runner(room):
    from hptracing_amd.runner import TraceRunner
    return TraceRunner(room)


def _solo(runner, views):
    return [runner.render(v, return_accum=True) for v in views]


def test_batch_equals_solo(runner):
    views = _views()
    solo = _solo(runner, views)
    imgs, accs = runner.render_batch(views, return_accum=True)
    for (img0, acc0), img, acc in zip(solo, imgs, accs):
        assert np.array_equal(acc0.view(np.int32), acc.view(np.int32))
        assert np.array_equal(img0, img)


def _edge_views():
    return [
        _view(sample_x=1, sample_y=1, sample=1, seed=21),
        _view(sample_x=24, sample_y=8, sample=2, seed=22, depth=1),
        _view(sample_x=16, sample_y=16, sample=2, seed=23, right_dir=(1.0, 0.0, 0.0)),
        _view(sample_x=9, sample_y=7, sample=33, seed=24, depth=3),
        _view(sample_x=40, sample_y=30, sample=17, seed=25),
    ]


def test_batch_edge_views_equal_solo(runner):
    """A batch mixing one pixel, depth 1, a view whose rays all miss, and
    sample counts that end in different launch groups (33 and 17 against
    1 and 2): each view equals its stand-alone render bit for bit, and so
    does every request of the same views on the server."""
    from hptracing_amd.runner import RenderServer
    views = _edge_views()
    solo = _solo(runner, views)
    imgs, accs = runner.render_batch(views, return_accum=True)
    for (img0, acc0), img, acc in zip(solo, imgs, accs):
        assert np.array_equal(acc0.view(np.int32), acc.view(np.int32))
        assert np.array_equal(img0, img)
    srv = RenderServer(runner, device_budget_bytes=64 << 20)
    try:
        ids = [srv.submit(v, return_accum=True) for v in views]
        for (img0, acc0), rid in zip(solo, ids):
            img, acc, _ = srv.wait(rid)
            assert np.array_equal(acc0.view(np.int32), acc.view(np.int32))
            assert np.array_equal(img0, img)
    finally:
        srv.close()


def test_graph_and_direct_launches_agree(runner):
    views = _views()
    a = runner.render_batch(views, return_accum=True, use_graph=True)[1]
    b = runner.render_batch(views, return_accum=True, use_graph=False)[1]
    for x, y in zip(a, b):
        assert np.array_equal(x.view(np.int32), y.view(np.int32))


def test_server_results_equal_solo_and_overlap(runner):
    from hptracing_amd.runner import RenderServer
    views = _views() * 2
    solo = _solo(runner, views)
    srv = RenderServer(runner, device_budget_bytes=64 << 20)
    ids = [srv.submit(v, return_accum=True) for v in views]
    results = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    for (img0, acc0), (img, acc, info) in zip(solo, results):
        assert np.array_equal(acc0.view(np.int32), acc.view(np.int32))
        assert np.array_equal(img0, img)
        assert info["latency_ms"] > 0
    assert st["completed"] == len(views)
    # requests shared passes: fewer macro passes than the sum of their spp
    assert st["macro_passes"] < sum(v.sample for v in views)


def test_server_budget_and_timeout(runner):
    from hptracing_amd.runner import RenderServer
    srv = RenderServer(runner, device_budget_bytes=8 << 20)  # ~7k camera rays
    with pytest.raises(N.HptError) as e:
        srv.submit(_view(sample_x=200, sample_y=200))
    assert e.value.code == N.HPT_ERR_BUDGET
    rid = srv.submit(_view(sample_x=64, sample_y=64, sample=8))
    info = N.RequestInfo()
    rc = N.device().hpt_server_wait(srv._h, rid, 0.0, info)
    assert rc in (N.HPT_OK, N.HPT_ERR_TIMEOUT)
    if rc == N.HPT_ERR_TIMEOUT:
        srv.wait(rid)
    srv.close()


@pytest.mark.parametrize("vr", [False, True])
def test_progressive_request_equals_render_at_each_pass_count(runner, vr):
    """A progressive request's image after S passes is hpt_render's image of
    the same view with sample = S, bit for bit -- including the background
    of camera rays that miss (the equirectangular view looks out of the
    room's open side) -- and its last image is the request's result."""
    from hptracing_amd.runner import RenderServer
    kw = dict(sample_x=40, sample_y=24, sample=20, seed=21)
    if vr:
        kw.update(height=-1.0, right_dir=(3.0, 0.0, 0.0), view_p=(5.0, 5.0, 5.0))
    v = _view(**kw)
    srv = RenderServer(runner, device_budget_bytes=64 << 20)
    rid = srv.submit(v, progressive=True)
    seen, passes = {}, 0
    while passes < v.sample:
        img, passes = srv.peek(rid, min_passes=passes + 1)
        seen[passes] = img
    final, _, _ = srv.wait(rid)
    srv.close()
    assert set(seen) <= {8, 16, 20} and 20 in seen
    assert np.array_equal(seen[20], final)
    for s, img in seen.items():
        ref = runner.render(_view(**{**kw, "sample": s}))
        assert np.array_equal(img, ref), f"image after {s} passes"
    if vr:
        bg = np.round(255 * (1 - np.exp(-np.array([0.05, 0.1, 0.2])))).astype(np.uint8)
        assert (final[..., :3] == bg).all(axis=-1).any(), "no background pixel"


def test_peek_rules(runner):
    from hptracing_amd.runner import RenderServer
    srv = RenderServer(runner, device_budget_bytes=64 << 20)
    plain = srv.submit(_view(sample=2))
    with pytest.raises(N.HptError):
        srv.peek(plain)
    rid = srv.submit(_view(sample=12, seed=5), progressive=True)
    img, passes = srv.peek(rid, min_passes=12)
    assert passes == 12 and img is not None
    srv.wait(plain)
    srv.wait(rid)
    srv.close()


def test_preview_session_frames(runner):
    """The GUI counterpart: frames refine until the full sample count, the
    last equals hpt_render of the session's view; a key press changes the
    view and the next frames show it."""
    from hptracing_amd.preview import PreviewSession
    from hptracing_amd.runner import RenderServer
    srv = RenderServer(runner, device_budget_bytes=64 << 20)
    r = scenes.ROOM_VIEW
    s = PreviewSession(srv, view=r["view_p"], up=r["top_dir"], right=r["right_dir"],
                       width=32, height=24, sample=17, depth=4, seed=3)
    frames = list(s.frames())
    assert [p for _, p in frames][-1] == 17
    assert all(a < b for a, b in zip([p for _, p in frames], [p for _, p in frames][1:]))
    assert np.array_equal(frames[-1][0], runner.render(s.hpt_view()))
    assert s.key("d")
    moved = s.render()
    assert np.array_equal(moved, runner.render(s.hpt_view()))
    assert not np.array_equal(moved, frames[-1][0])
    srv.close()
