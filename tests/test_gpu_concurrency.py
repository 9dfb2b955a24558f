"""Concurrent callers of one scene: every launch has a work counter that no
other caller can reach (render and server workspaces own theirs, each
hpt_cast call takes a stream-ordered one), so results do not depend on what
else runs on the device.  Each check is against the same call made alone."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from hptracing_amd import scenes

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.int32), np.asarray(b).view(np.int32))


def test_render_while_other_threads_cast_on_their_streams(room):
    """A 64-spp render on the scene's stream while two threads cast batches
    on streams of their own, 150 casts each (more launches than any fixed
    pool of counters would hold): the image, its accumulator and every cast
    equal the solo calls bit for bit."""
    import torch
    from hptracing_amd.runner import TraceRunner, make_view
    r = TraceRunner(room)
    view = make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=96, sample_y=96,
                     sample=64, depth=5, seed=3)
    solo_img, solo_acc = r.render(view, return_accum=True)
    batches = [scenes.random_rays(20000, (0, 0, -2), (10, 10, 10), seed=70 + k) for k in range(2)]
    solo = []
    for org, d in batches:
        ids, t = r.cast(org, d)
        torch.cuda.synchronize()
        solo.append((ids.cpu().numpy(), t.cpu().numpy()))

    errors = []
    results = {}

    def caster(k):
        try:
            org, d = batches[k]
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                org_t = torch.as_tensor(org).cuda()
                dir_t = torch.as_tensor(d).cuda()
                outs = [r.cast(org_t, dir_t) for _ in range(150)]
                s.synchronize()
            results[k] = [(i.cpu().numpy(), t.cpu().numpy()) for i, t in outs]
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    def renderer():
        try:
            results["render"] = r.render(view, return_accum=True)
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    threads = [threading.Thread(target=renderer)] + \
              [threading.Thread(target=caster, args=(k,)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=300)
    assert not any(th.is_alive() for th in threads), "a thread did not finish"
    assert not errors, errors
    img, acc = results["render"]
    assert np.array_equal(img, solo_img)
    assert _same(acc, solo_acc)
    for k in range(2):
        for ids, t in results[k]:
            assert np.array_equal(ids, solo[k][0])
            assert _same(t, solo[k][1])


def test_server_and_casts_share_a_scene(room):
    """A render server working through requests while the caller casts on
    its own stream: request images equal hpt_render's, casts equal the solo
    cast."""
    import torch
    from hptracing_amd.runner import RenderServer, TraceRunner, make_view
    r = TraceRunner(room)
    views = [make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=48 + 8 * k,
                       sample_y=40, sample=24, depth=5, seed=11 + k) for k in range(3)]
    solo_imgs = [r.render(v) for v in views]
    org, d = scenes.random_rays(20000, (0, 0, -2), (10, 10, 10), seed=90)
    ids0, t0 = r.cast(org, d)
    torch.cuda.synchronize()
    ids0, t0 = ids0.cpu().numpy(), t0.cpu().numpy()
    srv = RenderServer(r, device_budget_bytes=1 << 30)
    try:
        rids = [srv.submit(v) for v in views]
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            org_t, dir_t = torch.as_tensor(org).cuda(), torch.as_tensor(d).cuda()
            outs = [r.cast(org_t, dir_t) for _ in range(100)]
            s.synchronize()
        imgs = [srv.wait(rid, 120.0)[0] for rid in rids]
    finally:
        srv.close()
    for img, solo in zip(imgs, solo_imgs):
        assert np.array_equal(img, solo)
    for ids, t in outs:
        assert np.array_equal(ids.cpu().numpy(), ids0)
        assert _same(t.cpu().numpy(), t0)


def test_tuning_changes_while_casts_and_a_server_run(room):
    """The launch settings change (fetch forms, grids, node steps, refill,
    sort_rays) from one thread while another casts and a server renders:
    every launch reads a consistent copy of the settings, and since no
    setting changes a result, every cast equals the solo cast and every
    request image equals hpt_render's."""
    import torch
    from hptracing_amd.runner import RenderServer, TraceRunner, make_view
    r = TraceRunner(room)
    base = r.get_tuning()
    views = [make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=40 + 8 * k,
                       sample_y=40, sample=24, depth=5, seed=21 + k) for k in range(3)]
    solo_imgs = [r.render(v) for v in views]
    org, d = scenes.random_rays(70000, (0, 0, -2), (10, 10, 10), seed=91)  # sortable batch
    ids0, t0 = r.cast(org, d)
    torch.cuda.synchronize()
    ids0, t0 = ids0.cpu().numpy(), t0.cpu().numpy()
    forms = [dict(fetch=f, cast_blocks_per_cu=b, node_steps=ns, refill_threshold=rf, sort_rays=sr)
             for f, b, ns, rf, sr in [(3517, 6, 24, 40, 0), (1469, 6, 24, 48, 1), (157, 5, 16, 32, 1), (220, 4, 64, 64, 1),
                                      (5, 8, 8, 16, 0), (92, 5, 32, 48, 1)]]
    stop = threading.Event()
    errors = []

    def tuner():
        k = 0
        try:
            while not stop.is_set():
                r.set_tuning(**forms[k % len(forms)])
                k += 1
                time.sleep(0.002)
        except Exception as e:  # noqa: BLE001
            errors.append(repr(e))

    srv = RenderServer(r, device_budget_bytes=1 << 30)
    th = threading.Thread(target=tuner)
    th.start()
    try:
        rids = [srv.submit(v) for v in views]
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            org_t, dir_t = torch.as_tensor(org).cuda(), torch.as_tensor(d).cuda()
            outs = [r.cast(org_t, dir_t) for _ in range(60)]
            s.synchronize()
        imgs = [srv.wait(rid, 120.0)[0] for rid in rids]
    finally:
        stop.set()
        th.join(timeout=60)
        srv.close()
        r.set_tuning(**base)
    assert not errors, errors
    for img, solo in zip(imgs, solo_imgs):
        assert np.array_equal(img, solo)
    for ids, t in outs:
        assert np.array_equal(ids.cpu().numpy(), ids0)
        assert _same(t.cpu().numpy(), t0)
