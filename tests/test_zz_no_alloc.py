"""hpt_cast_noalloc, the allocation-free variant of hpt_cast (its work
counter and sort buffers live in a static device array), against hpt_cast
through the C ABI: two calls in a row, the second smaller, on the sorted
and the unsorted path, bit for bit; and batches beyond its limit refused."""
from __future__ import annotations

import numpy as np
import pytest

from hptracing_amd import _native as N
from hptracing_amd import scenes

pytestmark = pytest.mark.gpu


def _cast(fn, runner, org_t, dir_t):
    import torch
    n = org_t.shape[0]
    ids = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    t = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    N.check_device(fn(runner.handle, org_t.data_ptr(), dir_t.data_ptr(), n, ids.data_ptr(),
                      t.data_ptr(), stream))
    torch.cuda.synchronize()
    return ids.cpu().numpy(), t.cpu().numpy()


@pytest.mark.parametrize("sort_rays,sizes", [(1, (400_000, 100_000)), (0, (300_000, 5_000)),
                                             (1, (N.HPT_NOALLOC_MAX_RAYS, 70_000))],
                         ids=["sorted", "unsorted", "sorted-max-then-small"])
def test_noalloc_cast_equals_hpt_cast(soup_small, sort_rays, sizes):
    import torch
    from hptracing_amd.runner import TraceRunner
    r = TraceRunner(soup_small)
    r.set_tuning(sort_rays=sort_rays)
    dev = N.device()
    for k, n in enumerate(sizes):
        org, d = scenes.random_rays(n, (0, 0, 0), (1, 1, 1), seed=60 + k)
        org_t = torch.as_tensor(org).cuda()
        dir_t = torch.as_tensor(d).cuda()
        print(f"[noalloc] sort_rays={sort_rays} call {k}: {n} rays", flush=True)
        ids_a, t_a = _cast(dev.hpt_cast, r, org_t, dir_t)
        ids_b, t_b = _cast(dev.hpt_cast_noalloc, r, org_t, dir_t)
        assert (ids_a >= 0).mean() > 0.05
        assert np.array_equal(ids_a, ids_b)
        assert np.array_equal(t_a.view(np.int32), t_b.view(np.int32))


def test_noalloc_cast_refuses_batches_beyond_its_scratch(room):
    import torch
    from hptracing_amd.runner import TraceRunner
    r = TraceRunner(room)
    dev = N.device()
    n = N.HPT_NOALLOC_MAX_RAYS + 1
    org_t = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    dir_t = torch.ones((n, 3), dtype=torch.float32, device="cuda")
    ids = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    t = torch.full((n,), -7.0, dtype=torch.float32, device="cuda")
    rc = dev.hpt_cast_noalloc(r.handle, org_t.data_ptr(), dir_t.data_ptr(), n, ids.data_ptr(),
                              t.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == N.HPT_ERR_INVALID
    assert b"HPT_NOALLOC_MAX_RAYS" in dev.hpt_last_error()
    assert (ids == -7).all() and (t == -7.0).all()   # nothing launched
    assert dev.hpt_cast_noalloc(r.handle, None, None, 0, None, None, None) == N.HPT_OK
    assert dev.hpt_cast_noalloc(None, None, None, 1, None, None, None) == N.HPT_ERR_INVALID


def test_noalloc_null_stream_and_counter_reuse(room):
    """Many calls in a row on the null stream: each clears the shared
    counter first, so none depends on how the previous one ended."""
    import torch
    from hptracing_amd.runner import TraceRunner
    r = TraceRunner(room)
    dev = N.device()
    org, d = scenes.random_rays(50_000, (0, 0, -2), (10, 10, 10), seed=3)
    org_t, dir_t = torch.as_tensor(org).cuda(), torch.as_tensor(d).cuda()
    ids0, t0 = _cast(dev.hpt_cast, r, org_t, dir_t)
    for m in (50_000, 1, 64, 12_345, 50_000):
        ids = torch.empty(m, dtype=torch.int32, device="cuda")
        t = torch.empty(m, dtype=torch.float32, device="cuda")
        N.check_device(dev.hpt_cast_noalloc(r.handle, org_t.data_ptr(), dir_t.data_ptr(), m,
                                            ids.data_ptr(), t.data_ptr(), None))
        torch.cuda.synchronize()
        assert np.array_equal(ids.cpu().numpy(), ids0[:m])
        assert np.array_equal(t.cpu().numpy().view(np.int32), t0[:m].view(np.int32))
