"""The ray-cast stage at full size: a 200k-triangle soup, 2^22 rays.

The soup has the ray-cast workload's statistics (centres uniform in the unit
cube, edges of log-uniform length in [1e-3, 1e-1]); at 1M triangles the
reference's KD builder duplicates triangles super-linearly (7.9, 22, 78, 284
leaf entries per triangle at 25k, 50k, 100k, 200k) and does not fit in
memory, so 200k is the largest size it builds.

The oracle cannot walk 4M rays in a test's time, so the full batch is
checked through properties that do not depend on size: every reported t is
reproduced bit for bit by Moller-Trumbore on the reported triangle (float32,
the kernel's operation order with dot/cross fused as in the default
arithmetic, evaluated in numpy by tests/f32ops.py), every miss and hit
flag is consistent, and a seeded random subset of the rays matches the
oracle exactly (ids and t)."""
from __future__ import annotations

import numpy as np
import pytest

from f32ops import moller_trumbore
from hptracing_amd import scenes

pytestmark = pytest.mark.gpu


def test_cast_full_size_properties(soup_200k, oracle_mod):
    import torch
    from hptracing_amd.runner import TraceRunner
    kd = soup_200k
    pts = kd.points[:, :3]
    lo, hi = pts.min(0), pts.max(0)
    org, d = scenes.random_rays(1 << 22, lo, hi, seed=2)
    r = TraceRunner(kd)
    ids, t = r.cast(org, d)
    torch.cuda.synchronize()
    ids, t = ids.cpu().numpy(), t.cpu().numpy()
    hit = ids >= 0
    assert 0.05 < hit.mean() < 0.999
    assert (t[~hit] == np.float32(-42)).all()
    tri = kd.geometries[ids[hit]]
    t_re, _, _ = moller_trumbore(org[hit], d[hit], pts[tri[:, 0]], pts[tri[:, 1]],
                                 pts[tri[:, 2]], fused=True)
    assert np.array_equal(t_re.view(np.int32), t[hit].view(np.int32))
    # a seeded subset against the oracle, bit for bit
    rng = np.random.default_rng(9)
    sub = rng.choice(len(ids), 20000, replace=False)
    ref_ids, ref_t = oracle_mod.Scene(kd).kd_cast(org[sub], d[sub])
    assert np.array_equal(ids[sub], ref_ids)
    assert np.array_equal(t[sub].view(np.int32), ref_t.view(np.int32))
