"""Images judged statistically against the float64 reference render
(tests/golden/room_reference: room scene, 64x64, 16384 spp, depth 5, made by
oracle/oracle64.c).  At equal spp at least 99% of per-pixel means must fall
within 4 standard errors of the reference, and the RMSE is reported."""
from __future__ import annotations

import os

import numpy as np
import pytest

from hptracing_amd import scenes

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "room_reference",
                   "reference.npz")


from conftest import within_4se


@pytest.mark.parametrize("spp", [256, 16384])
def test_gpu_render_matches_float64_reference(room, spp):
    from hptracing_amd.runner import TraceRunner, make_view
    ref = np.load(REF)
    r = TraceRunner(room)
    v = make_view(**scenes.ROOM_VIEW, width=1.0, height=1.0, sample_x=int(ref["width"]),
                  sample_y=int(ref["height"]), sample=spp, depth=int(ref["depth"]), seed=777)
    _, acc = r.render(v, return_accum=True)
    frac, rmse = within_4se(acc.astype(np.float64) / spp, spp, ref)
    print(f"\n[stats] spp {spp}: {frac:.4f} of pixel channels within 4 SE, RMSE {rmse:.5f} "
          f"(reference mean {ref['mean'].mean():.4f})")
    assert frac >= 0.99
