"""Writes reference.npz: the room scene (seed 0) rendered by the float64
oracle (oracle/oracle64.c) at 64x64, 16384 samples per pixel, depth 5, with
per-pixel means and standard errors.  Run from the repository root:

    python tests/golden/room_reference/make_reference.py
"""
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402
from hptracing_amd import scenes  # noqa: E402
from hptracing_amd.scene import KDTree  # noqa: E402

SPP, W, H, DEPTH, SEED = 16384, 64, 64, 5, 20240613


def main():
    with tempfile.TemporaryDirectory() as d:
        p = scenes.write_room(d, 0)
        kd = KDTree(p, d + "/")
    v = O.make_view(view=scenes.ROOM_VIEW["view_p"], up=scenes.ROOM_VIEW["top_dir"],
                    right=scenes.ROOM_VIEW["right_dir"], width=1.0, height=1.0, sample_x=W,
                    sample_y=H, sample=SPP, depth=DEPTH, seed=SEED)
    t0 = time.time()
    s, s2 = O.render64(O.Scene(kd), v)
    mean = s / SPP
    var = np.maximum(s2 / SPP - mean ** 2, 0.0) * SPP / (SPP - 1)
    se = np.sqrt(var / SPP)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference.npz")
    np.savez_compressed(out, mean=mean, se=se, std=np.sqrt(var), spp=SPP, width=W, height=H,
                        depth=DEPTH, seed=SEED)
    print(f"wrote {out} in {time.time() - t0:.1f}s; mean {mean.mean(0)}")


if __name__ == "__main__":
    main()
