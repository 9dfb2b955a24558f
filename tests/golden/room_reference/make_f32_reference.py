"""Writes reference_f32.npz: the same room view as reference.npz rendered by
the float32 oracle (oracle/oracle.c, the reference algorithm's arithmetic;
dot/cross of the intersection test fused as ROCm's OpenCL builtins fuse
them, the oracle's and the library's default) at 16384 samples per pixel,
with its RMSE against the float64 render.

Float32 and float64 renders differ deterministically on pixels whose camera
ray meets an edge shared by two surfaces (the room's corner lines): the two
precisions assign the pixel to different surfaces (and with the fused forms
a few camera rays slip through the crack of a quad's diagonal, as the
reference kernel's do), and since the primary hit is the same for every
pass, no number of samples removes it.  So the per-pixel
4-standard-error test at equal spp is made against this float32 render, and
the float64 render is the RMSE yardstick.  Run from the repository root:

    python tests/golden/room_reference/make_f32_reference.py
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

SEED = 99


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    ref = np.load(os.path.join(here, "reference.npz"))
    spp = int(ref["spp"])
    with tempfile.TemporaryDirectory() as d:
        p = scenes.write_room(d, 0)
        kd = KDTree(p, d + "/")
    v = O.make_view(view=scenes.ROOM_VIEW["view_p"], up=scenes.ROOM_VIEW["top_dir"],
                    right=scenes.ROOM_VIEW["right_dir"], width=1.0, height=1.0,
                    sample_x=int(ref["width"]), sample_y=int(ref["height"]), sample=spp,
                    depth=int(ref["depth"]), seed=SEED)
    t0 = time.time()
    O.set_arithmetic(O.ARITH_ROCM_OPENCL)
    acc, _, _ = O.Scene(kd).render(v)
    mean = acc.astype(np.float64) / spp
    rmse = float(np.sqrt(np.mean((mean - ref["mean"]) ** 2)))
    out = os.path.join(here, "reference_f32.npz")
    np.savez_compressed(out, mean=mean.astype(np.float32), spp=spp, seed=SEED,
                        rmse_vs_f64=rmse, arithmetic=O.ARITH_ROCM_OPENCL)
    print(f"wrote {out} in {time.time() - t0:.1f}s; RMSE vs float64 {rmse:.5f}")


if __name__ == "__main__":
    main()
