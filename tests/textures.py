"""The texture pipeline of scene/base.cpp:98-110 restated in numpy for the
image tests: CImg's normalize(0, 255) and nearest resize of decoded
samples, and a one-triangle scene that loads an image as its texture."""
from __future__ import annotations

import shutil

import numpy as np

from hptracing_amd.scene import KDTree


def cimg_layer(samples):
    """normalize(0, 255) over every channel, then resize(512, 512, 1, 3)
    with nearest neighbour (source index floor(dst * src / dst) on each
    axis, channels included) and truncation to uint8 (CImg.h:19093-19105,
    21063-21138)."""
    f32 = np.float32
    src = samples.astype(f32)
    m, M = f32(src.min()), f32(src.max())
    if m == M:
        norm = np.zeros_like(src)
    elif m != 0 or M != 255:
        norm = ((src - m) / (M - m) * f32(255) + f32(0)).astype(f32)
    else:
        norm = src
    h, w, c = src.shape
    ys, xs = (np.arange(512) * h) // 512, (np.arange(512) * w) // 512
    cs = (np.arange(3) * c) // 3
    return norm[ys][:, xs][:, :, cs].astype(np.uint8)


def load_textured(tmp_path, image_name, src_path):
    shutil.copyfile(src_path, tmp_path / image_name)
    (tmp_path / "m.mtl").write_text(f"newmtl tex\nKd 1 1 1\nmap_Kd {image_name}\n")
    (tmp_path / "s.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\n"
                                    "vt 0 0\nvt 1 0\nvt 0 1\nusemtl tex\nf 1/1 2/2 3/3\n")
    return KDTree(str(tmp_path / "s.obj"), str(tmp_path) + "/")
