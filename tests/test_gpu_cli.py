"""The hptracing command line (hptracing.cpp:213-298 semantics): options,
argument mapping and the y-mirrored image file, against hpt_render."""
from __future__ import annotations

import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from hptracing_amd import scenes
from hptracing_amd._native import LIB_DIR

pytestmark = pytest.mark.gpu

CLI = os.path.join(LIB_DIR, "hptracing")


def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat = 8, b""
    w = h = 0
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 3 * w + 1)
    assert (raw[:, 0] == 0).all()
    return raw[:, 1:].reshape(h, w, 3)


def read_ppm(path):
    data = open(path, "rb").read()
    head = data.split(b"\n", 3)
    w, h = map(int, head[1].split())
    return np.frombuffer(head[3], np.uint8).reshape(h, w, 3)


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
@pytest.mark.parametrize("ext", ["png", "ppm"])
def test_cli_matches_render(tmp_path, room, ext):
    from hptracing_amd.runner import TraceRunner, make_view
    obj = scenes.write_room(str(tmp_path / "scene"), 0)
    out = str(tmp_path / f"out.{ext}")
    cmd = [CLI, "-i", obj, "-o", out, "--width", "40", "--height", "24", "-s", "3", "-d", "4",
           "--view", "5,5,-6", "--right", "-1,0,0", "--up", "0,1,0", "--angle", "1.0",
           "--background", "0.1,0.2,0.3", "--seed", "5", "-x"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    img = read_png(out) if ext == "png" else read_ppm(out)
    # the same request through the library: width = W/H * angle, height = angle
    r = TraceRunner(room)
    v = make_view((5, 5, -6), (0, 1, 0), (-1, 0, 0), 40 / 24 * 1.0, 1.0, 40, 24, 2, 2,
                  (0.1, 0.2, 0.3), 3, 4, 0, 1.0, seed=5)
    rgba = r.render(v)
    assert np.array_equal(img, rgba[::-1, :, :3])


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_errors(tmp_path):
    res = subprocess.run([CLI, "-i", str(tmp_path / "missing.obj"), "-o", "x.png"],
                         capture_output=True, text=True, timeout=60)
    assert res.returncode == 1 and "missing.obj" in res.stderr
    res = subprocess.run([CLI, "-i", "scene.obj"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 2


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_progressive_frames(tmp_path, room):
    """--progressive writes one file per published pass count; the frame of S
    passes is hpt_render's image of the same view with S samples."""
    from hptracing_amd.runner import TraceRunner, make_view
    obj = scenes.write_room(str(tmp_path / "scene"), 0)
    prefix = str(tmp_path / "frame_")
    out = str(tmp_path / "final.png")
    cmd = [CLI, "-i", obj, "-o", out, "--progressive", prefix, "--width", "32", "--height",
           "20", "-s", "20", "-d", "4", "--view", "5,5,-6", "--right", "-1,0,0", "--up",
           "0,1,0", "--seed", "3"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    frames = sorted(int(f[len("frame_"):-4]) for f in os.listdir(tmp_path)
                    if f.startswith("frame_"))
    assert frames and frames[-1] == 20 and frames == sorted(set(frames))
    r = TraceRunner(room)
    for s in frames:
        v = make_view((5, 5, -6), (0, 1, 0), (-1, 0, 0), 32 / 20, 1.0, 32, 20, 1, 1,
                      (0, 0, 0), s, 4, 0, 1.0, seed=3)
        img = read_png(f"{prefix}{s}.png")
        assert np.array_equal(img, r.render(v)[::-1, :, :3]), s
    assert np.array_equal(read_png(out), read_png(f"{prefix}20.png"))


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_tune_same_image(tmp_path):
    """--tune changes launch forms only: the file is the untuned one's, byte
    for byte."""
    obj = scenes.write_room(str(tmp_path / "scene"), 0)
    base = [CLI, "-i", obj, "--width", "40", "--height", "24", "-s", "3", "-d", "4",
            "--view", "5,5,-6", "--right", "-1,0,0", "--seed", "9"]
    a, b = str(tmp_path / "a.ppm"), str(tmp_path / "b.ppm")
    res = subprocess.run(base + ["-o", a], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    res = subprocess.run(base + ["-o", b, "--tune"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    assert "Tuned:" in res.stderr
    assert open(a, "rb").read() == open(b, "rb").read()
