"""The C++ boundary (include/hptracing.hpp) in use: tests/cpp/boundary_client
builds an hpt::KDTree and an hpt::TraceRunner and calls run() with the
reference CLI's argument mapping (src/hptracing.cpp:74-78: width =
W/H * angle, height = angle, 2x2 supersampling for -x), and its image must
equal hpt_render's for the same view and seed, byte for byte.  The client is
built by __graft_entry__.build() (g++ against the two libraries)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from hptracing_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "cpp", "boundary_client")


def _client(args, timeout=120):
    if not os.path.exists(CLIENT):
        pytest.fail("tests/cpp/boundary_client is not built: run __graft_entry__.build()")
    return subprocess.run([CLIENT, *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


@pytest.fixture(scope="module")
def room_file(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("cpp_room"))
    return scenes.write_room(d, 0), d + "/"


@pytest.mark.parametrize("W,H,angle,ss,spp,depth,no_diffuse,bright,seed,eye,up,right,bg", [
    (96, 72, 1.0, 1, 8, 5, 0, 1.0, 7, (5, 5, -6), (0, 1, 0), (-1, 0, 0), (0, 0, 0)),
    (64, 48, 0.8, 1, 4, 6, 1, 1.5, 9, (5, 4, -5), (0, 1, 0), (-1, 0, 0), (0.1, 0.2, 0.3)),
    (48, 32, 1.0, 0, 6, 5, 0, 1.0, 3, (5, 5, -6), (0, 1, 0), (-1, 0, 0), (0, 0, 0)),
    (64, 32, -1.0, 0, 4, 5, 0, 1.0, 5, (5, 5, 5), (0, 1, 0), (3, 0, 0), (0.2, 0.1, 0.05)),
], ids=["pinhole-x", "no-diffuse-bright", "no-supersample", "equirectangular"])
def test_cpp_client_equals_hpt_render(room_file, tmp_path, W, H, angle, ss, spp, depth,
                                      no_diffuse, bright, seed, eye, up, right, bg):
    from hptracing_amd.runner import TraceRunner, make_view
    from hptracing_amd.scene import KDTree
    obj, base = room_file
    out = str(tmp_path / "img.rgba")
    # -x is 2x2 supersampling; the equirectangular case runs without it
    supersample = 1 if ss else 0
    r = _client([obj, base, out, W, H, angle, supersample, spp, depth, no_diffuse, bright, seed,
                 *eye, *up, *right, *bg])
    assert r.returncode == 0, r.stderr
    img_cpp = np.fromfile(out, np.uint8).reshape(H, W, 4)
    kd = KDTree(obj, base)
    nodes = int(r.stdout.split()[1])
    assert nodes == len(kd.nodes)
    k = 2 if supersample else 1
    v = make_view(eye, up, right, float(W) / float(H) * angle, angle, W, H, k, k, bg, spp, depth,
                  no_diffuse, bright, seed=seed)
    img_py = TraceRunner(kd).render(v)
    assert np.array_equal(img_cpp, img_py)
    assert img_py[..., :3].max() > 0


def test_cpp_client_reports_errors_as_hpt_error(tmp_path):
    r = _client([str(tmp_path / "missing.obj"), str(tmp_path) + "/", str(tmp_path / "o"), 8, 8,
                 1.0, 0, 1, 2, 0, 1.0, 1, 0, 0, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0])
    assert r.returncode == 3
    assert "hpt::Error" in r.stderr
