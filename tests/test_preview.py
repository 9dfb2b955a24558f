"""Camera state and key bindings of the progressive preview client
(hptracing_amd/preview.py), the headless counterpart of the reference's GUI
(src/hptracing.cpp:124-211).  CPU only: no server is contacted."""
from __future__ import annotations

import numpy as np

from hptracing_amd.preview import PreviewSession


def _session(**kw):
    return PreviewSession(None, **kw)


def test_rotation_about_up_keeps_up_and_turns_right():
    s = _session()
    assert s.key("d")
    np.testing.assert_array_equal(s.up_dir, np.float32([0, 1, 0]))
    # 15 degrees about up: right = (cos 15, 0, sin 15) up to the sign convention
    assert abs(abs(float(s.right_dir[2])) - np.sin(np.radians(15))) < 1e-6
    assert abs(float(s.right_dir[0]) - np.cos(np.radians(15))) < 1e-6
    # 24 presses of 15 degrees come back round, the frame stays orthonormal
    s2 = _session()
    for _ in range(24):
        s2.key("a")
    np.testing.assert_allclose(s2.right_dir, [1, 0, 0], atol=1e-5)
    assert abs(float(np.dot(s2.up_dir, s2.right_dir))) < 1e-5


def test_opposite_keys_undo_each_other():
    for a, b in (("w", "s"), ("q", "e"), ("a", "d")):
        s = _session(up=(0.0, 1.0, 0.0), right=(0.6, 0.0, 0.8))
        s.key(a)
        s.key(b)
        np.testing.assert_allclose(s.up_dir, [0, 1, 0], atol=1e-6)
        np.testing.assert_allclose(s.right_dir, [0.6, 0, 0.8], atol=1e-6)


def test_front_rotation_turns_both_up_and_right():
    s = _session()
    s.key("e")
    assert abs(float(s.up_dir[0])) > 0.2 and abs(float(s.right_dir[1])) > 0.2
    # front = up x right is the rotation axis and does not move
    np.testing.assert_allclose(np.cross(s.up_dir, s.right_dir), [0, 0, -1], atol=1e-6)


def test_translation_and_step():
    s = _session(view=(1.0, 2.0, 3.0))
    s.key("k")  # along front = up x right = (0, 0, -1)
    np.testing.assert_allclose(s.view_point, [1, 2, 2])
    s.key("[")  # halve the step: no re-render
    assert not s.key("]") and s.translate_step == np.float32(1.0)
    s.key("[")
    s.key("up")
    np.testing.assert_allclose(s.view_point, [1, 2.5, 2])
    s.key("left")
    np.testing.assert_allclose(s.view_point, [0.5, 2.5, 2])


def test_angle_keys_and_view_arguments():
    s = _session(width=640, height=480, angle=1.0)
    s.key("x")
    s.key("+")
    assert abs(float(s.angle) - 1.2) < 1e-6
    s.key("z")
    s.key("-")
    assert abs(float(s.angle) - 1.0) < 1e-6
    for _ in range(20):  # 'z' stops above 0.1, '-' does not
        s.key("z")
    assert 0.0 < float(s.angle) <= 0.1 + 1e-6
    v = s.hpt_view()
    assert v.sample_x == 640 and v.sample_y == 480
    assert abs(v.width - 640 / 480 * v.height) < 1e-6   # W/H * angle, angle (runit)
    assert v.supersample_x == 1 and v.sample == 10 and v.depth == 6
