"""Progressive preview: a headless counterpart of the reference's GLUT window
(src/hptracing.cpp:40-211) as a client of the render server.

The reference keeps one scene resident, moves its camera on key presses
(keyFunc / specKeyFunc, :172-211) and re-renders the whole view with
``TraceRunner::run`` whenever the camera changed (timerFunc, :124-130).
Here the camera state and the key bindings are the same; a frame is a
progressive server request (HPT_SERVE_PROGRESSIVE) whose image is published
after every macro pass, so a caller sees a first image after the first
passes and a refined one after each further macro pass.  The image after S
passes equals ``hpt_render`` of the same view with S samples, bit for bit.
"""
from __future__ import annotations

import math
from typing import Iterator, Optional, Tuple

import numpy as np

from . import _native as N
from .runner import RenderServer, make_view

PI = 3.141592653589  # src/hp/common.h:20


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]], np.float32)


class PreviewSession:
    """Camera state of the reference's GUI plus progressive frames.

    Arguments are the CLI options the GUI starts from (hptracing.cpp:214-272):
    view point, up and right vectors, output size, angle, supersampling,
    background, samples, depth, --no-diffuse and brightness."""

    def __init__(self, server: RenderServer, view=(0.0, 0.0, -500.0), up=(0.0, 1.0, 0.0),
                 right=(1.0, 0.0, 0.0), width=500, height=500, angle=1.0, supersample=False,
                 background=(0.0, 0.0, 0.0), sample=10, depth=6, no_diffuse=False,
                 brightness=1.0, seed=1):
        self.server = server
        self.view_point = np.array(view, np.float32)
        self.up_dir = np.array(up, np.float32)
        self.right_dir = np.array(right, np.float32)
        self.width, self.height = int(width), int(height)
        self.angle = np.float32(angle)
        self.supersample = bool(supersample)
        self.background = tuple(float(c) for c in background)
        self.sample, self.depth = int(sample), int(depth)
        self.no_diffuse = bool(no_diffuse)
        self.brightness = float(brightness)
        self.translate_step = np.float32(1.0)
        self.seed = int(seed)

    # ------------------------------------------------------------ camera
    def _axis(self, ax: int):
        if ax == 2:
            return self.up_dir
        if ax == 3:
            return _cross(self.up_dir, self.right_dir)  # front = up x right
        return self.right_dir

    def rotate(self, alpha_deg: float, ax: int) -> None:
        """rotateView (hptracing.cpp:132-155): rotate the up/right/front
        frame by alpha degrees about right (1), up (2) or front (3), with
        m = I + sin(a) L + (1 - cos a) L^2, L the reference's cross matrix."""
        up, right = self.up_dir.copy(), self.right_dir.copy()
        front = _cross(up, right)
        about = {1: right, 2: up, 3: front}[ax]
        L = np.array([[0, about[2], -about[1]],
                      [-about[2], 0, about[0]],
                      [about[1], -about[0], 0]], np.float32)
        a = alpha_deg / 180.0 * PI
        m = (np.eye(3, dtype=np.float32) + np.float32(math.sin(a)) * L
             + np.float32(1.0 - math.cos(a)) * (L @ L)).astype(np.float32)
        if ax != 2:
            self.up_dir = (m @ up).astype(np.float32)
        if ax != 1:
            self.right_dir = (m @ right).astype(np.float32)

    def translate(self, dist: float, ax: int) -> None:
        """translateView (hptracing.cpp:157-170): move the view point along
        right (1), up (2) or front (3)."""
        self.view_point = (self.view_point + np.float32(dist) * self._axis(ax)).astype(np.float32)

    def key(self, k: str) -> bool:
        """Apply a key of keyFunc / specKeyFunc (hptracing.cpp:172-211);
        returns whether the view must be rendered again (need_rerun).
        Arrow keys are 'up', 'down', 'left', 'right'."""
        rot = {"a": (-15, 2), "d": (15, 2), "w": (-15, 1), "s": (15, 1), "q": (-15, 3),
               "e": (15, 3)}
        if k in rot:
            self.rotate(*rot[k])
            return True
        # 'z' and 'x' step by the double 0.1, '+' and '-' by the float 0.1f
        if k == "z":
            if float(self.angle) > 0.1:
                self.angle = np.float32(float(self.angle) - 0.1)
            return True
        if k == "x":
            self.angle = np.float32(float(self.angle) + 0.1)
            return True
        if k in ("+", "-"):
            step = np.float32(0.1) if k == "+" else -np.float32(0.1)
            self.angle = np.float32(self.angle + step)
            return True
        moves = {"k": 3, "j": 3, "up": 2, "down": 2, "right": 1, "left": 1}
        if k in moves:
            sign = -1.0 if k in ("j", "down", "left") else 1.0
            self.translate(sign * float(self.translate_step), moves[k])
            return True
        if k == "[":
            self.translate_step = np.float32(self.translate_step / np.float32(2.0))
        elif k == "]":
            self.translate_step = np.float32(self.translate_step * np.float32(2.0))
        return False

    def hpt_view(self) -> N.View:
        """The run() arguments of runit() (hptracing.cpp:60-83)."""
        ss = 2 if self.supersample else 1
        w_ang = float(np.float32(self.width) / np.float32(self.height) * self.angle)
        return make_view(tuple(self.view_point), tuple(self.up_dir), tuple(self.right_dir),
                         w_ang, float(self.angle), self.width, self.height, ss, ss,
                         self.background, self.sample, self.depth, int(self.no_diffuse),
                         self.brightness, seed=self.seed)

    # ------------------------------------------------------------ frames
    def frames(self, timeout_s: float = -1.0) -> Iterator[Tuple[np.ndarray, int]]:
        """Render the current view progressively: yields (rgba, passes) each
        time the server publishes more passes, ending with the full-sample
        image.  rgba rows are camera rows (the CLI flips them for files)."""
        rid = self.server.submit(self.hpt_view(), progressive=True)
        seen = 0
        try:
            while seen < self.sample:
                img, passes = self.server.peek(rid, min_passes=seen + 1, timeout_s=timeout_s)
                if passes <= seen:
                    if timeout_s >= 0:
                        raise TimeoutError("no new passes before the timeout")
                    break  # without a timeout peek only returns early once the request ended
                seen = passes
                yield img, passes
        finally:
            self.server.wait(rid)

    def render(self) -> Optional[np.ndarray]:
        """The full-sample image of the current view (the last frame)."""
        last = None
        for img, _ in self.frames():
            last = img
        return last
