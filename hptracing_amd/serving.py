"""Load generation for the render server: Poisson arrivals of a request mix,
latency percentiles, and the same request list rendered one view at a time
for comparison.

Request mix of the serving workload (2M-triangle scene):
  60 %  256 x 256 previews, 4 spp
  30 %  1280 x 720 views, 32 spp
  10 %  2048 x 1024 equirectangular panoramas (angle = -1), 16 spp
Views are drawn uniformly inside the room (camera positions and yaw), depth 8.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _native as N
from .runner import RenderServer, TraceRunner, make_view
from .scenes import ROOM

MIX = [  # (weight, width, height, spp, vr)
    (0.6, 256, 256, 4, False),
    (0.3, 1280, 720, 32, False),
    (0.1, 2048, 1024, 16, True),
]


def request_views(n: int, seed: int = 0, depth: int = 8, scale: float = 1.0) -> List[N.View]:
    """n requests drawn from MIX; `scale` shrinks image sides (for tests)."""
    rng = np.random.default_rng(seed)
    weights = np.array([m[0] for m in MIX])
    kinds = rng.choice(len(MIX), size=n, p=weights / weights.sum())
    views = []
    for i, k in enumerate(kinds):
        _, w, h, spp, vr = MIX[k]
        w, h = max(8, int(w * scale)), max(8, int(h * scale))
        pos = rng.uniform([1.5, 1.5, 1.5], [ROOM - 1.5, ROOM - 2.0, ROOM - 1.5])
        yaw = rng.uniform(0, 2 * math.pi)
        right = (math.cos(yaw), 0.0, math.sin(yaw))
        if vr:
            # angle = -1: the equirectangular camera; right.x carries the
            # horizontal pixel shift (trace_runner.cpp:132-133)
            v = make_view(pos, (0, 1, 0), (float(int(rng.integers(0, w))), 0, 0), -1.0, -1.0,
                          w, h, 1, 1, (0, 0, 0), spp, depth, 0, 1.0, seed=1000 + i)
        else:
            v = make_view(pos, (0, 1, 0), right, w / h, 1.0, w, h, 1, 1, (0, 0, 0), spp, depth,
                          0, 1.0, seed=1000 + i)
        views.append(v)
    return views


def batch_views(n: int = 64, seed: int = 7, depth: int = 8) -> List[N.View]:
    """The batching workload: n pinhole views of the full scene with
    resolutions uniform in [128, 1024] x [128, 768], spp uniform in
    [4, 64], positions uniform inside the room and view/up vectors of
    random yaw and pitch (|pitch| <= 0.3 rad)."""
    rng = np.random.default_rng(seed)
    views = []
    for i in range(n):
        w = int(rng.integers(128, 1025))
        h = int(rng.integers(128, 769))
        spp = int(rng.integers(4, 65))
        pos = rng.uniform([1.5, 1.5, 1.5], [ROOM - 1.5, ROOM - 2.0, ROOM - 1.5])
        yaw = rng.uniform(0, 2 * math.pi)
        pitch = rng.uniform(-0.3, 0.3)
        up = (math.sin(pitch) * -math.sin(yaw), math.cos(pitch), math.sin(pitch) * math.cos(yaw))
        right = (math.cos(yaw), 0.0, math.sin(yaw))
        views.append(make_view(tuple(pos), up, right, w / h, 1.0, w, h, 1, 1, (0, 0, 0), spp,
                               depth, 0, 1.0, seed=500 + i))
    return views


def camera_samples(v: N.View) -> int:
    return v.sample_x * v.supersample_x * v.sample_y * v.supersample_y * v.sample


@dataclass
class ServeResult:
    rate: float
    n: int
    p50_ms: float
    p99_ms: float
    mean_ms: float
    wall_s: float
    rays: int
    mrays_s: float
    macro_passes: int


def serve(runner: TraceRunner, views: List[N.View], rate: float, seed: int = 1,
          budget: int = 48 << 30) -> ServeResult:
    """Submit `views` with exponential inter-arrival gaps (rate per second;
    rate <= 0 submits them all at once) and wait for every result."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(views)) if rate > 0 else np.zeros(len(views))
    srv = RenderServer(runner, device_budget_bytes=budget)
    ids: List[int] = [0] * len(views)
    submitted = threading.Event()

    def producer():
        t_next = time.perf_counter()
        for i, v in enumerate(views):
            t_next += gaps[i]
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids[i] = srv.submit(v)
            if i == 0:
                submitted.set()
        submitted.set()

    t0 = time.perf_counter()
    th = threading.Thread(target=producer)
    th.start()
    lat = []
    submitted.wait()
    for i in range(len(views)):
        while ids[i] == 0:
            time.sleep(0.0005)
        _, _, info = srv.wait(ids[i])
        lat.append(info["latency_ms"])
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    return ServeResult(rate, len(views), float(np.percentile(lat, 50)),
                       float(np.percentile(lat, 99)), float(lat.mean()), wall,
                       int(st["rays_cast"]), st["rays_cast"] / wall / 1e6, int(st["macro_passes"]))


def one_at_a_time(runner: TraceRunner, views: List[N.View], pass_at_a_time: bool = False):
    """Render the same requests sequentially with hpt_render (optionally
    launching one sample pass at a time, the reference's structure)."""
    t0 = time.perf_counter()
    rays = 0
    for v in views:
        runner.render(v, pass_at_a_time=pass_at_a_time)
        rays += int(runner.last_stats.rays_cast)
    wall = time.perf_counter() - t0
    return wall, rays
