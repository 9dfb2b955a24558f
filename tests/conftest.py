"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs
on the CPU.  The oracle (oracle/) is imported here as the checker only."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def frac_within_4se(mean, ref_mean, se, rel_floor=1e-4):
    """Fraction of pixel channels whose mean lies within 4 standard errors
    `se` of `ref_mean`, and the RMSE over finite channels.  Pixels of
    constant radiance (a direct view of an emitter: sample std 0) differ only
    by rounding, which a relative floor admits (1e-4; for long float32 sums
    the caller passes the summation bound, terms x 2^-24).  A NaN pixel (the
    reference's intersection test accepts a ray lying in a triangle's plane
    at t = inf, and a zero-strength path then adds 0 * inf to its pixel)
    counts as outside."""
    diff = np.abs(mean - ref_mean)
    ok = diff <= 4 * se + rel_floor * np.abs(ref_mean) + 1e-9
    return float(ok.mean()), float(np.sqrt(np.nanmean((mean - ref_mean) ** 2)))


def within_4se(mean, spp, ref):
    """An spp-sample image against the float64 reference render
    (golden/room_reference/reference.npz, per-pixel mean, std and se)."""
    se = np.sqrt(ref["se"] ** 2 + (ref["std"] / np.sqrt(spp)) ** 2)
    return frac_within_4se(mean, ref["mean"], se)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("scenes"))


@pytest.fixture(scope="session")
def cornell():
    from hptracing_amd.scene import KDTree
    d = os.path.join(GOLDEN, "cornell_box")
    return KDTree(os.path.join(d, "cornell_box.obj"), d + "/")


@pytest.fixture(scope="session")
def room(scene_dir):
    from hptracing_amd import scenes
    from hptracing_amd.scene import KDTree
    p = scenes.write_room(os.path.join(scene_dir, "room"), 0)
    return KDTree(p, os.path.dirname(p) + "/")


@pytest.fixture(scope="session")
def soup_small():
    from hptracing_amd import scenes
    from hptracing_amd.scene import KDTree, material
    pts, mesh = scenes.soup(20000, seed=3)
    return KDTree.from_arrays(pts, mesh, np.array([material(kd=(0.5, 0.5, 0.5))]))


@pytest.fixture(scope="session")
def soup_200k():
    """The ray-cast workload's soup at the largest size the reference's KD
    builder fits in memory (200 000 triangles, 57 M leaf entries)."""
    from hptracing_amd import scenes
    from hptracing_amd.scene import KDTree, material
    pts, mesh = scenes.soup(200_000, seed=1)
    return KDTree.from_arrays(pts, mesh, np.array([material(kd=(0.5, 0.5, 0.5))]))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle
