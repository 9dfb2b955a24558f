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
def oracle_mod():
    import oracle
    return oracle
