"""Every scene the reference ships under scene/ goes through the host
pipeline (OBJ/MTL, textures, emitter table, KD build) as the reference's
Scene and KDTree constructors take it (scene/base.cpp:37-158,
kdtree.cpp:235-246): the scenes load, and a scene whose mtllib file is
missing is refused, as the reference rethrows tinyobj's error
(base.cpp:41-45, tiny_obj_loader.cpp:643-651).  Skipped where the reference
tree is absent."""
from __future__ import annotations

import os

import numpy as np
import pytest

from hptracing_amd import _native as N
from hptracing_amd.scene import KDTree

SCENES = "/root/reference/scene"

pytestmark = pytest.mark.skipif(not os.path.isdir(SCENES), reason="reference scenes absent")

# (scene dir, obj, triangles) for every OBJ of the reference whose mtllib exists
LOADABLE = [
    ("cornell-box", "untitled.obj", 22792),
    ("cornell-box-orig", "CornellBox-Original.obj", 34),
    ("cornell-box-orig", "CornellBox-Mirror.obj", 34),
    ("cornell-box-orig", "CornellBox-Glossy.obj", 1112),
    ("cornell-box-orig", "CornellBox-Sphere.obj", 2188),
    ("cornell-box-orig", "CornellBox-Water.obj", 7088),
    ("cornell-box-orig", "CornellBox-Empty-CO.obj", 12),
    ("cornell-box-orig", "CornellBox-Empty-RG.obj", 12),
    ("cornell-box-orig", "CornellBox-Empty-Squashed.obj", 12),
    ("cornell-box-orig", "CornellBox-Empty-White.obj", 12),
    ("geometries", "geometries.obj", 58),
    ("simple-cornell-box", "cornell_box.obj", 34),
    ("vr-scene", "final.obj", 36006),
]


@pytest.mark.parametrize("scene,obj,n_tris", LOADABLE, ids=[o for _, o, _ in LOADABLE])
def test_reference_scene_loads(scene, obj, n_tris):
    d = os.path.join(SCENES, scene)
    if not os.path.exists(os.path.join(d, obj)):
        pytest.skip(f"{scene}/{obj} absent")
    kd = KDTree(os.path.join(d, obj), d + "/")
    assert kd.n_triangles == n_tris
    nodes = kd.nodes
    # the flattened tree is in pre-order: a child follows its parent, and
    # every leaf list lies inside the leaf data
    inner = nodes["child"] >= 0
    assert np.array_equal(nodes["child"][inner], np.nonzero(inner)[0] + 1)
    leaf = ~inner & (nodes["data"] >= 0)
    counts = kd.leaf_data[nodes["data"][leaf]]
    assert (counts > 0).all()
    assert (nodes["data"][leaf] + counts < len(kd.leaf_data)).all()


def test_missing_material_library_is_refused():
    d = os.path.join(SCENES, "cornell-box-orig")
    if not os.path.exists(os.path.join(d, "water.obj")) or os.path.exists(os.path.join(d, "water.mtl")):
        pytest.skip("water.obj absent or its mtllib present")
    with pytest.raises(N.HptError) as e:
        KDTree(os.path.join(d, "water.obj"), d + "/")
    assert e.value.code == N.HPT_ERR_IO
    assert "material library" in str(e.value)
