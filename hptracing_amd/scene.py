"""Host-side scene: the Python face of hp::Scene / hp::KDTree.

``KDTree(filename, mtl_basepath)`` loads an OBJ/MTL scene and builds the KD
tree exactly as the reference's ``KDTree`` constructor does
(src/hp/scene/kdtree.cpp:235-246), through libhpt_host.so.  Its attributes
mirror the reference's public vectors (scene/base.h:30-40) as numpy arrays,
and ``get_data()`` returns the flattened (nodes, leaf ints) pair of
``KDTree::getData`` (kdtree.cpp:248-292).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native as N

MATERIAL_DTYPE = np.dtype([
    ("emission", "<f4", 4), ("ambient", "<f4", 4), ("diffuse", "<f4", 4),
    ("specular", "<f4", 4), ("optical_density", "<f4"), ("dissolve", "<f4"),
    ("shininess", "<f4"), ("diffuse_texture_id", "<i4"), ("ambient_texture_id", "<i4"),
    ("specular_possibility", "<f4"), ("refract_possibility", "<f4"),
    ("diffuse_possibility", "<f4"),
])
assert MATERIAL_DTYPE.itemsize == 96

NODE_DTYPE = np.dtype([
    ("box_start", "<f4", 4), ("box_end", "<f4", 4), ("child", "<i4"),
    ("parent", "<i4"), ("sibling", "<i4"), ("data", "<i4"),
])
assert NODE_DTYPE.itemsize == 48


def _array(ptr, count, dtype, width=1):
    dtype = np.dtype(dtype)
    if count == 0 or not ptr:
        shape = (0, width) if width > 1 else (0,)
        return np.zeros(shape, dtype=dtype)
    nbytes = count * width * dtype.itemsize
    buf = (C.c_char * nbytes).from_address(ptr)
    a = np.frombuffer(buf, dtype=dtype).copy()
    return a.reshape(count, width) if width > 1 else a


def read_image(path: str) -> np.ndarray:
    """An image file as the texture pipeline reads it (CImg<float>(path),
    scene/base.cpp:100): (height, width, channels) float32 samples, gray
    images with one channel and colour images with three."""
    w, h, c = C.c_int32(), C.c_int32(), C.c_int32()
    lib = N.host()
    N.check_host(lib.hpt_host_read_image(path.encode(), C.byref(w), C.byref(h), C.byref(c),
                                         None, 0))
    out = np.empty((c.value, h.value, w.value), np.float32)
    N.check_host(lib.hpt_host_read_image(path.encode(), C.byref(w), C.byref(h), C.byref(c),
                                         out.ctypes.data, out.size))
    return np.ascontiguousarray(out.transpose(1, 2, 0))


def material(ka=(0, 0, 0), kd=(0, 0, 0), ks=(0, 0, 0), ke=(0, 0, 0), ior=1.0,
             dissolve=1.0, shininess=1.0) -> np.void:
    """One material record with the reference's branch probabilities
    (scene/base.cpp:85-95), as a MATERIAL_DTYPE scalar."""
    out = N.Material()
    f3 = C.c_float * 3
    N.check_host(N.host().hpt_host_material_from_mtl(
        f3(*ka), f3(*kd), f3(*ks), f3(*ke), ior, dissolve, shininess, C.byref(out)))
    return np.frombuffer(bytes(out), dtype=MATERIAL_DTYPE)[0]


class KDTree:
    """Loaded scene plus KD tree (hp::KDTree)."""

    def __init__(self, filename: Optional[str] = None, mtl_basepath: str = "", *,
                 n_threads: int = 0, _handle=None):
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            N.check_host(N.host().hpt_host_load_obj(
                filename.encode(), mtl_basepath.encode(), n_threads, C.byref(self._h)))
        self._read_arrays()

    @classmethod
    def from_arrays(cls, points: np.ndarray, mesh: np.ndarray, materials: np.ndarray,
                    n_threads: int = 0) -> "KDTree":
        """Scene from in-memory arrays: points (P,3) float32, mesh (T,4) int32
        (v0, v1, v2, material), materials MATERIAL_DTYPE records."""
        points = np.ascontiguousarray(points, dtype=np.float32)
        mesh = np.ascontiguousarray(mesh, dtype=np.int32)
        materials = np.ascontiguousarray(materials, dtype=MATERIAL_DTYPE)
        h = C.c_void_p()
        N.check_host(N.host().hpt_host_from_arrays(
            points.ctypes.data, len(points), mesh.ctypes.data, len(mesh),
            materials.ctypes.data, len(materials), n_threads, C.byref(h)))
        return cls(_handle=h)

    @staticmethod
    def count(points: np.ndarray, mesh: np.ndarray, n_threads: int = 0) -> dict:
        """Size of the tree from_arrays would build over these triangles,
        without building it (hpt_host_kd_count): nodes, non-empty leaves,
        leaf entries, leaf ints, largest leaf, depth and seconds."""
        points = np.ascontiguousarray(points, dtype=np.float32)
        mesh = np.ascontiguousarray(mesh, dtype=np.int32)
        out = N.KdCounts()
        N.check_host(N.host().hpt_host_kd_count(points.ctypes.data, len(points),
                                                mesh.ctypes.data, len(mesh), n_threads,
                                                C.byref(out)))
        return {k: getattr(out, k) for k, _ in N.KdCounts._fields_ if k != "reserved"}

    def _read_arrays(self) -> None:
        d = N.SceneDesc()
        N.check_host(N.host().hpt_host_describe(self._h, C.byref(d)))
        self.points = _array(d.points, d.n_points, np.float32, 4)
        self.normals = _array(d.normals, d.n_normals, np.float32, 4)
        self.texcoords = _array(d.texcoords, d.n_texcoords, np.float32, 2)
        self.geometries = _array(d.mesh, d.n_triangles, np.int32, 4)
        self.materials = _array(d.materials, d.n_materials, MATERIAL_DTYPE)
        self.lights = _array(d.lights, d.n_lights, np.int32, 4)
        self.nodes = _array(d.nodes, d.n_nodes, NODE_DTYPE)
        self.leaf_data = _array(d.leaf_data, d.n_leaf_data, np.int32)
        self.n_textures = int(d.n_textures)
        self.texture_data = _array(d.textures, d.n_textures * N.HPT_TEXTURE_SIZE ** 2 * 4,
                                   np.uint8)
        self.build_seconds = float(N.host().hpt_host_build_seconds(self._h))

    def get_data(self):
        """(nodes, leaf_data): KDTree::getData()."""
        return self.nodes, self.leaf_data

    def desc(self) -> N.SceneDesc:
        """Descriptor pointing into the native scene (valid while self lives)."""
        d = N.SceneDesc()
        N.check_host(N.host().hpt_host_describe(self._h, C.byref(d)))
        return d

    def close(self) -> None:
        if self._h:
            N.host().hpt_host_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_triangles(self) -> int:
        return len(self.geometries)
