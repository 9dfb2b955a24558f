"""Seeded procedural scenes and ray batches of the benchmark workloads.

Every scene is written as OBJ + MTL (+ a binary PPM texture) so that it goes
through the same loader path as a user scene.  All shapes carry `vn` and
`vt` records: the reference reads normals and texture coordinates by global
vertex index (kernel.cl:306-318), so the streams must cover every vertex.

  room        1 316 triangles: 5 diffuse wall quads (albedo U[0.2,0.8]^3), a
              ceiling light quad (Ke 15), two axis-aligned diffuse boxes and a
              glass icosphere of 3 subdivisions (IOR 1.5).
  full        room + a heightfield water grid (sum of 8 seeded sines, IOR
              1.33), a surface of revolution with a seeded Bezier profile and
              a 1024x1024 value-noise texture, and a mirror sphere.
  soup        triangles with centres uniform in the unit cube and edges of
              random direction and log-uniform length in [1e-3, 1e-1].
"""
from __future__ import annotations

import hashlib
import os
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

ROOM = 10.0  # the room spans [0, ROOM]^3, open towards -z


@dataclass
class Mesh:
    v: np.ndarray      # (P,3) float
    vn: np.ndarray     # (P,3) float
    vt: np.ndarray     # (P,2) float
    f: np.ndarray      # (T,3) int, 0-based into v/vn/vt of this mesh
    material: str
    name: str


@dataclass
class Material:
    name: str
    kd: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    ks: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    ke: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    ka: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    ni: float = 1.0
    d: float = 1.0
    ns: float = 1.0
    map_kd: str = ""


# ------------------------------------------------------------------ shapes
def quad(p0, p1, p2, p3, material, name) -> Mesh:
    v = np.array([p0, p1, p2, p3], np.float64)
    n = np.cross(v[1] - v[0], v[2] - v[0])
    n = n / np.linalg.norm(n)
    return Mesh(v, np.tile(n, (4, 1)), np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float),
                np.array([[0, 1, 2], [0, 2, 3]]), material, name)


def box(lo, hi, material, name) -> Mesh:
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], float)
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (3, 7, 6, 2), (0, 4, 7, 3), (1, 2, 6, 5)]
    vs, ns, fs = [], [], []
    for k, (a, b, cc, d) in enumerate(faces):
        q = quad(c[a], c[b], c[cc], c[d], material, name)
        vs.append(q.v)
        ns.append(q.vn)
        fs.append(q.f + 4 * k)
    v = np.concatenate(vs)
    return Mesh(v, np.concatenate(ns), np.zeros((len(v), 2)), np.concatenate(fs), material, name)


def icosphere(center, radius, subdivisions, material, name) -> Mesh:
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    n = np.array(v)
    return Mesh(np.asarray(center) + radius * n, n, np.zeros((len(n), 2)), np.array(f),
                material, name)


def uv_sphere(center, radius, stacks, slices, material, name) -> Mesh:
    th = np.linspace(0, np.pi, stacks + 1)
    ph = np.linspace(0, 2 * np.pi, slices + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    idx = np.arange((stacks + 1) * (slices + 1)).reshape(stacks + 1, slices + 1)
    a, b = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel()
    c, d = idx[1:, 1:].ravel(), idx[1:, :-1].ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    uv = np.stack([P / (2 * np.pi), T / np.pi], -1).reshape(-1, 2)
    return Mesh(np.asarray(center) + radius * n, n, uv, f, material, name)


def grid_mesh(xs, zs, height_fn, normal_fn, material, name, uv=False) -> Mesh:
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = height_fn(X, Z)
    v = np.stack([X, Y, Z], -1).reshape(-1, 3)
    n = normal_fn(X, Z).reshape(-1, 3)
    nx, nz = len(xs), len(zs)
    idx = np.arange(nx * nz).reshape(nx, nz)
    a, b = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    c, d = idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    t = np.zeros((len(v), 2))
    if uv:
        t = np.stack([(X - xs[0]) / (xs[-1] - xs[0]), (Z - zs[0]) / (zs[-1] - zs[0])],
                     -1).reshape(-1, 2)
    return Mesh(v, n, t, f, material, name)


def water_grid(rng, n_side, level, material, name) -> Mesh:
    k = rng.uniform(0.5, 3.0, (8, 2)) * rng.choice([-1, 1], (8, 2))
    amp = rng.uniform(0.01, 0.06, 8)
    ph = rng.uniform(0, 2 * np.pi, 8)

    def h(X, Z):
        return level + sum(amp[i] * np.sin(k[i, 0] * X + k[i, 1] * Z + ph[i]) for i in range(8))

    def nrm(X, Z):
        dx = sum(amp[i] * k[i, 0] * np.cos(k[i, 0] * X + k[i, 1] * Z + ph[i]) for i in range(8))
        dz = sum(amp[i] * k[i, 1] * np.cos(k[i, 0] * X + k[i, 1] * Z + ph[i]) for i in range(8))
        n = np.stack([-dx, np.ones_like(dx), -dz], -1)
        return n / np.linalg.norm(n, axis=-1, keepdims=True)

    xs = np.linspace(0.02, ROOM - 0.02, n_side + 1)
    return grid_mesh(xs, xs, h, nrm, material, name)


def revolution(rng, center, n_around, n_along, material, name) -> Mesh:
    ctrl = np.stack([rng.uniform(0.3, 1.4, 4), np.sort(rng.uniform(0.0, 3.0, 4))], 1)
    ctrl[0, 1], ctrl[-1, 1] = 0.0, 3.0
    s = np.linspace(0, 1, n_along + 1)
    B = np.stack([(1 - s) ** 3, 3 * s * (1 - s) ** 2, 3 * s ** 2 * (1 - s), s ** 3], 1)
    prof = B @ ctrl                      # (K,2): radius, height
    dB = np.stack([-3 * (1 - s) ** 2, 3 * (1 - s) ** 2 - 6 * s * (1 - s),
                   6 * s * (1 - s) - 3 * s ** 2, 3 * s ** 2], 1)
    dprof = dB @ ctrl
    ang = np.linspace(0, 2 * np.pi, n_around + 1)
    A, S = np.meshgrid(ang, np.arange(n_along + 1), indexing="ij")
    r, y = prof[S, 0], prof[S, 1]
    dr, dy = dprof[S, 0], dprof[S, 1]
    v = np.stack([center[0] + r * np.cos(A), center[1] + y, center[2] + r * np.sin(A)], -1)
    n = np.stack([dy * np.cos(A), -dr, dy * np.sin(A)], -1)
    n = n / np.maximum(np.linalg.norm(n, axis=-1, keepdims=True), 1e-12)
    uv = np.stack([A / (2 * np.pi), S / n_along], -1)
    idx = np.arange((n_around + 1) * (n_along + 1)).reshape(n_around + 1, n_along + 1)
    a, b = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    c, d = idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    return Mesh(v.reshape(-1, 3), n.reshape(-1, 3), uv.reshape(-1, 2), f, material, name)


def value_noise(rng, size=1024, octaves=4) -> np.ndarray:
    img = np.zeros((size, size, 3))
    amp, cells = 1.0, 8
    for _ in range(octaves):
        lat = rng.uniform(0, 1, (cells + 1, cells + 1, 3))
        x = np.linspace(0, cells, size, endpoint=False)
        i0 = x.astype(int)
        fr = x - i0
        fr = fr * fr * (3 - 2 * fr)
        a = lat[i0][:, i0] * (1 - fr)[None, :, None] + lat[i0][:, i0 + 1] * fr[None, :, None]
        b = lat[i0 + 1][:, i0] * (1 - fr)[None, :, None] + lat[i0 + 1][:, i0 + 1] * fr[None, :, None]
        img += amp * (a * (1 - fr)[:, None, None] + b * fr[:, None, None])
        amp *= 0.5
        cells *= 2
    img = (img - img.min()) / (img.max() - img.min())
    return (img * 230 + 10).astype(np.uint8)


# ------------------------------------------------------------------ writer
def _fmt(a: np.ndarray, prefix: str, fmt: str) -> str:
    if len(a) == 0:
        return ""
    cols = a.shape[1]
    line = prefix + " " + " ".join([fmt] * cols)
    return "\n".join(line % tuple(r) for r in a.tolist()) + "\n"


def write_obj(path: str, meshes: List[Mesh], materials: List[Material]) -> str:
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    base = os.path.splitext(os.path.basename(path))[0]
    mtl = base + ".mtl"
    with open(os.path.join(d, mtl), "w") as fh:
        for m in materials:
            fh.write(f"newmtl {m.name}\n")
            fh.write("Ka %g %g %g\nKd %g %g %g\nKs %g %g %g\n" % (*m.ka, *m.kd, *m.ks))
            if any(m.ke):
                fh.write("Ke %g %g %g\n" % m.ke)
            fh.write(f"Ni {m.ni:g}\nd {m.d:g}\nNs {m.ns:g}\n")
            if m.map_kd:
                fh.write(f"map_Kd {m.map_kd}\n")
            fh.write("\n")
    with open(path, "w") as fh:
        fh.write(f"mtllib {mtl}\n")
        off = 1
        for m in meshes:
            fh.write(f"o {m.name}\nusemtl {m.material}\n")
            fh.write(_fmt(m.v, "v", "%.7g"))
            fh.write(_fmt(m.vn, "vn", "%.6g"))
            fh.write(_fmt(m.vt, "vt", "%.6g"))
            f = m.f + off
            fh.write("\n".join("f %d/%d/%d %d/%d/%d %d/%d/%d" % (a, a, a, b, b, b, c, c, c)
                               for a, b, c in f.tolist()) + "\n")
            off += len(m.v)
    return path


def write_ppm(path: str, img: np.ndarray) -> None:
    h, w, _ = img.shape
    with open(path, "wb") as fh:
        fh.write(b"P6\n%d %d\n255\n" % (w, h))
        fh.write(np.ascontiguousarray(img, np.uint8).tobytes())


# ------------------------------------------------------------------ scenes
def room_parts(rng) -> Tuple[List[Mesh], List[Material]]:
    R = ROOM
    mats = [Material(f"wall{i}", kd=tuple(rng.uniform(0.2, 0.8, 3))) for i in range(5)]
    mats.append(Material("light", ke=(15.0, 15.0, 15.0)))
    mats += [Material("box0", kd=tuple(rng.uniform(0.2, 0.8, 3))),
             Material("box1", kd=tuple(rng.uniform(0.2, 0.8, 3)))]
    mats.append(Material("glass", ks=(0.1, 0.1, 0.1), ni=1.5, d=0.0))
    meshes = [
        quad((0, 0, 0), (0, 0, R), (R, 0, R), (R, 0, 0), "wall0", "floor"),
        quad((0, R, 0), (R, R, 0), (R, R, R), (0, R, R), "wall1", "ceiling"),
        quad((0, 0, R), (0, R, R), (R, R, R), (R, 0, R), "wall2", "back"),
        quad((0, 0, 0), (0, R, 0), (0, R, R), (0, 0, R), "wall3", "left"),
        quad((R, 0, 0), (R, 0, R), (R, R, R), (R, R, 0), "wall4", "right"),
        quad((4, R - 0.01, 4), (6, R - 0.01, 4), (6, R - 0.01, 6), (4, R - 0.01, 6), "light", "light"),
        box((1.5, 0, 5.5), (3.5, 3.0, 7.5), "box0", "tall_box"),
        box((6.5, 0, 2.5), (8.5, 1.5, 4.5), "box1", "short_box"),
        icosphere((5.0, 2.0, 4.0), 1.5, 3, "glass", "glass_ball"),
    ]
    return meshes, mats


def write_room(directory: str, seed: int = 0) -> str:
    rng = np.random.default_rng(seed)
    meshes, mats = room_parts(rng)
    return write_obj(os.path.join(directory, f"room_s{seed}.obj"), meshes, mats)


def write_full(directory: str, seed: int = 0, water_side: int = 500, rev=(400, 250),
               sphere=(100, 100), textured: bool = True) -> str:
    """The full path-trace scene (water_side=500: 500k water triangles;
    water_side=1000 gives the 2M-triangle serving variant).  textured=False
    gives the vase a plain grey diffuse colour instead of its texture (same
    geometry)."""
    rng = np.random.default_rng(seed)
    meshes, mats = room_parts(rng)
    vase = (Material("vase", kd=(1.0, 1.0, 1.0), map_kd=f"noise_s{seed}.ppm") if textured
            else Material("vase", kd=(0.6, 0.6, 0.6)))
    mats += [Material("water", ks=(0.05, 0.05, 0.05), ni=1.33, d=0.0), vase,
             Material("mirror", ks=(0.95, 0.95, 0.95), ns=1000.0)]
    meshes.append(water_grid(rng, water_side, 0.8, "water", "water"))
    meshes.append(revolution(rng, (7.5, 0.0, 7.5), rev[0], rev[1], "vase", "vase"))
    meshes.append(uv_sphere((2.5, 5.0, 3.0), 1.2, sphere[0], sphere[1], "mirror", "mirror_ball"))
    os.makedirs(directory, exist_ok=True)
    write_ppm(os.path.join(directory, f"noise_s{seed}.ppm"), value_noise(rng))
    name = f"full_s{seed}_w{water_side}{'' if textured else '_plain'}.obj"
    return write_obj(os.path.join(directory, name), meshes, mats)


def cached(directory: str, writer, *args, **kw) -> str:
    """Write a scene once per argument set into `directory` and reuse it."""
    key = hashlib.sha1(repr((writer.__name__, args, sorted(kw.items()))).encode()).hexdigest()[:10]
    d = os.path.join(directory, key)
    marker = os.path.join(d, "done")
    if os.path.exists(marker):
        with open(marker) as fh:
            return fh.read().strip()
    path = writer(d, *args, **kw)
    with open(marker, "w") as fh:
        fh.write(path)
    return path


# --------------------------------------------------------------- ray cast
def soup(n_tris: int, seed: int = 0):
    """Triangle soup: (points (3n,3) float32, mesh (n,4) int32)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (n_tris, 3))

    def edges():
        d = rng.normal(size=(n_tris, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return d * (10.0 ** rng.uniform(-3, -1, (n_tris, 1)))

    e1, e2 = edges(), edges()
    v0 = c - (e1 + e2) / 3
    pts = np.stack([v0, v0 + e1, v0 + e2], 1).reshape(-1, 3).astype(np.float32)
    mesh = np.zeros((n_tris, 4), np.int32)
    mesh[:, :3] = np.arange(3 * n_tris).reshape(n_tris, 3)
    return pts, mesh


def random_rays(n: int, lo, hi, seed: int = 1):
    """Origins uniform in the box [lo, hi], directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    org = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org, d.astype(np.float32)


# Camera of the room / full scenes: looking +z into the room from its open side.
ROOM_VIEW = dict(view_p=(5.0, 5.0, -6.0), top_dir=(0.0, 1.0, 0.0), right_dir=(-1.0, 0.0, 0.0))
