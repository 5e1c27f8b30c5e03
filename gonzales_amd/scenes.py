"""Seeded synthetic workloads (triangle scenes and ray batches).

* :func:`triangle_soup` -- N triangles, centroids uniform in [0,1]^3, edge lengths log-normal
  (median 2e-3, sigma 0.5), uniformly random orientations (the 1M-triangle traversal workload).
* :func:`incoherent_rays` -- origins uniform in the unit cube, directions uniform on the sphere,
  tMax uniform in [0.1, 2].
* :func:`coherent_rays` -- one pinhole primary ray per pixel centre of a square image looking
  into the unit cube (eye (0.5, 0.5, -0.5), vertical fov 60 degrees), tMax uniform in [0.1, 2].
* :func:`cornell_box` -- the Cornell-box geometry of the reference's Scenes/cornell-box.pbrt
  (5 wall quads, two boxes rotated about y, one emissive ceiling quad, radiance (17, 12, 4)).

All generators use numpy's PCG64 bit generator with fixed seeds, so every run (here and on the
GPU box) sees the same inputs.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .accelerator import Primitives, Scene, TriangleMesh, build_bounding_hierarchy


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def triangle_soup(n: int = 1 << 20, seed: int = 1) -> TriangleMesh:
    """Triangle soup: centroid c ~ U[0,1]^3, edge scale e ~ LogNormal(ln 2e-3, 0.5), vertices
    c + e R t_k with t_k the unit-edge equilateral triangle centred at 0 and R a uniform random
    rotation (normalised Gaussian quaternion)."""
    g = _rng(seed)
    c = g.random((n, 3))
    e = np.exp(np.log(2e-3) + 0.5 * g.standard_normal(n))
    q = g.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    s3 = np.sqrt(3.0)
    tmpl = np.array([[1 / s3, 0.0, 0.0], [-1 / (2 * s3), 0.5, 0.0], [-1 / (2 * s3), -0.5, 0.0]])
    verts = c[:, None, :] + e[:, None, None] * np.einsum("nij,kj->nki", R, tmpl)
    pts = np.ones((3 * n, 4), np.float32)
    pts[:, :3] = verts.reshape(-1, 3).astype(np.float32)
    return TriangleMesh(pts, np.arange(3 * n, dtype=np.int64))


def _tmax(n: int, seed: int) -> np.ndarray:
    return _rng(seed).uniform(0.1, 2.0, n).astype(np.float32)


def incoherent_rays(n: int = 1 << 24, seed_org: int = 2, seed_dir: int = 3, seed_tmax: int = 4):
    """Origins U[0,1]^3, directions uniform on S^2 (z = 1 - 2u, phi = 2 pi v), tMax U[0.1, 2]."""
    rays = np.empty(n, abi.RAY)
    o = _rng(seed_org).random((n, 3)).astype(np.float32)
    uv = _rng(seed_dir).random((n, 2))
    z = 1.0 - 2.0 * uv[:, 0]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * uv[:, 1]
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)
    rays["orgX"], rays["orgY"], rays["orgZ"] = o[:, 0], o[:, 1], o[:, 2]
    rays["dirX"], rays["dirY"], rays["dirZ"] = d[:, 0], d[:, 1], d[:, 2]
    return rays, _tmax(n, seed_tmax)


def coherent_rays(res: int = 4096, seed_tmax: int = 4):
    """Pinhole primary rays through pixel centres: eye (0.5, 0.5, -0.5) looking at (0.5, 0.5, 0.5),
    up +y, vertical fov 60 degrees, square image; directions normalised."""
    n = res * res
    j, i = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    tan = np.tan(np.radians(30.0))
    sx = (2.0 * (i + 0.5) / res - 1.0) * tan
    sy = (1.0 - 2.0 * (j + 0.5) / res) * tan
    d = np.stack([sx, sy, np.ones_like(sx)], axis=-1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    rays = np.empty(n, abi.RAY)
    rays["orgX"], rays["orgY"], rays["orgZ"] = 0.5, 0.5, -0.5
    rays["dirX"], rays["dirY"], rays["dirZ"] = d[:, 0], d[:, 1], d[:, 2]
    return rays, _tmax(n, seed_tmax)


def soup_scene(n: int = 1 << 20, seed: int = 1):
    """Single-mesh triangle soup as geometric primitives with one diffuse material."""
    mesh = triangle_soup(n, seed)
    mats = np.zeros(1, abi.MATERIAL)
    mats["type"] = 1
    mats["albedo"] = 0.5
    scene = Scene([mesh], mats)
    prims = Primitives.geometric(0, mesh.number_triangles, 0)
    return scene, prims


def soup_hierarchy(n: int = 1 << 20, seed: int = 1):
    scene, prims = soup_scene(n, seed)
    return build_bounding_hierarchy(scene, prims)


# ---------------------------------------------------------------------------------------------
# Cornell box (Scenes/cornell-box.pbrt:49-162)
# ---------------------------------------------------------------------------------------------

def _quad(p0, p1, p2, p3):
    return [p0, p1, p2, p3], [0, 1, 2, 0, 2, 3]


def _box(cx, cz, half_w, height, angle_deg, with_floor=True):
    """Axis box of half width `half_w`, rotated about y by angle, standing on y = 0."""
    a = np.radians(angle_deg)
    ca, sa = np.cos(a), np.sin(a)

    def P(x, y, z):
        return [cx + ca * x + sa * z, y, cz - sa * x + ca * z]

    h = half_w
    v = [P(-h, 0, -h), P(h, 0, -h), P(h, 0, h), P(-h, 0, h),
         P(-h, height, -h), P(h, height, -h), P(h, height, h), P(-h, height, h)]
    faces = [(4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    if with_floor:
        faces.append((0, 3, 2, 1))
    idx = []
    for f in faces:
        idx += [f[0], f[1], f[2], f[0], f[2], f[3]]
    return v, idx


def cornell_box():
    """34-triangle Cornell box: walls x, z in [-1, 1], y in [0, 2]; light quad at y = 1.98 with
    radiance (17, 12, 4); a tall box rotated +18 degrees and a short box rotated -18 degrees
    (the short box's floor face is omitted: it is never visible and 36 - 2 = 34 triangles).
    Materials: white (0.725, 0.71, 0.68), red (0.63, 0.065, 0.05) and green (0.14, 0.45, 0.091)
    walls clamped into [0.1, 0.8]. Returns (scene, primitives, camera dict)."""
    meshes = []
    parts = []
    mats = []

    def add(verts, idx, material):
        mi = len(meshes)
        meshes.append(TriangleMesh.from_xyz(np.asarray(verts, np.float32), np.asarray(idx, np.int64)))
        parts.append(Primitives.geometric(mi, len(idx) // 3, material))

    white = (0.725, 0.71, 0.68)
    red = tuple(np.clip((0.63, 0.065, 0.05), 0.1, 0.8))
    green = tuple(np.clip((0.14, 0.45, 0.091), 0.1, 0.8))
    for c in (white, red, green):
        m = np.zeros(1, abi.MATERIAL)
        m["type"] = 1
        m["albedo"] = c
        mats.append(m)
    add(*_quad([-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]), 0)   # floor
    add(*_quad([-1, 2, -1], [-1, 2, 1], [1, 2, 1], [1, 2, -1]), 0)   # ceiling
    add(*_quad([-1, 0, -1], [-1, 2, -1], [1, 2, -1], [1, 0, -1]), 0)  # back wall
    add(*_quad([-1, 0, -1], [-1, 0, 1], [-1, 2, 1], [-1, 2, -1]), 1)  # left (red)
    add(*_quad([1, 0, -1], [1, 2, -1], [1, 2, 1], [1, 0, 1]), 2)      # right (green)
    add(*_box(-0.335, -0.29, 0.3, 1.2, 18.0, with_floor=True), 0)   # tall box
    add(*_box(0.33, 0.37, 0.3, 0.6, -18.0, with_floor=False), 0)    # short box
    # Light quad (two triangles) at y = 1.98, facing down.
    li = len(meshes)
    lv, lidx = _quad([-0.24, 1.98, -0.22], [0.23, 1.98, -0.22], [0.23, 1.98, 0.16], [-0.24, 1.98, 0.16])
    meshes.append(TriangleMesh.from_xyz(np.asarray(lv, np.float32), np.asarray(lidx, np.int64)))
    parts.append(Primitives.area_lights(li, 2, 0))
    scene = Scene(meshes, abi.concat(mats, abi.MATERIAL),
                  light_emission=np.array([[17, 12, 4], [17, 12, 4]], np.float32))
    camera = {"eye": (0.0, 1.0, 6.8), "look": (0.0, 1.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 19.5}
    return scene, Primitives.concat(parts), camera


def cornell_hierarchy():
    scene, prims, camera = cornell_box()
    return build_bounding_hierarchy(scene, prims), camera


def total_triangles(scene: Scene) -> int:
    return int(sum(m.number_triangles for m in scene.meshes))
