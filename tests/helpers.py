"""Shared test helpers: small seeded scenes, the oracle's view of a BoundingHierarchy."""
from __future__ import annotations

import ctypes

import numpy as np

import oracle
from gonzales_amd import abi
from gonzales_amd.accelerator import Primitives, Scene, TriangleMesh, build_bounding_hierarchy
from gonzales_amd import scenes


def random_soup_hierarchy(n: int, seed: int, spread: float = 1.0, scale: float = 0.05):
    """n random triangles in [0, spread]^3 with edges ~ scale (two meshes, to exercise mesh ids)."""
    g = np.random.default_rng(seed)
    c = g.random((n, 3)) * spread
    v = c[:, None, :] + scale * (g.random((n, 3, 3)) - 0.5)
    half = n // 2
    meshes = []
    parts = []
    for mi, sl in enumerate((slice(0, half), slice(half, n))):
        vv = v[sl].reshape(-1, 3).astype(np.float32)
        meshes.append(TriangleMesh.from_xyz(vv, np.arange(vv.shape[0], dtype=np.int64)))
        parts.append(Primitives.geometric(mi, vv.shape[0] // 3, mi % 2))
    mats = np.zeros(2, abi.MATERIAL)
    mats["type"] = 1
    mats["albedo"] = [[0.5, 0.4, 0.3], [0.2, 0.7, 0.6]]
    scene = Scene(meshes, mats)
    return build_bounding_hierarchy(scene, Primitives.concat(parts))


def random_rays(n: int, seed: int, lo=-0.2, hi=1.2, tmax=(0.1, 2.0)):
    g = np.random.default_rng(seed)
    rays = np.zeros(n, abi.RAY)
    o = g.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    for k, name in enumerate(("orgX", "orgY", "orgZ")):
        rays[name] = o[:, k]
    for k, name in enumerate(("dirX", "dirY", "dirZ")):
        rays[name] = d[:, k]
    t = g.uniform(tmax[0], tmax[1], n).astype(np.float32)
    return rays, t


class OracleView:
    """Keeps the ctypes descriptor of a BoundingHierarchy alive for oracle calls."""

    def __init__(self, bh):
        self.bh = bh
        self.desc, self._meshes = bh.descriptor()

    @property
    def ptr(self) -> int:
        return ctypes.addressof(self.desc)

    def closest(self, rays, tmax, nthreads=8, stats=False):
        return oracle.traverse(self.ptr, len(self.bh.nodes), rays, tmax, abi.INTERSECTION, nthreads, stats)

    def occluded(self, rays, tmax, nthreads=8, stats=False):
        return oracle.occluded(self.ptr, len(self.bh.nodes), rays, tmax, nthreads, stats)

    def brute(self, rays, tmax):
        return oracle.brute_closest(self.ptr, len(self.bh.prim_ids), rays, tmax)


def f32_bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_isect_equal(a: np.ndarray, b: np.ndarray) -> None:
    """Intersection_C records equal field by field (floats bitwise; struct padding ignored, since
    numpy does not keep padding bytes through np.load)."""
    assert len(a) == len(b)
    for f in ("id1", "id2", "materialIndex", "type"):
        assert np.array_equal(a["primId"][f], b["primId"][f]), f
    for f in ("tHit", "u", "v"):
        assert np.array_equal(f32_bits(a[f]), f32_bits(b[f])), f
    assert np.array_equal(a["hit"], b["hit"])


__all__ = ["random_soup_hierarchy", "random_rays", "OracleView", "f32_bits", "scenes"]
