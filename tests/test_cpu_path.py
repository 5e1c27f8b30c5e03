"""The CPU half of the boundary (gz_traverse_bvh2, gz_cpu_traverse_batch, gz_cpu_shade_batch, the
replacements of mojo_traverse_bvh2 / mojo_cpu_traverse_batch / mojo_cpu_shade_batch,
kernel.mojo:369-403): product host code, checked on the CPU against the oracle and the committed
golden records.

Bar: traversal records bit-exact (the whole 48-byte Intersection_C: PrimId, t, u, v, hit flag);
shade: PCG state, throughput, estimate and the active flag bit-exact, ray origin and direction
bit-exact too (both sides call the C library's cosf/sinf/sqrtf in the same order).
"""
import ctypes
import threading
from pathlib import Path

import numpy as np
import pytest

import oracle
from gonzales_amd import abi, scenes
from gonzales_amd.accelerator import BoundingHierarchy, Scene, TriangleMesh

from helpers import OracleView, assert_isect_equal, random_rays, random_soup_hierarchy


def _cpu_batch(bh, rays, tmax):
    desc, keep = bh.descriptor()
    out = np.zeros(len(rays), abi.INTERSECTION)
    rays = np.ascontiguousarray(rays, abi.RAY)
    tmax = np.ascontiguousarray(tmax, np.float32)
    abi.load().gz_cpu_traverse_batch(ctypes.addressof(desc), abi.ptr(rays), abi.ptr(tmax), len(rays), abi.ptr(out))
    return out


def _with_axis_aligned(rays):
    """Some rays get zero direction components (the slab test's 0 * inf = NaN cases)."""
    rays = rays.copy()
    rays["dirX"][::7] = 0.0
    rays["dirY"][3::11] = 0.0
    rays["dirZ"][5::13] = 0.0
    return rays


@pytest.mark.parametrize("n,seed", [(300, 1), (5000, 31)])
def test_batch_equals_oracle(n, seed):
    bh = random_soup_hierarchy(n, seed)
    rays, tmax = random_rays(20000, seed + 100)
    rays = _with_axis_aligned(rays)
    got = _cpu_batch(bh, rays, tmax)
    ref, _ = OracleView(bh).closest(rays, tmax)
    assert_isect_equal(got, ref)
    assert got["hit"].sum() > 100
    assert abi.last_error() == ""


@pytest.mark.parametrize("name", ["soup300", "cornell"])
def test_golden_intersection_records(name):
    data = np.load(Path(__file__).resolve().parent / "golden" / "traversal_vectors.npz")
    nmesh = len([k for k in data.files if k.startswith(name + "_mesh") and k.endswith("_points")])
    meshes = [TriangleMesh(data[f"{name}_mesh{k}_points"], data[f"{name}_mesh{k}_vtx"]) for k in range(nmesh)]
    bh = BoundingHierarchy(Scene(meshes, np.zeros(0, abi.MATERIAL)), data[name + "_prims"].copy(),
                           data[name + "_nodes"].copy())
    got = _cpu_batch(bh, data[name + "_rays"], data[name + "_tmax"])
    assert_isect_equal(got, data[name + "_isect"])


def test_single_ray_equals_batch():
    bh = random_soup_hierarchy(2000, 7)
    rays, tmax = random_rays(300, 8)
    batch = _cpu_batch(bh, rays, tmax)
    desc, keep = bh.descriptor()
    lib = abi.load()
    one = np.zeros(1, abi.INTERSECTION)
    for i in range(len(rays)):
        r = np.ascontiguousarray(rays[i:i + 1])
        lib.gz_traverse_bvh2(ctypes.addressof(desc), abi.ptr(r), float(tmax[i]), abi.ptr(one))
        assert_isect_equal(one, batch[i:i + 1])


def test_concurrent_callers_get_the_same_records():
    """Re-entrant: one call per thread at once (ctypes releases the interpreter lock)."""
    bh = random_soup_hierarchy(20000, 5)
    batches = [random_rays(40000, 60 + k) for k in range(6)]
    ref = [_cpu_batch(bh, r, t) for r, t in batches]
    out = [None] * len(batches)

    def work(k):
        out[k] = _cpu_batch(bh, *batches[k])

    th = [threading.Thread(target=work, args=(k,)) for k in range(len(batches))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for a, b in zip(out, ref):
        assert_isect_equal(a, b)


def test_empty_descriptor_misses():
    desc = abi.SceneDescriptor2C()
    rays, tmax = random_rays(16, 2)
    out = np.zeros(16, abi.INTERSECTION)
    abi.load().gz_cpu_traverse_batch(ctypes.addressof(desc), abi.ptr(rays), abi.ptr(tmax), 16, abi.ptr(out))
    assert not out["hit"].any() and (out["primId"]["id1"] == -1).all() and (out["primId"]["id2"] == -1).all()
    assert np.array_equal(out["tHit"], tmax)


def test_stack_overflow_is_reported():
    """A chain of 70 interior nodes whose children overlap needs 70 > 64 stack entries."""
    K = 70
    nodes = np.zeros(2 * K + 1, abi.BVH2_NODE)
    for f in ("boundsMinX", "boundsMinY", "boundsMinZ"):
        nodes[f] = -1.0
    for f in ("boundsMaxX", "boundsMaxY", "boundsMaxZ"):
        nodes[f] = 1.0
    for i in range(K):
        nodes["offset"][i] = K + 1 + i
    nodes["count"][K:] = 1
    nodes["offset"][K:] = 0
    mesh = TriangleMesh.from_xyz(np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32), np.arange(3))
    prims = np.zeros(1, abi.PRIM_ID)
    bh = BoundingHierarchy(Scene([mesh], np.zeros(0, abi.MATERIAL)), prims, nodes)
    rays, tmax = random_rays(8, 3, lo=-0.5, hi=0.5)
    _cpu_batch(bh, rays, tmax)
    assert "stack" in abi.last_error()


def test_shade_equals_oracle():
    """Cornell box camera rays: closest hits by the CPU traversal, then one diffuse shade step
    (emitters add and stop, diffuse surfaces take a cosine step from PCG32) against the oracle."""
    bh, _ = scenes.cornell_hierarchy()
    bh.prepare_materials()
    g = np.random.default_rng(0)
    n = 20000
    rays = np.zeros(n, abi.RAY)
    eye = np.array([0.0, 1.0, 6.8])
    d = g.uniform([-1, 0, -1], [1, 2, 1], (n, 3)) - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for k, f in enumerate(("orgX", "orgY", "orgZ")):
        rays[f] = eye[k]
    for k, f in enumerate(("dirX", "dirY", "dirZ")):
        rays[f] = d[:, k].astype(np.float32)
    tmax = np.full(n, np.inf, np.float32)
    isect = _cpu_batch(bh, rays, tmax)
    paths = np.zeros(n, abi.PATH_STATE)
    paths["ray"] = rays
    paths["throughput"] = g.uniform(0.2, 1.0, (n, 3))
    paths["estimate"] = g.uniform(0.0, 0.5, (n, 3))
    paths["pcgState"] = g.integers(0, 2**63, n, dtype=np.uint64)
    paths["pcgInc"] = g.integers(0, 2**63, n, dtype=np.uint64)
    paths["active"] = 1
    paths["active"][::9] = 0
    ref = paths.copy()
    desc, meshes_c = bh.descriptor()
    oracle.shade(ref, isect, ctypes.addressof(meshes_c), bh.materials_c)
    abi.load().gz_cpu_shade_batch(abi.ptr(paths), n, abi.ptr(isect), ctypes.addressof(meshes_c),
                                  abi.ptr(bh.materials_c))
    for f in ("pcgState", "pcgInc", "active"):
        assert np.array_equal(paths[f], ref[f]), f
    for f in ("throughput", "estimate"):
        assert np.array_equal(paths[f].view(np.uint32), ref[f].view(np.uint32)), f
    for f in ("orgX", "orgY", "orgZ", "dirX", "dirY", "dirZ"):
        assert np.array_equal(paths["ray"][f].view(np.uint32), ref["ray"][f].view(np.uint32)), f
    mats = bh.materials_c["type"][isect["primId"]["materialIndex"]]
    live = paths["active"] == 1
    assert live.sum() > 1000 and (isect["hit"][live] == 1).all() and (mats[live] == 1).all()
    assert ((mats == 2) & (isect["hit"] == 1)).any()  # some rays end on the light
