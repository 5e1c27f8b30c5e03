"""Host-side mirror of the reference's Accelerator API for triangle scenes.

The reference's Swift host (Sources/libgonzales/Accelerators/BoundingHierarchy.swift and
BoundingHierarchyBuilder.swift, Api/SceneDescription.swift:4-36) builds a SAH BVH2 over its
primitives, packs one PrimId per primitive, prepares a Material_C table and hands all of it to
the device through ``mojo_gpu_upload_scene``; render loops then call ``intersectBatchGPU``,
``intersect`` and the shading batch. This module keeps those names and semantics on top of this
library's C-ABI (``gz_*``), with numpy arrays in the reference's byte layouts:

* :class:`TriangleMesh` -- world-space points (4 floats per vertex, w = 1, Triangle.swift:4-88),
  int64 vertex indices, optional int64 face indices;
* :class:`Primitives` -- the primitive list (triangle / geometric primitive / area light);
* :func:`build_bounding_hierarchy` -- the SAH build (C++ restatement of the Swift builder) and
  PrimId packing (BoundingHierarchy.swift:21-58, 137-148);
* :class:`BoundingHierarchy` -- upload (``uploadToGPU``), ``intersect_batch_gpu``, ``intersect``,
  ``occluded``, and device-resident variants taking torch tensors.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import abi

PRIM_TRIANGLE = 0
PRIM_GEOMETRIC = 1
PRIM_TRANSFORMED = 2
PRIM_AREA_LIGHT = 3


@dataclass
class TriangleMesh:
    """TriangleMesh (Shape/Triangle.swift:4-88): points already in world space."""

    points: np.ndarray          # (P, 4) float32, w = 1
    vertex_indices: np.ndarray  # (3T,) int64
    face_indices: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    normals: np.ndarray | None = None  # (P, 3) float32 (shading normals), optional
    uvs: np.ndarray | None = None      # (P, 2) float32, optional

    def __post_init__(self):
        self.points = np.ascontiguousarray(self.points, dtype=np.float32)
        if self.points.ndim != 2 or self.points.shape[1] != 4:
            raise ValueError("points must be (P, 4) float32 (x, y, z, 1)")
        self.vertex_indices = np.ascontiguousarray(self.vertex_indices, dtype=np.int64)
        self.face_indices = np.ascontiguousarray(self.face_indices, dtype=np.int64)
        if self.vertex_indices.size % 3:
            raise ValueError("vertex indices must be a multiple of 3")
        if self.vertex_indices.size and (self.vertex_indices.min() < 0 or
                                         self.vertex_indices.max() >= self.points.shape[0]):
            raise ValueError("vertex index out of range")  # Triangle.swift:626-639

    @staticmethod
    def from_xyz(xyz: np.ndarray, indices: np.ndarray, **kw) -> "TriangleMesh":
        xyz = np.asarray(xyz, dtype=np.float32).reshape(-1, 3)
        pts = np.ones((xyz.shape[0], 4), np.float32)
        pts[:, :3] = xyz
        return TriangleMesh(pts, np.asarray(indices, np.int64).reshape(-1), **kw)

    @property
    def number_triangles(self) -> int:
        return self.vertex_indices.size // 3

    def c_struct(self) -> abi.TriangleMeshC:
        return abi.TriangleMeshC(self.points.ctypes.data,
                                 self.face_indices.ctypes.data if self.face_indices.size else None,
                                 self.vertex_indices.ctypes.data)

    def triangle_vertices(self) -> np.ndarray:
        """(T, 3, 3) float32 vertices of every triangle."""
        return self.points[self.vertex_indices.reshape(-1, 3), :3]


@dataclass
class Primitives:
    """Primitive list in the order the scene produced it.

    kind: PRIM_TRIANGLE | PRIM_GEOMETRIC | PRIM_AREA_LIGHT; mesh/tri: mesh index and triangle
    number; idx: the GeometricPrimitive/AreaLight index (id1); material: material index for
    geometric primitives.
    """

    kind: np.ndarray
    mesh: np.ndarray
    tri: np.ndarray
    idx: np.ndarray
    material: np.ndarray

    def __len__(self) -> int:
        return int(self.kind.size)

    @staticmethod
    def geometric(mesh_index: int, n: int, material: int | np.ndarray, first_idx: int = 0) -> "Primitives":
        mat = np.broadcast_to(np.asarray(material, np.int64), (n,)).copy()
        return Primitives(np.full(n, PRIM_GEOMETRIC, np.int8), np.full(n, mesh_index, np.int64),
                          np.arange(n, dtype=np.int64), np.arange(first_idx, first_idx + n, dtype=np.int64), mat)

    @staticmethod
    def area_lights(mesh_index: int, n: int, first_light: int = 0) -> "Primitives":
        return Primitives(np.full(n, PRIM_AREA_LIGHT, np.int8), np.full(n, mesh_index, np.int64),
                          np.arange(n, dtype=np.int64), np.arange(first_light, first_light + n, dtype=np.int64),
                          np.zeros(n, np.int64))

    @staticmethod
    def concat(parts: list["Primitives"]) -> "Primitives":
        return Primitives(*[np.concatenate([getattr(p, f) for p in parts]) for f in
                            ("kind", "mesh", "tri", "idx", "material")])

    def take(self, order: np.ndarray) -> "Primitives":
        return Primitives(self.kind[order], self.mesh[order], self.tri[order], self.idx[order], self.material[order])


def take_records(a: np.ndarray, index=None) -> np.ndarray:
    """a[index] (or a copy of a) for a structured array, with the padding bytes zero: NumPy copies
    the fields of a dtype with holes and leaves the holes of the new array uninitialised."""
    src = a if index is None else a[index]
    out = np.zeros(len(src), a.dtype)
    for f in a.dtype.names:
        out[f] = src[f]
    return out


def pack_prim_ids(prims: Primitives) -> np.ndarray:
    """PrimId packing of BoundingHierarchy.swift:21-58 (before prepareMaterials)."""
    out = np.zeros(len(prims), abi.PRIM_ID)
    out["type"] = prims.kind
    packed = (prims.mesh << 32) | prims.tri
    tri_kind = prims.kind == PRIM_TRIANGLE
    out["id1"] = np.where(tri_kind, prims.mesh, prims.idx)
    out["id2"] = np.where(tri_kind, 3 * prims.tri, packed)
    out["materialIndex"] = np.where(prims.kind == PRIM_GEOMETRIC, prims.material, 0)
    return out


def primitive_bounds(meshes: list[TriangleMesh], prims: Primitives) -> np.ndarray:
    """Triangle.worldBound (Triangle.swift:249-256): (n, 6) float32 min/max of the 3 vertices."""
    out = np.empty((len(prims), 6), np.float32)
    for m_index, mesh in enumerate(meshes):
        sel = np.nonzero(prims.mesh == m_index)[0]
        if sel.size == 0:
            continue
        v = mesh.triangle_vertices()[prims.tri[sel]]
        out[sel, :3] = v.min(axis=1)
        out[sel, 3:] = v.max(axis=1)
    return out


def build_bvh2(bounds: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """SAH BVH2 over primitive bounds (BoundingHierarchyBuilder.swift:66-339, via gz_bvh2_build).

    Returns (nodes as BVH2_NODE array, order) where BVH position k holds primitive order[k].
    """
    lib = abi.load()
    bounds = np.ascontiguousarray(bounds, dtype=np.float32)
    n = bounds.shape[0]
    nodes = np.zeros(max(2 * n - 1, 1), abi.BVH2_NODE)
    order = np.zeros(max(n, 1), np.int64)
    count = lib.gz_bvh2_build(abi.ptr(bounds), n, abi.ptr(nodes), abi.ptr(order))
    if count < 0:
        abi.check(int(count), "gz_bvh2_build")
    return nodes[:count].copy(), order[:n].copy()


@dataclass
class Scene:
    """The parts of Core/Scene.swift the hot path reads: meshes, materials, area-light radiance."""

    meshes: list[TriangleMesh]
    materials: np.ndarray  # MATERIAL table for the geometric primitives' material indices
    light_emission: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))

    def meshes_c(self):
        arr = (abi.TriangleMeshC * max(len(self.meshes), 1))()
        for i, m in enumerate(self.meshes):
            arr[i] = m.c_struct()
        return arr


class BoundingHierarchy:
    """BoundingHierarchy (Accelerators/BoundingHierarchy.swift): BVH2 nodes + PrimIds, GPU handle."""

    def __init__(self, scene: Scene, prim_ids: np.ndarray, nodes: np.ndarray):
        self.scene = scene
        self.prim_ids = np.ascontiguousarray(prim_ids)
        self.nodes = np.ascontiguousarray(nodes)
        self.materials_c = None
        self.handle = None

    # prepareMaterials (BoundingHierarchy.swift:91-150): geometric materials, then one emissive
    # Material_C per area light; area-light PrimIds point at baseMatIndex + lightIdx.
    def prepare_materials(self) -> None:
        if self.materials_c is not None:
            return
        mats = np.asarray(self.scene.materials, abi.MATERIAL)
        lights = np.zeros(len(self.scene.light_emission), abi.MATERIAL)
        lights["type"] = 2
        lights["emission"] = self.scene.light_emission
        self.materials_c = abi.concat([mats, lights], abi.MATERIAL)
        base = len(mats)
        sel = self.prim_ids["type"] == PRIM_AREA_LIGHT
        self.prim_ids["materialIndex"][sel] = base + self.prim_ids["id1"][sel]

    def upload_to_gpu(self) -> None:
        """uploadToGPU (BoundingHierarchy.swift:152-202) through gz_gpu_upload_scene; a no-op when
        the hierarchy already has a handle (built on the device, or uploaded before)."""
        if self.handle is not None:
            return
        lib = abi.load()
        self.prepare_materials()
        meshes = self.scene.meshes
        meshes_c = self.scene.meshes_c()
        pts = np.array([m.points.size for m in meshes] or [0], np.int64)
        faces = np.array([m.face_indices.size for m in meshes] or [0], np.int64)
        vtx = np.array([m.vertex_indices.size for m in meshes] or [0], np.int64)
        h = lib.gz_gpu_upload_scene(abi.ptr(self.nodes), len(self.nodes), abi.ptr(self.prim_ids), len(self.prim_ids),
                                    ctypes.addressof(meshes_c), len(meshes), abi.ptr(pts), abi.ptr(faces), abi.ptr(vtx),
                                    abi.ptr(self.materials_c), len(self.materials_c))
        if not h:
            raise abi.GzError(abi.GZ_E_INVALID, "gz_gpu_upload_scene: " + abi.last_error())
        self.handle = h

    def free(self) -> None:
        if self.handle:
            abi.load().gz_gpu_free_scene(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self) -> abi.GzSceneInfo:
        info = abi.GzSceneInfo()
        abi.check(abi.load().gz_gpu_scene_info(self.handle, ctypes.byref(info)), "gz_gpu_scene_info")
        return info

    # intersectBatchGPU (BoundingHierarchy.swift:242-275): host rays -> Intersection_C records.
    def intersect_batch_gpu(self, rays: np.ndarray, t_hits: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, abi.RAY)
        t_hits = np.ascontiguousarray(t_hits, np.float32)
        out = np.zeros(len(rays), abi.INTERSECTION)
        lib = abi.load()
        lib.gz_gpu_traverse_batch(self.handle, abi.ptr(rays), abi.ptr(t_hits), len(rays), abi.ptr(out))
        err = abi.last_error()
        if err:
            raise abi.GzError(abi.GZ_E_HIP, "gz_gpu_traverse_batch: " + err)
        return out

    # intersect(scene:ray:tHit:) -> SurfaceInteraction? (BoundingHierarchy.swift:387-456), record form.
    def intersect(self, ray, t_hit: float):
        rays = np.zeros(1, abi.RAY)
        rays[0] = tuple(ray)
        r = self.intersect_batch_gpu(rays, np.array([t_hit], np.float32))[0]
        return r if r["hit"] else None

    # intersect(scene:ray:tHit:) -> Bool (BoundingHierarchy.swift:348-384), batched.
    def occluded(self, rays: np.ndarray, t_max: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, abi.RAY)
        t_max = np.ascontiguousarray(t_max, np.float32)
        out = np.zeros(len(rays), np.uint8)
        abi.check(abi.load().gz_gpu_occluded_batch(self.handle, abi.ptr(rays), abi.ptr(t_max), len(rays),
                                                    abi.ptr(out)), "gz_gpu_occluded_batch")
        return out.astype(bool)

    # Device-resident forms: torch tensors in HBM, asynchronous on the current torch stream.
    def closest_device(self, rays_t, tmax_t, hits_t, stats_t=None, stream=None) -> None:
        lib = abi.load()
        n = int(tmax_t.numel())
        if stats_t is None:
            rc = lib.gz_gpu_traverse_closest_device(self.handle, abi.ptr(rays_t), abi.ptr(tmax_t), n,
                                                    abi.ptr(hits_t), abi.stream_ptr(stream))
        else:
            rc = lib.gz_gpu_traverse_closest_stats_device(self.handle, abi.ptr(rays_t), abi.ptr(tmax_t), n,
                                                          abi.ptr(hits_t), abi.ptr(stats_t), abi.stream_ptr(stream))
        abi.check(rc, "closest_device")

    def any_device(self, rays_t, tmax_t, occ_t, stats_t=None, stream=None) -> None:
        lib = abi.load()
        n = int(tmax_t.numel())
        if stats_t is None:
            rc = lib.gz_gpu_traverse_any_device(self.handle, abi.ptr(rays_t), abi.ptr(tmax_t), n, abi.ptr(occ_t),
                                                abi.stream_ptr(stream))
        else:
            rc = lib.gz_gpu_traverse_any_stats_device(self.handle, abi.ptr(rays_t), abi.ptr(tmax_t), n,
                                                      abi.ptr(occ_t), abi.ptr(stats_t), abi.stream_ptr(stream))
        abi.check(rc, "any_device")

    def set_trace_config(self, persistent: int = 3, lds_stack: int = 16, refill_lanes: int = 32,
                         grid_scale: int = 1, chunk: int = 128, bin_min_rays: int = 1 << 18,
                         record_times: int = 0) -> None:
        cfg = abi.GzTraceConfig(persistent, lds_stack, refill_lanes, grid_scale, chunk, bin_min_rays, record_times)
        abi.check(abi.load().gz_gpu_set_trace_config(self.handle, ctypes.byref(cfg)), "gz_gpu_set_trace_config")

    def trace_times(self, stream=None) -> tuple[float, float]:
        """(binning ms, traversal kernel ms) of the last traversal on `stream` (needs record_times)."""
        b, t = ctypes.c_float(), ctypes.c_float()
        abi.check(abi.load().gz_gpu_trace_times(self.handle, abi.stream_ptr(stream), ctypes.byref(b), ctypes.byref(t)),
                  "gz_gpu_trace_times")
        return b.value, t.value

    def trace_config(self) -> abi.GzTraceConfig:
        cfg = abi.GzTraceConfig()
        abi.check(abi.load().gz_gpu_get_trace_config(self.handle, ctypes.byref(cfg)), "gz_gpu_get_trace_config")
        return cfg

    def check_traversal(self, stream=None) -> None:
        abi.check(abi.load().gz_gpu_check_traversal(self.handle, abi.stream_ptr(stream)), "gz_gpu_check_traversal")

    def descriptor(self) -> tuple[abi.SceneDescriptor2C, object]:
        """SceneDescriptor2_C over the host arrays (for CPU consumers of the same layout)."""
        meshes_c = self.scene.meshes_c()
        desc = abi.SceneDescriptor2C(abi.ptr(self.nodes) if len(self.nodes) else None, abi.ptr(self.prim_ids),
                                     ctypes.addressof(meshes_c), len(self.scene.meshes))
        return desc, meshes_c


def build_bounding_hierarchy(scene: Scene, prims: Primitives) -> BoundingHierarchy:
    """makeAccelerator (Api/SceneDescription.swift:4-36) for a triangle scene: SAH build, then the
    PrimIds in BVH order (BoundingHierarchyBuilder.getBoundingHierarchy, :102-108)."""
    bounds = primitive_bounds(scene.meshes, prims)
    nodes, order = build_bvh2(bounds)
    ids = pack_prim_ids(prims.take(order))
    bh = BoundingHierarchy(scene, ids, nodes)
    bh.prepare_materials()
    return bh


def build_bounding_hierarchy_gpu(scene: Scene, prims: Primitives, max_leaf: int = 4,
                                 sah: bool = True) -> BoundingHierarchy:
    """The same hierarchy object built on the device (gz_gpu_build_scene): with sah (the default)
    the host builder's binned-SAH tree, its decisions made top-down on the device (max_leaf does
    not apply: leaves are where the host builder puts them); otherwise a Morton-order linear BVH
    with leaves of at most max_leaf primitives. The handle is ready for every device entry point;
    the tree is also read back in the BVH2 layout (gz_gpu_export_bvh2) with the PrimIds in leaf
    order, for host consumers."""
    lib = abi.load()
    bh = BoundingHierarchy(scene, pack_prim_ids(prims), np.zeros(0, abi.BVH2_NODE))
    bh.prepare_materials()
    meshes = scene.meshes
    meshes_c = scene.meshes_c()
    pts = np.array([m.points.size for m in meshes] or [0], np.int64)
    faces = np.array([m.face_indices.size for m in meshes] or [0], np.int64)
    vtx = np.array([m.vertex_indices.size for m in meshes] or [0], np.int64)
    opt = abi.GzBuildOptions(max_leaf, 1 if sah else 0)
    h = lib.gz_gpu_build_scene(abi.ptr(bh.prim_ids), len(bh.prim_ids), ctypes.addressof(meshes_c), len(meshes),
                               abi.ptr(pts), abi.ptr(faces), abi.ptr(vtx), abi.ptr(bh.materials_c),
                               len(bh.materials_c), ctypes.byref(opt))
    if not h:
        raise abi.GzError(abi.GZ_E_INVALID, "gz_gpu_build_scene: " + abi.last_error())
    bh.handle = h
    n = len(bh.prim_ids)
    nodes = np.zeros(max(2 * n, 1), abi.BVH2_NODE)
    order = np.zeros(max(n, 1), np.int32)
    count = lib.gz_gpu_export_bvh2(h, abi.ptr(nodes), len(nodes), abi.ptr(order))
    if count < 0:
        abi.check(int(count), "gz_gpu_export_bvh2")
    bh.nodes = nodes[:count].copy()
    bh.prim_ids = take_records(bh.prim_ids, order[:n])
    return bh


def export_bvh2(bh: BoundingHierarchy) -> tuple[np.ndarray, np.ndarray]:
    """(BVH2 nodes, leaf-order -> caller index) of a handle (gz_gpu_export_bvh2)."""
    n = len(bh.prim_ids)
    nodes = np.zeros(max(2 * n, 1), abi.BVH2_NODE)
    order = np.zeros(max(n, 1), np.int32)
    count = abi.load().gz_gpu_export_bvh2(bh.handle, abi.ptr(nodes), len(nodes), abi.ptr(order))
    if count < 0:
        abi.check(int(count), "gz_gpu_export_bvh2")
    return nodes[:count].copy(), order[:n].copy()


def sah_cost(nodes: np.ndarray, traversal: float = 0.5, intersect: float = 1.0) -> float:
    """Surface-area cost of a BVH2 in the host builder's model (a node visit costs 1/2 of a
    primitive test, BoundingHierarchyBuilder.swift:225-249): sum over interior nodes of
    SA(node) / SA(root) x traversal + over leaves of SA(leaf) / SA(root) x count x intersect."""
    lo = np.stack([nodes["boundsMinX"], nodes["boundsMinY"], nodes["boundsMinZ"]], 1).astype(np.float64)
    hi = np.stack([nodes["boundsMaxX"], nodes["boundsMaxY"], nodes["boundsMaxZ"]], 1).astype(np.float64)
    d = np.maximum(hi - lo, 0.0)
    sa = 2 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
    leaf = nodes["count"] > 0
    return float((sa[~leaf].sum() * traversal + (sa[leaf] * nodes["count"][leaf]).sum() * intersect) / sa[0])
