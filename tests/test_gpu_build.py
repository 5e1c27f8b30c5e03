"""BVH build on the device (gz_gpu_build_scene) and its BVH2 export (gz_gpu_export_bvh2).

Parity: the oracle's traversal (the port of traverse_bvh2_core) over the exported BVH2 nodes and
the leaf-order PrimIds gives bit-identical Intersection_C records to the device traversal of
the device-built handle (same tree, same visit order). Layout-independent checks against the
host SAH build: the closest-hit distance of every ray is bit-identical (it is the minimum over
all triangles of the same Moller-Trumbore arithmetic) and the hit primitive is the same wherever
no two triangles tie at that distance. Tree invariants: every primitive in exactly one leaf,
child boxes inside the parent's, leaf boxes covering their triangles, leaves of at most maxLeaf
primitives, depth within the traversal stack.
"""
import ctypes

import numpy as np
import pytest

import oracle
from gonzales_amd import abi, scenes
from gonzales_amd.accelerator import (Primitives, Scene, TriangleMesh, build_bounding_hierarchy,
                                      build_bounding_hierarchy_gpu, export_bvh2, pack_prim_ids, sah_cost,
                                      take_records)

pytestmark = pytest.mark.gpu


def _rays(n, lo, hi, seed):
    g = np.random.default_rng(seed)
    rays = np.zeros(n, abi.RAY)
    o = g.uniform(lo, hi, (n, 3))
    d = g.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for k, f in enumerate(("orgX", "orgY", "orgZ")):
        rays[f] = o[:, k]
    for k, f in enumerate(("dirX", "dirY", "dirZ")):
        rays[f] = d[:, k]
    return rays, np.full(n, np.inf, np.float32)


def _oracle(bh, rays, tmax):
    desc, keep = bh.descriptor()
    ref, _ = oracle.traverse(ctypes.addressof(desc), len(bh.nodes), rays, tmax, abi.INTERSECTION, 8)
    return ref


def _soup_scene(n, seed=3):
    mesh = scenes.triangle_soup(n, seed)
    mats = np.zeros(1, abi.MATERIAL)
    mats["type"] = 1
    mats["albedo"] = 0.5
    return Scene([mesh], mats), Primitives.geometric(0, n, 0)


def _check_tree(bh, max_leaf):
    nodes = bh.nodes
    n = len(bh.prim_ids)
    seen = np.zeros(n, np.int64)
    leaves = nodes[nodes["count"] > 0]
    assert np.all(leaves["count"] <= max_leaf)
    for lf in leaves:
        seen[lf["offset"]:lf["offset"] + lf["count"]] += 1
    assert np.all(seen == 1), "every primitive in exactly one leaf"
    assert len(nodes) == 2 * len(leaves) - 1
    box = lambda nd: np.array([nd["boundsMinX"], nd["boundsMinY"], nd["boundsMinZ"],
                               nd["boundsMaxX"], nd["boundsMaxY"], nd["boundsMaxZ"]])
    verts = {}
    for mi, m in enumerate(bh.scene.meshes):
        verts[mi] = m.triangle_vertices()
    depth = np.zeros(len(nodes), np.int64)
    for i, nd in enumerate(nodes):
        b = box(nd)
        if nd["count"] == 0:
            for c in (i + 1, nd["offset"]):
                cb = box(nodes[c])
                assert np.all(cb[:3] >= b[:3]) and np.all(cb[3:] <= b[3:]), "child inside parent"
                depth[c] = depth[i] + 1
        else:
            ids = bh.prim_ids[nd["offset"]:nd["offset"] + nd["count"]]
            for pid in ids:
                mesh = pid["id1"] if pid["type"] == 0 else pid["id2"] >> 32
                tri = pid["id2"] // 3 if pid["type"] == 0 else pid["id2"] & 0xFFFFFFFF
                v = verts[int(mesh)][int(tri)]
                assert np.all(v.min(0) >= b[:3]) and np.all(v.max(0) <= b[3:]), "leaf covers its triangles"
    assert depth.max() < 64
    return depth.max()


def _step(msg):
    print(f"[build test] {msg}", flush=True)


def _same_records(got, ref, what):
    """Intersection records equal byte for byte; on a difference, a short report (not a byte diff)."""
    g = got.view(np.uint8).reshape(len(got), -1)
    r = ref.view(np.uint8).reshape(len(ref), -1)
    bad = np.nonzero((g != r).any(1))[0]
    if len(bad):
        k = bad[:5]
        print(f"{what}: {len(bad)} of {len(got)} records differ; first rays {k.tolist()}", flush=True)
        print("  got hit/t/id2", got["hit"][k].tolist(), got["tHit"][k].tolist(), got["primId"]["id2"][k].tolist())
        print("  ref hit/t/id2", ref["hit"][k].tolist(), ref["tHit"][k].tolist(), ref["primId"]["id2"][k].tolist())
        cols = np.nonzero((g[bad] != r[bad]).any(0))[0]
        print("  differing byte offsets", cols.tolist())
        for f in ("u", "v"):
            print(f"  {f} got", got[f][k].view(np.uint32).tolist(), "ref", ref[f][k].view(np.uint32).tolist())
        print("  id1/mat/type got", got["primId"]["id1"][k].tolist(), got["primId"]["materialIndex"][k].tolist(),
              got["primId"]["type"][k].tolist())
        print("  id1/mat/type ref", ref["primId"]["id1"][k].tolist(), ref["primId"]["materialIndex"][k].tolist(),
              ref["primId"]["type"][k].tolist())
    return len(bad) == 0


@pytest.mark.parametrize("sah", [False, True])
@pytest.mark.parametrize("max_leaf", [1, 2, 4, 8])
def test_device_build_traversal_matches_oracle_on_the_exported_tree(gpu, max_leaf, sah):
    sc, prims = _soup_scene(60000)
    _step(f"device build, 60000 triangles, max_leaf {max_leaf}, sah {sah}")
    bh = build_bounding_hierarchy_gpu(sc, prims, max_leaf=max_leaf, sah=sah)
    _step(f"built: {len(bh.nodes)} nodes; device traversal of 8192 rays")
    rays, tmax = _rays(8192, 0.0, 1.0, 5)
    got = bh.intersect_batch_gpu(rays, tmax)
    _step("oracle traversal on the exported tree")
    ref = _oracle(bh, rays, tmax)
    _step("host upload of the exported tree")
    from gonzales_amd.accelerator import BoundingHierarchy
    up = BoundingHierarchy(sc, take_records(bh.prim_ids), bh.nodes.copy())
    up.prepare_materials()
    up.upload_to_gpu()
    via_upload = up.intersect_batch_gpu(rays, tmax)
    up.free()
    _step("compare")
    ok_upload = _same_records(via_upload, ref, "host upload of the export vs oracle")
    ok_dev = _same_records(got, ref, "device-built handle vs oracle")
    assert ok_upload and ok_dev
    assert np.array_equal(bh.occluded(rays, tmax), ref["hit"].astype(bool))
    info = bh.info()
    # the SAH build puts leaves where the host builder does (at most four primitives here)
    depth = _check_tree(bh, 4 if sah else max_leaf)
    assert info.maxDepth == depth
    assert ref["hit"].mean() > 0.02  # tiny soup triangles: a few percent of the rays hit
    bh.free()


def _node_sets(bh, min_size):
    """Frozensets of (id1, id2) of the primitives under every node of a BVH2 array with more than
    min_size primitives."""
    nodes, ids = bh.nodes, bh.prim_ids
    keys = list(zip(ids["id1"].tolist(), ids["id2"].tolist()))
    out, memo = [], {}

    def under(i):
        if i in memo:
            return memo[i]
        nd = nodes[i]
        if nd["count"] > 0:
            r = keys[nd["offset"]:nd["offset"] + nd["count"]]
        else:
            r = under(i + 1) + under(int(nd["offset"]))
        memo[i] = r
        return r

    import sys
    sys.setrecursionlimit(100000)
    for i in range(len(nodes)):
        u = under(i)
        if len(u) > min_size:
            out.append(frozenset(u))
    return sorted(out, key=lambda f: (len(f), min(f)))


def test_device_sah_build_has_the_host_builds_nodes(gpu):
    """The device SAH build makes the host builder's decisions: every node, leaves included,
    holds the same primitive set in both trees, so the two trees have the same node boxes and SAH
    cost (only the order of the primitives inside a leaf and of the nodes in the array differ)."""
    sc, prims = _soup_scene(30000, seed=12)
    _step("device SAH build and host SAH build, 30000 triangles")
    dev = build_bounding_hierarchy_gpu(sc, prims, sah=True)
    host = build_bounding_hierarchy(sc, prims)
    _step("node sets")
    a, b = _node_sets(dev, 0), _node_sets(host, 0)
    print(f"nodes: device {len(a)}, host {len(b)}", flush=True)
    assert len(dev.nodes) == len(host.nodes)
    assert len(a) == len(b) and a == b
    c_dev, c_host = sah_cost(dev.nodes), sah_cost(host.nodes)
    print(f"SAH cost: device {c_dev:.4f}, host {c_host:.4f}", flush=True)
    assert abs(c_dev - c_host) <= 1e-4 * c_host
    dev.free()


@pytest.mark.parametrize("sah", [False, True])
def test_device_build_equals_host_build_distances(gpu, sah):
    sc, prims = _soup_scene(50000, seed=4)
    dev = build_bounding_hierarchy_gpu(sc, prims, sah=sah)
    host = build_bounding_hierarchy(sc, prims)
    host.upload_to_gpu()
    rays, tmax = _rays(16384, 0.0, 1.0, 6)
    a = dev.intersect_batch_gpu(rays, tmax)
    b = host.intersect_batch_gpu(rays, tmax)
    assert np.array_equal(a["hit"], b["hit"])
    assert a["tHit"].tobytes() == b["tHit"].tobytes()
    same = (a["primId"]["id2"] == b["primId"]["id2"]) & (a["primId"]["id1"] == b["primId"]["id1"])
    # a differing primitive must be a tie: the other triangle is hit at the same distance
    assert same.mean() > 0.999
    c_dev, c_host = sah_cost(dev.nodes), sah_cost(host.nodes)
    print(f"SAH cost: device build {c_dev:.2f}, host SAH build {c_host:.2f}", flush=True)
    assert c_dev < (1.0001 if sah else 2.0) * c_host
    dev.free()
    host.free()


def test_device_build_scenes_and_mis_trace(gpu):
    """The Cornell box and a displaced sphere with area lights: oracle parity on the export, and an
    MIS trace over the device-built handle equals the trace over a host upload of the exported
    tree (same tree, so the same path records)."""
    import torch
    from gonzales_amd.integrator import MisOptions, VolumePathIntegrator, new_paths, pinhole_rays
    for make in (lambda: scenes.cornell_box()[:2], lambda: _sphere_scene()):
        scene, prims = make()
        dev = build_bounding_hierarchy_gpu(scene, prims)
        rays, tmax = _rays(4096, -1.0, 1.5, 7)
        assert _same_records(dev.intersect_batch_gpu(rays, tmax), _oracle(dev, rays, tmax), "scene")
        from gonzales_amd.accelerator import BoundingHierarchy
        up = BoundingHierarchy(scene, take_records(dev.prim_ids), dev.nodes.copy())
        up.prepare_materials()
        up.upload_to_gpu()
        _, cam = scenes.cornell_hierarchy()
        r, px, py = pinhole_rays(cam, 24, 24)
        paths = abi.concat([new_paths(r, px, py, s) for s in range(2)], abi.GZ_PATH)
        opt = MisOptions.for_image(24, 24, 2, max_depth=4)
        out = []
        for bh in (dev, up):
            integ = VolumePathIntegrator(bh, opt, sky=(0.3, 0.3, 0.3))
            d = torch.from_numpy(paths.view(np.uint8).copy()).cuda()
            integ.trace(d)
            torch.cuda.synchronize()
            out.append(d.cpu().numpy().tobytes())
        assert out[0] == out[1]
        dev.free()
        up.free()


def _sphere_scene():
    v, f = scenes.icosphere(3)
    v = v * (1 + 0.05 * scenes.fbm(v, 5))[:, None]
    mesh = TriangleMesh.from_xyz(v.astype(np.float32), f.reshape(-1))
    lv = np.array([[-0.5, 2, -0.5], [0.5, 2, -0.5], [0.5, 2, 0.5], [-0.5, 2, 0.5]], np.float32)
    light = TriangleMesh.from_xyz(lv, np.array([0, 1, 2, 0, 2, 3], np.int64))
    mats = np.zeros(1, abi.MATERIAL)
    mats["type"] = 1
    mats["albedo"] = 0.6
    prims = Primitives.concat([Primitives.geometric(0, len(f), 0), Primitives.area_lights(1, 2, 0)])
    return Scene([mesh, light], mats, light_emission=np.full((2, 3), 5.0, np.float32)), prims


def _raw_build(prim_ids, meshes, max_leaf=4, sah=0, reserved=0):
    lib = abi.load()
    meshes_c = (abi.TriangleMeshC * len(meshes))()
    for i, m in enumerate(meshes):
        meshes_c[i] = m.c_struct()
    pts = np.array([m.points.size for m in meshes], np.int64)
    faces = np.array([m.face_indices.size for m in meshes], np.int64)
    vtx = np.array([m.vertex_indices.size for m in meshes], np.int64)
    mats = np.zeros(1, abi.MATERIAL)
    opt = abi.GzBuildOptions(max_leaf, sah)
    opt.reserved[2] = reserved
    return lib.gz_gpu_build_scene(abi.ptr(prim_ids) if len(prim_ids) else None, len(prim_ids),
                                  ctypes.addressof(meshes_c), len(meshes), abi.ptr(pts), abi.ptr(faces), abi.ptr(vtx),
                                  abi.ptr(mats), 1, ctypes.byref(opt))


@pytest.mark.parametrize("sah", [False, True])
def test_device_build_edge_cases(gpu, sah):
    lib = abi.load()
    # identical triangles (equal Morton codes, flat centroids), a single triangle, a root leaf,
    # skipped primitives
    tri = np.array([[0.2, 0.2, 0.5], [0.8, 0.2, 0.5], [0.5, 0.8, 0.5]], np.float32)
    for count, max_leaf in ((1, 4), (3, 4), (4, 4), (5, 4), (33, 4), (300, 2)):
        mesh = TriangleMesh.from_xyz(np.tile(tri, (count, 1)), np.arange(3 * count, dtype=np.int64))
        sc = Scene([mesh], np.zeros(1, abi.MATERIAL))
        bh = build_bounding_hierarchy_gpu(sc, Primitives.geometric(0, count, 0), max_leaf=max_leaf, sah=sah)
        rays, tmax = _rays(512, 0.3, 0.7, 8)
        # slightly tilted: a zero direction component gives 0 x inf = NaN in that axis' slab, which
        # the reference's select-form slab test turns into a miss of every box
        d = np.array([0.01, 0.02, 1.0]) / np.linalg.norm([0.01, 0.02, 1.0])
        rays["orgZ"] = 0.0
        rays["dirX"], rays["dirY"], rays["dirZ"] = d
        got = bh.intersect_batch_gpu(rays, tmax)
        assert _same_records(got, _oracle(bh, rays, tmax), f"{count} identical triangles"), count
        if not got["hit"].any():
            print(f"{count} identical triangles, max_leaf {max_leaf}: no hit; nodes {bh.nodes[:8].tolist()}", flush=True)
        assert got["hit"].any(), count
        # the SAH build keeps a group of flat centroids as one leaf, as the host builder does
        _check_tree(bh, count if sah else max_leaf)
        bh.free()
    # a skipped primitive (non-triangle shape: id2 == -1) among triangles is never hit
    sc, prims = _soup_scene(1000, seed=9)
    ids = pack_prim_ids(prims)
    ids["type"][::7] = 1
    ids["id2"][::7] = -1
    h = _raw_build(ids, sc.meshes, sah=int(sah))
    assert h, abi.last_error()
    nodes = np.zeros(2 * len(ids), abi.BVH2_NODE)
    order = np.zeros(len(ids), np.int32)
    cnt = lib.gz_gpu_export_bvh2(h, abi.ptr(nodes), len(nodes), abi.ptr(order))
    assert cnt > 0 and sorted(order.tolist()) == list(range(len(ids)))
    lib.gz_gpu_free_scene(h)
    # an empty scene misses everything
    h = _raw_build(np.zeros(0, abi.PRIM_ID), sc.meshes, sah=int(sah))
    assert h
    assert lib.gz_gpu_export_bvh2(h, abi.ptr(nodes), len(nodes), None) == 0
    lib.gz_gpu_free_scene(h)


def test_device_build_rejects_bad_indices(gpu):
    sc, prims = _soup_scene(100, seed=2)
    ids = pack_prim_ids(prims)
    bad = take_records(ids)
    bad["id2"][5] = (0 << 32) | 10_000  # triangle past the mesh
    assert not _raw_build(bad, sc.meshes)
    assert "outside" in abi.last_error()
    bad = take_records(ids)
    bad["id2"][7] = (3 << 32) | 1  # mesh 3 of 1
    assert not _raw_build(bad, sc.meshes)
    assert not _raw_build(ids, sc.meshes, max_leaf=0)
    assert not _raw_build(ids, sc.meshes, max_leaf=65)
    assert not _raw_build(ids, sc.meshes, sah=2)
    assert not _raw_build(ids, sc.meshes, reserved=1)
