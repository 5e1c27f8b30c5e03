"""Host-side helpers of the BVH builds: the SAH cost model (a node visit costs 1/2 of a primitive
test, BoundingHierarchyBuilder.swift:225-249) on trees whose cost follows by hand."""
import numpy as np

from gonzales_amd import abi
from gonzales_amd.accelerator import build_bvh2, primitive_bounds, sah_cost
from gonzales_amd import scenes


def _node(lo, hi, offset, count):
    n = np.zeros(1, abi.BVH2_NODE)
    n["boundsMinX"], n["boundsMinY"], n["boundsMinZ"] = lo
    n["boundsMaxX"], n["boundsMaxY"], n["boundsMaxZ"] = hi
    n["offset"], n["count"] = offset, count
    return n


def test_sah_cost_of_hand_trees():
    leaf = _node((0, 0, 0), (1, 1, 1), 0, 3)
    assert sah_cost(leaf) == 3.0
    # root unit cube, children: half cube (SA 4 of 6) with 2 prims, quarter slab (SA 2.5/6... ) with 1
    root = _node((0, 0, 0), (1, 1, 1), 2, 0)
    a = _node((0, 0, 0), (0.5, 1, 1), 0, 2)
    b = _node((0.5, 0, 0), (1, 1, 0.5), 2, 1)
    tree = np.concatenate([root, a, b])
    sa_a = 2 * (0.5 * 1 + 1 * 1 + 1 * 0.5)
    sa_b = 2 * (0.5 * 1 + 1 * 0.5 + 0.5 * 0.5)
    expect = 0.5 * 1.0 + (sa_a * 2 + sa_b * 1) / 6.0
    assert abs(sah_cost(tree) - expect) < 1e-12


def test_host_sah_build_beats_a_single_leaf():
    sc, prims = scenes.soup_scene(4096, 2)
    nodes, order = build_bvh2(primitive_bounds(sc.meshes, prims))
    assert sorted(order.tolist()) == list(range(4096))
    c = sah_cost(nodes)
    assert 1.0 < c < 0.05 * 4096
