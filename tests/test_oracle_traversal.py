"""Pins the CPU oracle's traversal (restating kernel.mojo:122-364) and the host SAH builder.

The oracle is trusted only after these checks: closed-form triangle and box cases, Moller-Trumbore
invariants, the reference's own box known-answer cases (BoundsTests.swift:120-180) where the
Mojo slab test agrees with the Swift one, a brute-force closest hit over every primitive (BVH
layout independent), and any-hit == "a closest hit exists".
"""
import numpy as np
import pytest

import oracle
from gonzales_amd import abi
from gonzales_amd.accelerator import build_bvh2
from helpers import OracleView, f32_bits, random_rays, random_soup_hierarchy

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")

TRI = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


def test_triangle_closed_form_hit():
    hit, t, u, v = oracle.intersect_triangle([0.25, 0.25, 1], [0, 0, -1], *TRI, 10.0)
    assert hit and (t, u, v) == (1.0, 0.25, 0.25)


def test_triangle_back_face_hits():
    hit, t, u, v = oracle.intersect_triangle([0.25, 0.25, -1], [0, 0, 1], *TRI, 10.0)
    assert hit and (t, u, v) == (1.0, 0.25, 0.25)


def test_triangle_tmax_is_inclusive():
    # kernel.mojo:153 rejects t > tMax only: t == tMax is a hit.
    assert oracle.intersect_triangle([0.25, 0.25, 1], [0, 0, -1], *TRI, 1.0)[0]
    assert not oracle.intersect_triangle([0.25, 0.25, 1], [0, 0, -1], *TRI, float(np.nextafter(np.float32(1), 0)))[0]


def test_triangle_behind_and_parallel_miss():
    assert not oracle.intersect_triangle([0.25, 0.25, 1], [0, 0, 1], *TRI, 10.0)[0]  # t < 0
    assert not oracle.intersect_triangle([0.25, 0.25, 1], [1, 0, 0], *TRI, 10.0)[0]  # det = 0


def test_triangle_absolute_det_epsilon():
    # |det| < 1e-7 is rejected even for a geometrically valid hit (a quirk of kernel.mojo:141).
    small = ([0, 0, 0], [1e-4, 0, 0], [0, 1e-4, 0])  # det = 1e-8
    assert not oracle.intersect_triangle([2.5e-5, 2.5e-5, 1], [0, 0, -1], *small, 10.0)[0]
    big = ([0, 0, 0], [1e-3, 0, 0], [0, 1e-3, 0])  # det = 1e-6
    assert oracle.intersect_triangle([2.5e-4, 2.5e-4, 1], [0, 0, -1], *big, 10.0)[0]


def test_triangle_edges_inclusive():
    hit, t, u, v = oracle.intersect_triangle([0.0, 0.5, 1], [0, 0, -1], *TRI, 10.0)  # on edge p0-p2
    assert hit and u == 0.0 and v == 0.5
    assert not oracle.intersect_triangle([-0.01, 0.5, 1], [0, 0, -1], *TRI, 10.0)[0]


def test_triangle_random_hits_satisfy_invariants():
    g = np.random.default_rng(5)
    n_hit = 0
    for _ in range(3000):
        p = g.standard_normal((3, 3)).astype(np.float32)
        o = (g.standard_normal(3) * 3).astype(np.float32)
        target = (p.T @ g.dirichlet([1, 1, 1])).astype(np.float32)
        d = (target - o).astype(np.float32)
        d /= np.float32(np.linalg.norm(d))
        hit, t, u, v = oracle.intersect_triangle(o, d, p[0], p[1], p[2], 100.0)
        if not hit:
            continue
        n_hit += 1
        assert 0.0 <= u <= 1.0 and 0.0 <= v and u + v <= 1.0 and 0.0 < t <= 100.0
        # Exact solution of o + t d = p0 + u (p1 - p0) + v (p2 - p0) in float64.
        p64, o64, d64 = p.astype(np.float64), o.astype(np.float64), d.astype(np.float64)
        e1, e2 = p64[1] - p64[0], p64[2] - p64[0]
        n = np.cross(e1, e2)
        if abs(np.dot(n, d64)) < 0.2 * np.linalg.norm(n):
            continue  # grazing rays amplify rounding; the check targets gross mistakes
        tt, uu, vv = np.linalg.solve(np.stack([-d64, e1, e2], axis=1), o64 - p64[0])
        assert abs(u - uu) < 1e-4 and abs(v - vv) < 1e-4 and abs(t - tt) < 1e-4 * (1 + tt)
    assert n_hit > 2500


BOX = [-1, -1, -1, 1, 1, 1]


def _n(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def test_box_known_answers_shared_with_swift_slab_test():
    # BoundsTests.swift:130-173, the cases without a zero direction component.
    inf = float("inf")
    assert not oracle.intersect_aabb(BOX, [0, 5, 0], _n([0.01, 1, 0.01]), inf)[0]
    assert not oracle.intersect_aabb(BOX, [0, 0, 5], _n([0.001, 0.001, 1]), inf)[0]
    assert not oracle.intersect_aabb(BOX, [0, 0, 5], _n([0.001, 0.001, -1]), 3.0)[0]
    assert oracle.intersect_aabb(BOX, [0, 0, 5], _n([0.001, 0.001, -1]), 10.0)[0]
    assert oracle.intersect_aabb(BOX, [-5, -5, -5], _n([1, 1, 1]), inf)[0]


def test_box_entry_distance_closed_form():
    d = _n([1, 0.1, -0.1])
    hit, tn = oracle.intersect_aabb([0, 0, 0, 1, 1, 1], [-2, 0.5, 0.5], d, 10.0)
    assert hit and abs(tn - 2.0 / float(d[0])) < 1e-5
    hit, tn = oracle.intersect_aabb([0, 0, 0, 1, 1, 1], [0.5, 0.5, 0.5], d, 10.0)
    assert hit and tn == 0.0  # origin inside: tNear clamped to 0


def test_box_slack_admits_tmax_just_short_of_entry():
    # tFar = min(..., tMax) * 1.0000003: a box entered at t slightly past tMax still counts.
    d = _n([1, 0.25, 0.125])
    _, tn = oracle.intersect_aabb([0, -5, -5, 1, 5, 5], [-2, 0, 0], d, 100.0)
    assert oracle.intersect_aabb([0, -5, -5, 1, 5, 5], [-2, 0, 0], d, tn / 1.0000002)[0]
    assert not oracle.intersect_aabb([0, -5, -5, 1, 5, 5], [-2, 0, 0], d, tn / 1.000001)[0]


def test_box_axis_aligned_ray_nan_quirk():
    # kernel.mojo:233-239 computes org * (1/dir): a zero direction component with a zero origin
    # component gives 0 * inf = NaN, and the select-form max/min (kernel.mojo:193-204) then make
    # the box test fail. The Swift slab test (BoundsTests.swift:121-128, 163-170) reports these
    # rays as hits; the Mojo traversal the reference renders with does not.
    assert not oracle.intersect_aabb(BOX, [0, 0, 5], [0, 0, -1], float("inf"))[0]
    assert not oracle.intersect_aabb(BOX, [0, 0, 0], [1, 0, 0], float("inf"))[0]
    # Origin inside the slab of a zero-direction axis: (-1)*inf - (-0.5*inf) = NaN in tNear. The
    # chain `tNear = tNearX > tNearY ? tNearX : tNearY` drops a NaN in its first operand but keeps
    # one in its second, so a zero x component still hits and a zero y component misses.
    hit_x, tn = oracle.intersect_aabb(BOX, [-0.5, -0.2, 5], _n([0, 0.1, -1]), float("inf"))
    assert hit_x and abs(tn - 4.0 * np.sqrt(1.01)) < 1e-5
    assert not oracle.intersect_aabb(BOX, [-0.2, -0.5, 5], _n([0.1, 0, -1]), float("inf"))[0]


@pytest.mark.parametrize("n_tris,seed", [(1, 0), (2, 1), (7, 2), (64, 3), (500, 4), (3000, 5)])
def test_traversal_matches_brute_force(n_tris, seed):
    bh = random_soup_hierarchy(n_tris, seed, scale=0.2)
    view = OracleView(bh)
    rays, tmax = random_rays(3000, seed + 100)
    res, _ = view.closest(rays, tmax)
    bi, bt = view.brute(rays, tmax)
    hit = res["hit"].astype(bool)
    assert np.array_equal(hit, bi >= 0)
    assert np.array_equal(f32_bits(res["tHit"]), f32_bits(bt))
    # Hit identity: the same primitive, or a different one at exactly the same t (a tie).
    ids_trav = res["primId"]["id2"][hit]
    ids_brute = bh.prim_ids["id2"][bi[hit]]
    assert np.mean(ids_trav == ids_brute) > 0.999
    assert hit.sum() > 0 or n_tris < 8


def test_traversal_miss_record_layout():
    bh = random_soup_hierarchy(50, 9)
    view = OracleView(bh)
    rays, tmax = random_rays(500, 10, lo=5.0, hi=6.0, tmax=(0.1, 0.2))  # far away: all miss
    res, _ = view.closest(rays, tmax)
    assert not res["hit"].any()
    assert (res["primId"]["id1"] == -1).all() and (res["primId"]["id2"] == -1).all()
    assert np.array_equal(f32_bits(res["tHit"]), f32_bits(tmax)) and not res["u"].any() and not res["v"].any()


def test_any_hit_equals_closest_hit_exists():
    bh = random_soup_hierarchy(2000, 11, scale=0.1)
    view = OracleView(bh)
    rays, tmax = random_rays(5000, 12)
    res, st_c = view.closest(rays, tmax, stats=True)
    occ, st_a = view.occluded(rays, tmax, stats=True)
    assert np.array_equal(occ, res["hit"].astype(bool))
    # The early exit never does more work than the closest-hit traversal.
    assert (st_a["interior"] <= st_c["interior"]).all() and (st_a["tris"] <= st_c["tris"]).all()
    assert occ.mean() > 0.1


def test_single_threaded_equals_threaded():
    bh = random_soup_hierarchy(1000, 13)
    view = OracleView(bh)
    rays, tmax = random_rays(9000, 14)
    a, sa = view.closest(rays, tmax, nthreads=1, stats=True)
    b, sb = view.closest(rays, tmax, nthreads=8, stats=True)
    assert a.tobytes() == b.tobytes() and sa.tobytes() == sb.tobytes()


# ---- host SAH builder (product code: gz_bvh2_build) ------------------------------------------

def _leaf_ranges(nodes):
    leaves = nodes[nodes["count"] > 0]
    return leaves["offset"], leaves["count"]


def _check_bvh(bounds, nodes, order):
    n = len(bounds)
    assert sorted(order.tolist()) == list(range(n))
    # depth-first layout: left child = i + 1, right child index > i; every node reached once
    seen = np.zeros(len(nodes), int)
    stack = [0]
    while stack:
        i = stack.pop()
        seen[i] += 1
        if nodes["count"][i] == 0:
            r = nodes["offset"][i]
            assert i + 1 < len(nodes) and i < r < len(nodes)
            stack += [r, i + 1]
    assert (seen == 1).all()
    off, cnt = _leaf_ranges(nodes)
    cover = np.zeros(n, int)
    for o, c in zip(off, cnt):
        cover[o:o + c] += 1
    assert (cover == 1).all()  # every primitive in exactly one leaf
    # node bounds == union of its primitives' bounds (exactly)
    sb = bounds[order]

    def span(i):
        if nodes["count"][i] > 0:
            return nodes["offset"][i], nodes["offset"][i] + nodes["count"][i]
        a, _ = span(i + 1)
        _, b = span(nodes["offset"][i])
        return a, b

    for i in range(len(nodes)):
        a, b = span(i)
        nb = np.array([nodes[f][i] for f in ("boundsMinX", "boundsMinY", "boundsMinZ")], np.float32)
        xb = np.array([nodes[f][i] for f in ("boundsMaxX", "boundsMaxY", "boundsMaxZ")], np.float32)
        assert np.array_equal(nb, sb[a:b, :3].min(0)) and np.array_equal(xb, sb[a:b, 3:].max(0))
    # leaves hold at most 4 primitives unless the centroids cannot be split
    for o, c in zip(off, cnt):
        if c > 4:
            cen = 0.5 * sb[o:o + c, :3] + 0.5 * sb[o:o + c, 3:]
            ext = cen.max(0) - cen.min(0)
            assert ext.max() == 0.0 or np.all((sb[o:o + c, 3:] - sb[o:o + c, :3]) == 0)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 300, 5000])
def test_builder_invariants(n):
    g = np.random.default_rng(n)
    lo = g.random((n, 3)).astype(np.float32)
    hi = (lo + 0.05 * g.random((n, 3))).astype(np.float32)
    bounds = np.concatenate([lo, hi], axis=1)
    nodes, order = build_bvh2(bounds)
    _check_bvh(bounds, nodes, order)
    if n == 1:
        assert len(nodes) == 1 and nodes["count"][0] == 1


def test_builder_two_primitives_sorted_by_centre():
    # range.count <= 2: sorted by centre on the longest centroid axis, split in the middle.
    bounds = np.array([[2, 0, 0, 3, 1, 1], [0, 0, 0, 1, 1, 1]], np.float32)
    nodes, order = build_bvh2(bounds)
    assert len(nodes) == 3 and nodes["count"][0] == 0 and list(order) == [1, 0]


def test_builder_identical_centroids_make_one_leaf():
    bounds = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (10, 1))
    nodes, order = build_bvh2(bounds)
    assert len(nodes) == 1 and nodes["count"][0] == 10


def test_builder_sah_prefers_separated_clusters():
    # Two far-apart clusters of 8 tiny boxes: the root split must separate them.
    g = np.random.default_rng(3)
    a = g.random((8, 3)) * 0.01
    b = g.random((8, 3)) * 0.01 + [10, 0, 0]
    lo = np.concatenate([a, b]).astype(np.float32)
    bounds = np.concatenate([lo, lo + 0.001], axis=1).astype(np.float32)
    nodes, order = build_bvh2(bounds)
    left = set(order[:8].tolist())
    assert left in ({*range(8)}, {*range(8, 16)})
    _check_bvh(bounds, nodes, order)


def test_builder_rejects_bad_arguments():
    lib = abi.load()
    assert lib.gz_bvh2_build(None, -1, None, None) == abi.GZ_E_INVALID
    assert lib.gz_bvh2_build(None, 0, None, None) == 0
