"""Writes the committed golden vectors under tests/golden/ from the CPU oracle.

The reference (Swift + Mojo) cannot be built or run here, and its own tests hold no vectors for
traversal or Z-Sobol values, so these fixtures are produced by the oracle after it has been pinned
(tests/test_oracle_*.py); the traversal fixtures are additionally cross-checked against a
brute-force closest hit before they are written. They freeze the oracle's answers so every later
GPU run is compared with the same bits.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE.parent))

import oracle  # noqa: E402
from gonzales_amd import abi, scenes  # noqa: E402
from helpers import OracleView, random_rays, random_soup_hierarchy  # noqa: E402


def zsobol_cases():
    g = np.random.default_rng(2024)
    cases = {}
    pix4 = np.array([[0, 0], [5, 5], [10, 10], [1919, 1079]])
    for seed in (0, 1234):
        px = np.repeat(pix4[:, 0], 16)
        py = np.repeat(pix4[:, 1], 16)
        si = np.tile(np.arange(16), 4)
        cases[f"dims16_seed{seed}"] = (16, 1920, 1080, seed, 0, [1] * 16, px, py, si)
    px = g.integers(0, 512, 256)
    py = g.integers(0, 512, 256)
    si = g.integers(0, 64, 256)
    cases["bounce_pattern_spp64"] = (64, 512, 512, 1234, 0, [2, 2, 2, 3, 2, 3, 2, 3, 1], px, py, si)
    si8 = g.integers(0, 8, 256)
    cases["odd_log2spp_spp8"] = (8, 512, 512, 0, 0, [2, 3, 1, 2], px, py, si8)
    cases["wrap_from_1020"] = (16, 64, 64, 7, 1020, [3, 3, 2, 1], px % 64, py % 64, si % 16)
    cases["non_pow2_spp12"] = (12, 100, 60, 3, 0, [2, 2], px % 100, py % 60, si % 4)
    return cases


def write_zsobol(path: Path):
    table = oracle.sobol_matrices()
    out = {}
    for name, (spp, rx, ry, seed, start, pat, px, py, si) in zsobol_cases().items():
        vals = oracle.zsobol_batch(spp, rx, ry, seed, px, py, si, pat, start_dim=start, table=table)
        out[name + "_args"] = np.array([spp, rx, ry, seed, start], np.int64)
        out[name + "_pattern"] = np.array(pat, np.int8)
        out[name + "_px"] = np.asarray(px, np.int32)
        out[name + "_py"] = np.asarray(py, np.int32)
        out[name + "_si"] = np.asarray(si, np.int64)
        out[name + "_out"] = vals
    np.savez_compressed(path, **out)


def write_traversal(path: Path):
    out = {}
    for name, (bh, rays, tmax) in {
        "soup300": (random_soup_hierarchy(300, 77, scale=0.15),) + random_rays(4096, 78),
        "cornell": (scenes.cornell_hierarchy()[0],) + cornell_rays(4096),
    }.items():
        view = OracleView(bh)
        res, stats = view.closest(rays, tmax, nthreads=1, stats=True)
        occ, _ = view.occluded(rays, tmax, nthreads=1)
        bi, bt = view.brute(rays, tmax)
        assert np.array_equal(res["tHit"].view(np.uint32), bt.view(np.uint32)), name
        assert np.array_equal(res["hit"].astype(bool), bi >= 0) and np.array_equal(occ, bi >= 0), name
        out[name + "_nodes"] = bh.nodes
        out[name + "_prims"] = bh.prim_ids
        out[name + "_materials"] = bh.materials_c
        for k, m in enumerate(bh.scene.meshes):
            out[f"{name}_mesh{k}_points"] = m.points
            out[f"{name}_mesh{k}_vtx"] = m.vertex_indices
        out[name + "_rays"] = rays
        out[name + "_tmax"] = tmax
        out[name + "_isect"] = res
        out[name + "_occluded"] = occ.astype(np.uint8)
        out[name + "_stats"] = stats
    np.savez_compressed(path, **out)


def cornell_rays(n: int):
    """Camera-like rays from the Cornell eye towards random points in the box, plus rays from
    random interior points in random directions."""
    g = np.random.default_rng(79)
    rays = np.zeros(n, abi.RAY)
    half = n // 2
    eye = np.array([0.0, 1.0, 6.8])
    tgt = g.uniform([-1, 0, -1], [1, 2, 1], (half, 3))
    d1 = tgt - eye
    o2 = g.uniform([-0.9, 0.1, -0.9], [0.9, 1.9, 0.9], (n - half, 3))
    d2 = g.standard_normal((n - half, 3))
    o = np.concatenate([np.tile(eye, (half, 1)), o2])
    d = np.concatenate([d1, d2])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for k, f in enumerate(("orgX", "orgY", "orgZ")):
        rays[f] = o[:, k].astype(np.float32)
    for k, f in enumerate(("dirX", "dirY", "dirZ")):
        rays[f] = d[:, k].astype(np.float32)
    tmax = np.where(np.arange(n) % 3 == 0, np.float32(np.inf), g.uniform(0.5, 10.0, n)).astype(np.float32)
    return rays, tmax


def main():
    write_zsobol(HERE / "zsobol_vectors.npz")
    write_traversal(HERE / "traversal_vectors.npz")
    print("wrote", HERE / "zsobol_vectors.npz", HERE / "traversal_vectors.npz")


if __name__ == "__main__":
    main()
