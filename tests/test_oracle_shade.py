"""Pins the CPU oracle's diffuse wavefront shade (restating kernel.mojo:99-118, 676-773).

Checks: PCG32 against O'Neill's published pcg32 demo output (seed 42, sequence 54), then the
shade of single paths against closed forms -- the cosine-hemisphere direction built from the
first two PCG32 draws in Duff et al.'s orthonormal basis (JCGT 2017), the offset origin
o + d t + 1e-4 n, the albedo product, emission accumulation and deactivation rules.
"""
import ctypes

import numpy as np

import oracle
from gonzales_amd import abi
from gonzales_amd.accelerator import TriangleMesh

# pcg32-demo.c of the PCG reference implementation prints these for pcg32_srandom_r(42, 54).
PCG32_42_54 = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_known_answer():
    assert oracle.pcg32(42, 54, 6).tolist() == PCG32_42_54


def _scene():
    # one triangle in the z = 0 plane (normal +z), a diffuse material and a light material
    mesh = TriangleMesh.from_xyz(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.arange(3))
    meshes = (abi.TriangleMeshC * 1)(mesh.c_struct())
    mats = np.zeros(3, abi.MATERIAL)
    mats[0]["type"] = 1
    mats[0]["albedo"] = (0.5, 0.25, 0.75)
    mats[1]["type"] = 2
    mats[1]["emission"] = (17, 12, 4)
    mats[2]["type"] = 0
    return mesh, meshes, mats


def _path(o, d, state=42, inc=54):
    p = np.zeros(1, abi.PATH_STATE)
    p["ray"]["orgX"], p["ray"]["orgY"], p["ray"]["orgZ"] = o
    p["ray"]["dirX"], p["ray"]["dirY"], p["ray"]["dirZ"] = d
    p["throughput"] = (1.0, 0.5, 2.0)
    p["estimate"] = (0.1, 0.2, 0.3)
    p["pcgState"], p["pcgInc"], p["active"] = state, inc, 1
    return p


def _isect(material, t=1.0, hit=1):
    i = np.zeros(1, abi.INTERSECTION)
    i["primId"]["id1"], i["primId"]["id2"], i["primId"]["type"] = 0, 0, 1  # mesh 0, triangle 0
    i["primId"]["materialIndex"] = material
    i["tHit"], i["hit"] = t, hit
    return i


def _duff_basis(n):
    s = 1.0 if n[2] >= 0 else -1.0
    a = -1.0 / (s + n[2])
    b = n[0] * n[1] * a
    return np.array([1 + s * n[0] * n[0] * a, s * b, -s * n[0]]), np.array([b, s + n[1] * n[1] * a, -n[1]])


def test_diffuse_bounce_closed_form():
    _, meshes, mats = _scene()
    for d, n in (((0, 0, -1), (0, 0, 1)), ((0, 0, 1), (0, 0, -1))):
        p = _path((0.25, 0.25, 1.0 if d[2] < 0 else -1.0), d)
        oracle.shade(p, _isect(0), ctypes.addressof(meshes), mats)
        u1 = (PCG32_42_54[0] >> 8) / 2.0**24
        u2 = (PCG32_42_54[1] >> 8) / 2.0**24
        n = np.array(n, float)
        t, b = _duff_basis(n)
        local = np.array([np.sqrt(u1) * np.cos(2 * np.pi * u2), np.sqrt(u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1)])
        want = local[0] * t + local[1] * b + local[2] * n
        got = np.array([p["ray"][k][0] for k in ("dirX", "dirY", "dirZ")], float)
        assert np.abs(got - want).max() < 2e-6
        org = np.array([p["ray"][k][0] for k in ("orgX", "orgY", "orgZ")], float)
        assert np.abs(org - (np.array([0.25, 0.25, 0.0]) + 1e-4 * n)).max() < 1e-6
        assert np.allclose(p["throughput"][0], [0.5, 0.125, 1.5]) and p["active"][0] == 1
        assert np.allclose(p["estimate"][0], [0.1, 0.2, 0.3])


def test_pcg_state_advances_by_two_draws():
    _, meshes, mats = _scene()
    p = _path((0.25, 0.25, 1.0), (0, 0, -1), state=42, inc=54)
    oracle.shade(p, _isect(0), ctypes.addressof(meshes), mats)
    p2 = _path((0.25, 0.25, 1.0), (0, 0, -1), state=int(p["pcgState"][0]), inc=54)
    # Re-seeding from the advanced state gives a different draw: the state moved.
    assert int(p["pcgState"][0]) != 42
    mult, inc = 6364136223846793005, (54 << 1) | 1
    s = 0
    s = (s * mult + inc) % 2**64
    s = (s + 42) % 2**64
    for _ in range(3):
        s = (s * mult + inc) % 2**64
    assert int(p["pcgState"][0]) == s
    del p2


def test_random_paths_leave_in_the_normal_hemisphere():
    _, meshes, mats = _scene()
    g = np.random.default_rng(3)
    for k in range(500):
        d = g.standard_normal(3)
        d /= np.linalg.norm(d)
        p = _path((0.2, 0.2, 0.0), tuple(d.astype(np.float32)), state=int(g.integers(0, 2**63)),
                  inc=int(g.integers(0, 2**63)))
        oracle.shade(p, _isect(0, t=0.0), ctypes.addressof(meshes), mats)
        out = np.array([p["ray"][c][0] for c in ("dirX", "dirY", "dirZ")], float)
        n = np.array([0, 0, -1.0 if d[2] > 0 else 1.0])
        assert abs(np.linalg.norm(out) - 1) < 1e-6 and out @ n >= -1e-7


def test_emission_miss_and_inactive_paths():
    _, meshes, mats = _scene()
    p = _path((0.25, 0.25, 1.0), (0, 0, -1))
    oracle.shade(p, _isect(1), ctypes.addressof(meshes), mats)  # area light: two-sided, stop
    assert np.allclose(p["estimate"][0], [0.1 + 17, 0.2 + 6, 0.3 + 8]) and p["active"][0] == 0
    p = _path((0.25, 0.25, 1.0), (0, 0, -1))
    oracle.shade(p, _isect(0, hit=0), ctypes.addressof(meshes), mats)  # miss
    assert p["active"][0] == 0 and np.allclose(p["estimate"][0], [0.1, 0.2, 0.3])
    p = _path((0.25, 0.25, 1.0), (0, 0, -1))
    oracle.shade(p, _isect(2), ctypes.addressof(meshes), mats)  # unknown material type: stop
    assert p["active"][0] == 0
    p = _path((0.25, 0.25, 1.0), (0, 0, -1))
    p["active"] = 0
    before = p.tobytes()
    oracle.shade(p, _isect(0), ctypes.addressof(meshes), mats)
    assert p.tobytes() == before
