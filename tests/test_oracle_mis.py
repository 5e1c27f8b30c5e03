"""Pins the oracle's MIS integrator (oracle/oracle_mis.c): the reference's known-answer tests for
its sampling routines, and integrals with a known value — direct light from a constant sky and
from a triangle (float64 quadrature, including the reference's MIS-weight behaviour for area
lights), and multi-bounce sky lighting against an independent BSDF-sampling estimator."""
import numpy as np
import pytest

import oracle
from gonzales_amd import abi
from gonzales_amd.integrator import new_paths, pinhole_rays

from mis_helpers import (OracleMis, box_on_plane_scene, brute_force_sky, direct_quadrature, light_over_plane,
                         plane_scene, point_paths, sky_plane_quadrature)


# ---- the reference's KATs (Tests/libgonzalesTests/SamplingTests.swift:9-89) ------------------

def test_power_heuristic_kats():
    assert abs(oracle.power_heuristic(1.0, 1.0) - 0.5) <= 1e-6
    assert oracle.power_heuristic(0.0, 1.0) == 0.0
    assert oracle.power_heuristic(1.0, 0.0) == 0.0
    assert oracle.power_heuristic(0.0, 0.0) == 0.0
    assert oracle.power_heuristic(100.0, 1.0) > 0.99
    assert oracle.power_heuristic(1.0, 100.0) < 0.01
    assert abs(oracle.power_heuristic(2.0, 3.0) - 4.0 / 13.0) <= 1e-6


def test_concentric_disk_kats():
    c = oracle.concentric_disk(0.5, 0.5)
    assert abs(c[0]) <= 1e-6 and abs(c[1]) <= 1e-6
    for u in [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (1.0, 1.0), (0.25, 0.75), (0.8, 0.2)]:
        x, y = oracle.concentric_disk(*u)
        assert x * x + y * y <= 1.0 + 1e-6
    # Shirley-Chiu map: square rings go to circles (r = max(|2u-1|, |2v-1|)) and area is kept,
    # so r^2 of uniform points is uniform on [0, 1)
    u = np.random.default_rng(5).random((20000, 2))
    xy = np.array([oracle.concentric_disk(a, b) for a, b in u], np.float64)
    r = np.linalg.norm(xy, axis=1)
    assert np.allclose(r, np.abs(2 * u - 1).max(axis=1), atol=2e-6)
    hist = np.histogram(r * r, bins=8, range=(0, 1))[0] / len(r)
    assert np.allclose(hist, 1 / 8, atol=0.01)


def test_cosine_hemisphere_kats():
    v = oracle.cosine_hemisphere(0.3, 0.7)
    assert abs(np.linalg.norm(v.astype(np.float64)) - 1.0) <= 1e-5
    for u in [(0.1, 0.1), (0.5, 0.5), (0.9, 0.9), (0.0, 0.0)]:
        assert oracle.cosine_hemisphere(*u)[2] >= 0
    # Malley's method: z = cos(theta) with density 2 z, so E[z] = 2/3 over a uniform grid
    u = np.random.default_rng(6).random((20000, 2))
    z = np.array([oracle.cosine_hemisphere(a, b)[2] for a, b in u], np.float64)
    assert abs(z.mean() - 2.0 / 3.0) < 0.01


@pytest.mark.parametrize("unit", [False, True])
def test_equal_area_sphere_fixed_points(unit):
    assert np.allclose(oracle.equal_area_square_to_sphere(0.5, 0.5, unit), [0, 0, 1])
    assert np.allclose(oracle.equal_area_square_to_sphere(0.0, 0.0, unit), [0, 0, -1])
    assert np.allclose(oracle.equal_area_square_to_sphere(1.0, 0.5, unit), [1, 0, 0], atol=1e-6)
    assert np.allclose(oracle.equal_area_square_to_sphere(0.5, 1.0, unit), [0, 1, 0], atol=1e-6)
    assert np.allclose(oracle.equal_area_square_to_sphere(0.0, 0.5, unit), [-1, 0, 0], atol=1e-6)


def test_equal_area_sphere_unit_is_uniform():
    """With the sqrt(2 - r^2) factor the map is Clarici's equal-area octahedral map: unit vectors,
    uniform on the sphere, so each coordinate is uniform on [-1, 1] (Archimedes)."""
    u = np.random.default_rng(7).random((20000, 2))
    d = np.array([oracle.equal_area_square_to_sphere(a, b, True) for a, b in u], np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5)
    for axis in range(3):
        hist = np.histogram(d[:, axis], bins=8, range=(-1, 1))[0] / len(d)
        assert np.allclose(hist, 1 / 8, atol=0.01), axis


def test_equal_area_sphere_reference_form():
    """The reference's form (InfiniteLight.swift:112-127) has the same z and azimuth but x, y
    shorter by sqrt(2 - r^2), r = sqrt(1 - |z|)."""
    u = np.random.default_rng(8).random((2000, 2))
    for a, b in u:
        ref = oracle.equal_area_square_to_sphere(a, b, False).astype(np.float64)
        uni = oracle.equal_area_square_to_sphere(a, b, True).astype(np.float64)
        assert ref[2] == uni[2]
        r2 = 1 - abs(uni[2])
        assert np.allclose(ref[:2] * np.sqrt(2 - r2), uni[:2], atol=3e-6)


def test_rr_draw_uniform_and_distinct():
    u = np.array([oracle.rr_draw(7, x, 3, 2, b) for x in range(64) for b in range(32)])
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.02
    assert len(np.unique(u)) > 0.99 * len(u)
    assert oracle.rr_draw(7, 1, 2, 3, 4) == oracle.rr_draw(7, 1, 2, 3, 4)
    assert oracle.rr_draw(7, 1, 2, 3, 4) != oracle.rr_draw(8, 1, 2, 3, 4)


# ---- integrals with known values ---------------------------------------------------------------

def _mean(paths):
    e = paths["estimate"].astype(np.float64)
    return e.mean(0), e.std(0) / np.sqrt(len(e))


def test_sky_plane_quadrature_closed_form():
    albedo, sky = np.array([0.5, 0.25, 0.75]), np.array([1.0, 2.0, 3.0])
    assert np.allclose(sky_plane_quadrature(albedo, sky, True), albedo * sky, rtol=1e-5)
    assert np.all(sky_plane_quadrature(albedo, sky, False) < albedo * sky)


@pytest.mark.parametrize("max_depth", [1, 4])
@pytest.mark.parametrize("unit", [True, False])
def test_sky_over_plane(max_depth, unit):
    """A diffuse plane under a constant sky: light sampling (equal-area sphere map, pdf 1/(4 pi))
    and cosine BSDF sampling combined by the power heuristic; bounce rays leave the scene. With
    unit sky directions the answer is albedo * Le; in the reference's form, a float64 quadrature."""
    albedo, sky = (0.5, 0.25, 0.75), (1.0, 2.0, 3.0)
    om = OracleMis(plane_scene(albedo), sky=sky)
    n = 1 << 14
    p = point_paths((0.13, 1.0, 0.37), (0.2, -1.0, 0.1), n, log2spp=14)
    om.trace(p, om.params(max_depth, fix_sky=unit, log2spp=14, digits=7))
    mean, err = _mean(p)
    want = sky_plane_quadrature(albedo, sky, unit)
    if unit:
        assert np.allclose(want, np.array(albedo) * np.array(sky), rtol=1e-5)
    assert np.all(np.abs(mean - want) <= np.maximum(4 * err, 0.005 * want)), (mean, want)
    assert np.all(p["bounce"] <= 1)


@pytest.mark.parametrize("fix", [False, True])
def test_triangle_light_facing_point(fix):
    """Direct light from one triangle facing the shaded point: the light strategy with the
    reference's weight p_l^2 / (p_l^2 + p_b^2) (its BSDF strategy sees the light's back side, so
    adds nothing, with or without the pdf correction). Float64 quadrature over the triangle."""
    albedo, L = (0.5, 0.25, 0.75), (4.0, 2.0, 1.0)
    bh, tri = light_over_plane(True, L, albedo)
    om = OracleMis(bh)
    n = 1 << 14
    point = (0.05, 0.0, 0.1)
    p = point_paths((0.05, 2.0, 0.1 - 1.0), (0.0, -2.0, 1.0), n, log2spp=14)
    om.trace(p, om.params(1, fix=fix, log2spp=14, digits=7))
    mean, err = _mean(p)
    want = direct_quadrature(point, albedo, L, tri, True, fix)
    assert np.all(np.abs(mean - want) <= np.maximum(4 * err, 0.01 * want)), (mean, want)


def test_triangle_light_facing_away_needs_the_pdf_correction():
    """A triangle whose emitting side faces away from the point: the light strategy sees no
    radiance; the BSDF strategy sees it and is weighted by p_b^2 / (p_b^2 + p_l^2) with the pdf
    correction, and by 0 without it (the reference's light pdf of a BSDF direction is 0)."""
    albedo, L = (0.5, 0.25, 0.75), (4.0, 2.0, 1.0)
    bh, tri = light_over_plane(False, L, albedo)
    om = OracleMis(bh)
    n = 1 << 14
    point = (0.05, 0.0, 0.1)
    p = point_paths((0.05, 2.0, 0.1 - 1.0), (0.0, -2.0, 1.0), n, log2spp=14)
    om.trace(p, om.params(1, fix=False, log2spp=14, digits=7))
    assert np.all(p["estimate"] == 0)
    p = point_paths((0.05, 2.0, 0.1 - 1.0), (0.0, -2.0, 1.0), n, log2spp=14)
    om.trace(p, om.params(1, fix=True, log2spp=14, digits=7))
    mean, err = _mean(p)
    want = direct_quadrature(point, albedo, L, tri, False, True)
    assert np.all(want > 0)
    assert np.all(np.abs(mean - want) <= np.maximum(4 * err, 0.02 * want)), (mean, want)


def test_multibounce_sky_matches_brute_force():
    """Box on a plane under a constant sky, up to 4 bounces with Russian roulette: the MIS
    integrator and an independent BSDF-sampling estimator agree on the mean image."""
    sky = (1.0, 1.5, 2.0)
    bh = box_on_plane_scene()
    om = OracleMis(bh, sky=sky)
    cam = {"eye": (0.0, 1.6, 4.0), "look": (0.0, 0.4, 0.0), "up": (0.0, 1.0, 0.0), "fov": 30.0}
    w = h = 16
    rays, px, py = pinhole_rays(cam, w, h)
    spp = 256
    paths = abi.concat([new_paths(rays, px, py, s) for s in range(spp)], abi.GZ_PATH)
    om.trace(paths, om.params(4, fix_sky=True, log2spp=8, digits=8))
    est = paths["estimate"].astype(np.float64).reshape(spp, h * w, 3).mean(0)
    ref = brute_force_sky(bh, sky, np.tile(rays, 256), 4, seed=11).reshape(256, h * w, 3).mean(0)
    m_est, m_ref = est.mean(0), ref.mean(0)
    assert np.allclose(m_est, m_ref, rtol=0.02), (m_est, m_ref)
    # per pixel, loosely (a few hundred samples each)
    assert np.mean(np.abs(est - ref) <= 0.1 * ref + 0.02) > 0.95
    assert paths["bounce"].max() >= 3


def test_dimension_bookkeeping():
    """Each bounce that samples consumes 2 + 3 + 3 path dimensions and 1 light-chooser dimension."""
    bh, tri = light_over_plane(True)
    om = OracleMis(bh)
    p = point_paths((0.05, 2.0, -0.9), (0.0, -2.0, 1.0), 64, log2spp=6, start_dim=5)
    om.trace(p, om.params(3, log2spp=6, digits=3))
    done = p["bounce"]  # bounces that sampled (the last bounce of a path may stop before sampling)
    assert np.all((p["dim"] - 5) % 8 == 0)
    assert np.all(p["lightDim"] == (p["dim"] - 5) // 8)
    assert np.all(done <= 3)
    assert np.all(p["flags"] & abi.GZ_PATH_ACTIVE == 0)
