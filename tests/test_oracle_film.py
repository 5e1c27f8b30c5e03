"""Pins the oracle's camera, pixel filter and film (oracle/oracle_film.c): erfinv against erf,
the Gaussian filter's sampling distribution and weight, pinhole geometry, and the film sums."""
import math

import numpy as np
import pytest

import oracle
from gonzales_amd import abi
from gonzales_amd.camera import Film, PerspectiveCamera, cornell_camera, look_at


def test_erfinv_inverts_erf():
    for x in np.linspace(-2.5, 2.5, 201):
        y = math.erf(x)
        if abs(y) >= 1:
            continue
        assert abs(oracle.erfinv(y) - x) <= 2e-5 * max(1.0, abs(x)), x
    assert oracle.erfinv(1.0) == math.inf and oracle.erfinv(-1.0) == -math.inf
    assert math.isnan(oracle.erfinv(1.5))
    assert oracle.erfinv(0.0) == 0.0


def test_gaussian_filter_sampling_follows_truncated_gaussian():
    r, sigma = 1.5, 0.5
    g = np.random.default_rng(1)
    xs = np.array([oracle.gaussian_sample(a, b, r, r, sigma)[0] for a, b in g.random((20000, 2))])
    assert xs.min() >= -r and xs.max() <= r
    edges = np.linspace(-r, r, 13)
    cdf = lambda x: 0.5 * (1 + math.erf(x / (sigma * math.sqrt(2))))
    norm = cdf(r) - cdf(-r)
    want = np.array([(cdf(b) - cdf(a)) / norm for a, b in zip(edges[:-1], edges[1:])])
    hist = np.histogram(xs, bins=edges)[0] / len(xs)
    assert np.allclose(hist, want, atol=0.01)


def test_gaussian_filter_weight_is_constant_inside_support():
    """evaluate / pdf = (2 norm - 1)^2 where the sample is not clamped (normalised Gaussians)."""
    r, sigma = 1.5, 0.5
    norm = 0.5 * (1 + math.erf(r / (sigma * math.sqrt(2))))
    g = np.random.default_rng(2)
    for a, b in g.random((200, 2)):
        dx, dy, pdf = oracle.gaussian_sample(a, b, r, r, sigma)
        if abs(dx) < r and abs(dy) < r:
            assert abs(oracle.gaussian_evaluate(dx, dy, sigma) / pdf - (2 * norm - 1) ** 2) < 1e-4
    g0 = 1 / math.sqrt(2 * math.pi * sigma * sigma)
    assert abs(oracle.gaussian_evaluate(0.0, 0.0, sigma) - g0 * g0) < 1e-6


def _box_camera(res=(64, 48), fov=40.0, eye=(1.0, 2.0, 3.0), target=(0.0, 0.5, -1.0), spp=4):
    film = Film(*res)
    return PerspectiveCamera(look_at(eye, target, (0, 1, 0)), fov, film, spp, filter_type=abi.GZ_FILTER_BOX,
                             filter_radius=(0.5, 0.5))


def test_pinhole_geometry():
    eye, target = np.array([1.0, 2.0, 3.0]), np.array([0.0, 0.5, -1.0])
    cam = _box_camera(eye=eye, target=target)
    paths = oracle.camera_paths(cam.c, 0, cam.path_count, abi.GZ_PATH)
    view = (target - eye) / np.linalg.norm(target - eye)
    o = np.stack([paths["ray"][k] for k in ("orgX", "orgY", "orgZ")], 1)
    d = np.stack([paths["ray"][k] for k in ("dirX", "dirY", "dirZ")], 1).astype(np.float64)
    assert np.allclose(o, eye, atol=1e-5)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-5)
    # film position of each ray: angle from the view axis follows the pinhole model
    # tan(angle) = |(film - centre)| * 2 tan(fov/2) / min(res) for the shorter axis
    w, h = 64, 48
    film_x = paths["pixelX"] + 0.5  # box filter offsets are within half a pixel
    film_y = paths["pixelY"] + 0.5
    cos_a = d @ view
    tan_a = np.sqrt(np.maximum(0, 1 - cos_a ** 2)) / cos_a
    scale = 2 * math.tan(math.radians(40.0) / 2) / min(w, h)
    rad = np.hypot(film_x - w / 2, film_y - h / 2) * scale
    assert np.all(np.abs(tan_a - rad) <= scale * 0.75 + 1e-4)
    # raster x grows to the camera's right, raster y downwards
    right = np.cross((0, 1, 0), view)
    right /= np.linalg.norm(right)
    up = np.cross(view, right)
    assert np.corrcoef(d @ right, film_x)[0, 1] > 0.99
    assert np.corrcoef(d @ up, film_y)[0, 1] < -0.99
    assert np.all(paths["dim"] == 4) and np.all(paths["flags"] == abi.GZ_PATH_ACTIVE)
    assert np.all(paths["filterWeight"] == 1.0)


def test_cornell_camera_centre_ray():
    cam = cornell_camera(res=64, spp=1, filter_type=abi.GZ_FILTER_BOX, filter_radius=(0.0, 0.0))
    paths = oracle.camera_paths(cam.c, 0, cam.path_count, abi.GZ_PATH)
    centre = paths[(paths["pixelX"] == 32) & (paths["pixelY"] == 32)][0]["ray"]
    # pixel 32's centre is at raster 32.5, half a pixel right and below the image centre
    assert abs(centre["orgY"] - 1.0) < 1e-6 and abs(centre["orgZ"] - 6.8) < 1e-5
    assert centre["dirZ"] < -0.999 and abs(centre["dirX"]) < 0.01 and abs(centre["dirY"]) < 0.01


def test_film_sums_and_resolve():
    cam = _box_camera(res=(16, 8), spp=3)
    paths = oracle.camera_paths(cam.c, 0, cam.path_count, abi.GZ_PATH)
    g = np.random.default_rng(3)
    paths["estimate"] = g.random((len(paths), 3)) * 4
    paths["albedo"] = g.random((len(paths), 3))
    paths["firstNormal"] = g.standard_normal((len(paths), 3))
    paths["filterWeight"] = g.uniform(0.5, 1.5, len(paths))
    paths["estimate"][:3] = 0  # pixel 0 black
    n = 16 * 8
    light, albedo, normal = (np.zeros(3 * n, np.float32) for _ in range(3))
    weight = np.zeros(n, np.float32)
    oracle.film_add(cam.c, paths, 100.0, 3.0, light, weight, albedo, normal)
    idx = paths["pixelY"] * 16 + paths["pixelX"]
    est = paths["estimate"].astype(np.float64)
    mx = est.max(1, keepdims=True)
    est = np.where(mx > 3.0, est * (3.0 / np.maximum(mx, 1e-30)), est)
    want = np.zeros((n, 3))
    np.add.at(want, idx, est)
    assert np.allclose(light.reshape(n, 3), want, rtol=1e-5)
    wsum = np.zeros(n)
    np.add.at(wsum, idx, paths["filterWeight"])
    assert np.allclose(weight, wsum, rtol=1e-6)
    nn = paths["firstNormal"] / np.linalg.norm(paths["firstNormal"], axis=1, keepdims=True)
    nsum = np.zeros((n, 3))
    np.add.at(nsum, idx, np.abs(nn))
    assert np.allclose(normal.reshape(n, 3), nsum, rtol=1e-5)
    out, bad = oracle.film_resolve(light, weight)
    assert bad == 0
    assert np.all(out.reshape(n, 3)[0] == 0)
    assert np.allclose(out.reshape(n, 3)[1:], (want / wsum[:, None])[1:], rtol=1e-5)
