"""Pinhole perspective camera, pixel filter and film on the device: the host side of
gz_gpu_camera_paths_device / gz_gpu_film_add_device / gz_gpu_film_resolve_device.

The host builds the reference's camera matrices (Camera/PerspectiveCamera.swift:3-23,
Camera/CameraTransform.swift:3-27, Geometry/Transform.swift:83-95, Api/RenderConfiguration.swift:
84-127) in float64 and rounds them to float32; the kernels apply them as Transform * Point and
Transform * Vector do.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi


def _scale(x, y, z) -> np.ndarray:
    return np.diag([x, y, z, 1.0])


def _translate(x, y, z) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def perspective(fov: float, near: float = 0.01) -> np.ndarray:
    """Transform.makePerspective (Transform.swift:83-95); the reference passes the field of view
    where pbrt has the far plane, which changes only the depth row."""
    far = fov
    persp = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, far / (far - near), -far * near / (far - near)],
                      [0, 0, 1, 0]], np.float64)
    inv_tan = 1.0 / np.tan(np.radians(fov) / 2)
    return _scale(inv_tan, inv_tan, 1) @ persp


def screen_window(res_x: int, res_y: int) -> tuple:
    """makeCamera's default screen window (RenderConfiguration.swift:88-102)."""
    frame = res_x / res_y
    if frame > 1:
        return (-frame, frame, -1.0, 1.0)
    return (-1.0, 1.0, -1.0 / frame, 1.0 / frame)


def raster_to_camera(fov: float, res_x: int, res_y: int, window=None) -> np.ndarray:
    """CameraTransform.init: rasterToCamera = cameraToScreen^-1 * rasterToScreen."""
    x0, x1, y0, y1 = screen_window(res_x, res_y) if window is None else window
    screen_to_raster = _scale(res_x, res_y, 1) @ _scale(1 / (x1 - x0), 1 / (y0 - y1), 1) @ _translate(-x0, -y1, 0)
    return np.linalg.inv(perspective(fov)) @ np.linalg.inv(screen_to_raster)


def look_at(eye, target, up) -> np.ndarray:
    """lookAtTransform (Api/SceneDescription.swift:38-52): the world-to-camera matrix."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    d = target - eye
    d /= np.linalg.norm(d)
    up = up / np.linalg.norm(up)
    right = np.cross(up, d)
    right /= np.linalg.norm(right)
    new_up = np.cross(d, right)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = right, new_up, d, eye
    return np.linalg.inv(m)


def pbrt_transform(values16) -> np.ndarray:
    """A pbrt `Transform [...]` directive (16 numbers, column-major) as a row-major matrix."""
    return np.asarray(values16, np.float64).reshape(4, 4).T


@dataclass
class Film:
    res_x: int
    res_y: int
    iso: float = 100.0
    max_component_value: float = 0.0


class PerspectiveCamera:
    """A pinhole camera with a Gaussian (default support 1.5, sigma 0.5) or box filter."""

    def __init__(self, world_to_camera: np.ndarray, fov: float, film: Film, spp: int, seed: int = 0,
                 filter_type: int = abi.GZ_FILTER_GAUSSIAN, filter_radius=(1.5, 1.5), filter_sigma: float = 0.5,
                 crop=None, window=None):
        self.film = film
        self.spp = int(spp)
        c = abi.GzCamera()
        r2c = raster_to_camera(fov, film.res_x, film.res_y, window).astype(np.float32)
        c2w = np.linalg.inv(np.asarray(world_to_camera, np.float64)).astype(np.float32)
        c.rasterToCamera[:] = r2c.reshape(-1).tolist()
        c.cameraToWorld[:] = c2w.reshape(-1).tolist()
        c.resX, c.resY = film.res_x, film.res_y
        x0, x1, y0, y1 = (0, film.res_x, 0, film.res_y) if crop is None else crop
        c.x0, c.x1, c.y0, c.y1 = x0, x1, y0, y1
        c.spp = self.spp
        c.filterType = filter_type
        c.filterRadius[0], c.filterRadius[1] = filter_radius
        c.filterSigma = filter_sigma
        c.lensRadius = 0.0
        c.focalDistance = 1e6
        z = abi.GzZSobolParams()
        abi.check(abi.load().gz_zsobol_params(self.spp, film.res_x, film.res_y, seed, ctypes.byref(z)),
                  "gz_zsobol_params")
        c.zsSeed, c.zsLog2Spp, c.zsBase4Digits = z.seed, z.log2SamplesPerPixel, z.nBase4Digits
        self.c = c

    @property
    def path_count(self) -> int:
        return (self.c.x1 - self.c.x0) * (self.c.y1 - self.c.y0) * self.spp

    def paths_device(self, d_paths, first: int = 0, count: int | None = None, stream=None) -> None:
        n = self.path_count - first if count is None else count
        abi.check(abi.load().gz_gpu_camera_paths_device(ctypes.byref(self.c), first, n, abi.ptr(d_paths),
                                                         abi.stream_ptr(stream)), "gz_gpu_camera_paths_device")


class DeviceFilm:
    """Beauty, weight, albedo and normal sums of a film in HBM (torch tensors)."""

    def __init__(self, camera: PerspectiveCamera, device="cuda"):
        import torch
        self.camera = camera
        n = camera.film.res_x * camera.film.res_y
        self.pixels = n
        self.light = torch.zeros(n * 3, dtype=torch.float32, device=device)
        self.weight = torch.zeros(n, dtype=torch.float32, device=device)
        self.albedo = torch.zeros(n * 3, dtype=torch.float32, device=device)
        self.normal = torch.zeros(n * 3, dtype=torch.float32, device=device)
        self.out = torch.zeros(n * 3, dtype=torch.float32, device=device)
        self.bad = torch.zeros(1, dtype=torch.int32, device=device)

    def clear(self) -> None:
        for t in (self.light, self.weight, self.albedo, self.normal, self.bad):
            t.zero_()

    def add(self, d_paths, count: int, stream=None) -> None:
        f = self.camera.film
        abi.check(abi.load().gz_gpu_film_add_device(ctypes.byref(self.camera.c), abi.ptr(d_paths), count, f.iso,
                                                     f.max_component_value, abi.ptr(self.light),
                                                     abi.ptr(self.weight), abi.ptr(self.albedo),
                                                     abi.ptr(self.normal), abi.stream_ptr(stream)),
                  "gz_gpu_film_add_device")

    def resolve(self, stream=None):
        abi.check(abi.load().gz_gpu_film_resolve_device(abi.ptr(self.light), abi.ptr(self.weight), self.pixels,
                                                         abi.ptr(self.out), abi.ptr(self.bad),
                                                         abi.stream_ptr(stream)), "gz_gpu_film_resolve_device")
        return self.out

    def resolve_aov(self, which: str, stream=None):
        """The first-hit albedo ("albedo") or normal ("normal") image: each pixel's sum over its
        samples divided by their filter weight, as Film.writeImages normalizes the two before
        the denoiser takes them (Film.swift:41-58). A new (pixels * 3) tensor."""
        import torch
        src = {"albedo": self.albedo, "normal": self.normal}[which]
        out = torch.empty_like(self.out)
        bad = torch.zeros(1, dtype=torch.int32, device=out.device)
        abi.check(abi.load().gz_gpu_film_resolve_device(abi.ptr(src), abi.ptr(self.weight), self.pixels,
                                                         abi.ptr(out), abi.ptr(bad), abi.stream_ptr(stream)),
                  "gz_gpu_film_resolve_device")
        return out

    def write_exrs(self, directory, name: str = "gonzales.exr") -> list[str]:
        """Film.writeImages' three files (Film.swift:40-78): the beauty under `name`, and
        albedo.exr and normal.exr beside it, the auxiliary inputs of a denoiser. The reference
        first runs Open Image Denoise's RT filter on the beauty with those two (Denoiser.swift);
        that library is not part of this build, so the beauty is written as rendered and the
        three files are the hand-off to a denoiser run afterwards. Returns the three paths."""
        import os

        import torch

        from .exr import write_exr
        f = self.camera.film
        shape = (f.res_y, f.res_x, 3)
        images = [(name, self.resolve()), ("albedo.exr", self.resolve_aov("albedo")),
                  ("normal.exr", self.resolve_aov("normal"))]
        torch.cuda.synchronize()
        paths = []
        for fname, img in images:
            path = os.path.join(str(directory), fname)
            write_exr(path, img.cpu().numpy().reshape(shape))
            paths.append(path)
        return paths


def cornell_camera(res: int = 512, spp: int = 16, seed: int = 0, **kw) -> PerspectiveCamera:
    """The Cornell box camera: `Transform [1 -0 -0 -0 -0 1 -0 -0 -0 -0 -1 -0 -0 -1 6.8 1]`,
    fov 19.5 (Scenes/cornell-box.pbrt:2,20), with the reference's default Gaussian filter."""
    m = pbrt_transform([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, 0, 0, -1, 6.8, 1])
    return PerspectiveCamera(m, 19.5, Film(res, res), spp, seed, **kw)
