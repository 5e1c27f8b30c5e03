"""Wavefront MIS path integrator on the GPU: the bounce of VolumePathIntegrator
(Sources/libgonzales/Integrator/VolumePathIntegrator.swift:35-463) over GzPath records in HBM.

A path record (``abi.GZ_PATH``) carries its ray, throughput, running estimate, first-hit albedo and
normal, pixel and sample index, and the Z-Sobol dimensions of its two samplers (the path sampler
and the light chooser's, which always runs at Morton index 0 as PowerLightSampler's own sampler
does). ``bounce`` advances every active path by one bounce; ``trace`` runs them to completion.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi
from .accelerator import PRIM_AREA_LIGHT, BoundingHierarchy


# The reference's default conductor, namedSpectra "metal-Cu-eta" / "metal-Cu-k" reduced to RGB by
# PiecewiseLinearSpectrum.asRgb (values at 630, 532 and 465 nm; Core/Spectrum.swift:23-28, 81-160).
CU_ETA = (0.230813, 1.04, 1.155312)
CU_K = (3.458187, 2.583, 2.469187)


def _pair(r) -> tuple:
    return (float(r), float(r)) if np.isscalar(r) else (float(r[0]), float(r[1]))


def diffuse_material(reflectance) -> np.ndarray:
    m = np.zeros(1, abi.GZ_MATERIAL)
    m["type"] = abi.GZ_MAT_DIFFUSE
    m["reflectance"] = reflectance
    return m


def conductor_material(eta=CU_ETA, k=CU_K, roughness=0.5) -> np.ndarray:
    """Conductor.create defaults (Conductor.swift:25-40): Cu, roughness 0.5, not remapped."""
    m = np.zeros(1, abi.GZ_MATERIAL)
    m["type"] = abi.GZ_MAT_CONDUCTOR
    m["reflectance"] = 1.0
    m["roughness"] = _pair(roughness)
    m["conductorEta"] = eta
    m["conductorK"] = k
    return m


def dielectric_material(eta=1.5, roughness=0.0, remap=True) -> np.ndarray:
    """Dielectric.create defaults (Dielectric.swift:61-78): eta 1.5, roughness 0, remapped."""
    m = np.zeros(1, abi.GZ_MATERIAL)
    m["type"] = abi.GZ_MAT_DIELECTRIC
    m["flags"] = abi.GZ_MAT_REMAP_ROUGHNESS if remap else 0
    m["eta"] = eta
    m["roughness"] = _pair(roughness)
    return m


def coated_conductor_material(reflectance=(1.0, 1.0, 1.0), interface_roughness=0.5, conductor_roughness=0.5,
                              eta=1.5, thickness=0.01, albedo=(0.0, 0.0, 0.0), g=0.0, conductor_eta=CU_ETA,
                              conductor_k=CU_K, max_depth=10, n_samples=1, remap=True) -> np.ndarray:
    """CoatedConductor.create defaults (CoatedConductor.swift:91-153)."""
    m = np.zeros(1, abi.GZ_MATERIAL)
    m["type"] = abi.GZ_MAT_COATED_CONDUCTOR
    m["flags"] = abi.GZ_MAT_REMAP_ROUGHNESS if remap else 0
    m["reflectance"] = reflectance
    m["eta"] = eta
    m["roughness"] = _pair(interface_roughness)
    m["conductorRoughness"] = _pair(conductor_roughness)
    m["conductorEta"] = conductor_eta
    m["conductorK"] = conductor_k
    m["thickness"] = thickness
    m["albedo"] = albedo
    m["g"] = g
    m["maxDepth"] = max_depth
    m["nSamples"] = n_samples
    return m


def materials_from_scene(bh: BoundingHierarchy) -> np.ndarray:
    """GzMaterial table for the geometric primitives' material indices (diffuse only, as the
    MIS kernel supports today). Area-light surfaces shade with entry 0, as the reference does."""
    src = np.asarray(bh.scene.materials, abi.MATERIAL)
    out = np.zeros(max(len(src), 1), abi.GZ_MATERIAL)
    out["type"] = abi.GZ_MAT_DIFFUSE
    if len(src):
        if np.any(src["type"] != 1):
            raise ValueError("only diffuse materials (type 1) are supported by the MIS kernel")
        out["reflectance"] = src["albedo"]
    return out


def area_lights_from_scene(bh: BoundingHierarchy) -> np.ndarray:
    """One GzAreaLight per area-light primitive, in light-index order (Scene.areaLights)."""
    ids = bh.prim_ids
    sel = np.nonzero(ids["type"] == PRIM_AREA_LIGHT)[0]
    out = np.zeros(len(sel), abi.GZ_AREA_LIGHT)
    if len(sel) == 0:
        return out
    light = ids["id1"][sel]
    order = np.argsort(light, kind="stable")
    light = light[order]
    if not np.array_equal(light, np.arange(len(light))):
        raise ValueError("area-light indices must be 0..n-1, each once")
    id2 = ids["id2"][sel][order]
    out["mesh"] = (id2 >> 32).astype(np.int32)
    out["triangle"] = id2 & 0xFFFFFFFF
    out["brightness"] = np.asarray(bh.scene.light_emission, np.float32)[light]
    return out


def new_paths(rays: np.ndarray, px, py, sample_index, start_dim: int = 0, light_dim: int = 0) -> np.ndarray:
    """Fresh GzPath records: unit throughput, zero estimate, active, bounce 0."""
    rays = np.asarray(rays, abi.RAY)
    n = len(rays)
    p = np.zeros(n, abi.GZ_PATH)
    p["ray"] = rays
    p["throughput"] = 1.0
    p["filterWeight"] = 1.0
    p["pixelX"] = px
    p["pixelY"] = py
    p["sampleIndex"] = sample_index
    p["dim"] = start_dim
    p["lightDim"] = light_dim
    p["flags"] = abi.GZ_PATH_ACTIVE
    return p


@dataclass
class MisOptions:
    max_depth: int = 5
    fix_light_pdf: bool = False  # intersect the chosen triangle for the BSDF ray's light pdf
    fix_sky_direction: bool = False  # unit-length sky sample directions
    zsobol_seed: int = 0
    log2_spp: int = 0
    base4_digits: int = 0
    rr_seed: int = 0x5EED

    @classmethod
    def for_image(cls, width: int, height: int, spp: int, seed: int = 1234, **kw) -> "MisOptions":
        """Sampler constants of ZSobolSampler(samplesPerPixel:fullResolution:seed:)."""
        z = abi.GzZSobolParams()
        abi.check(abi.load().gz_zsobol_params(spp, width, height, seed, ctypes.byref(z)), "gz_zsobol_params")
        return cls(zsobol_seed=z.seed, log2_spp=z.log2SamplesPerPixel, base4_digits=z.nBase4Digits, **kw)

    def params(self) -> abi.GzMisParams:
        flags = (abi.GZ_MIS_FIX_LIGHT_PDF if self.fix_light_pdf else 0) | \
            (abi.GZ_MIS_FIX_SKY_DIRECTION if self.fix_sky_direction else 0)
        return abi.GzMisParams(self.max_depth, flags,
                               self.zsobol_seed, self.log2_spp, self.base4_digits, self.rr_seed)


class VolumePathIntegrator:
    """GPU MIS integrator over an uploaded BoundingHierarchy (surfaces, area lights, constant sky)."""

    def __init__(self, bh: BoundingHierarchy, options: MisOptions, sky=None, materials=None, lights=None):
        if bh.handle is None:
            raise ValueError("upload the hierarchy first (BoundingHierarchy.upload_to_gpu)")
        self.bh = bh
        self.options = options
        self.materials = np.ascontiguousarray(materials_from_scene(bh) if materials is None else materials,
                                              abi.GZ_MATERIAL)
        self.lights = np.ascontiguousarray(area_lights_from_scene(bh) if lights is None else lights,
                                           abi.GZ_AREA_LIGHT)
        self.sky = None if sky is None else np.ascontiguousarray(sky, np.float32).reshape(3)
        self._params = options.params()
        abi.check(abi.load().gz_gpu_set_mis_scene(
            bh.handle, abi.ptr(self.materials), len(self.materials),
            abi.ptr(self.lights) if len(self.lights) else None, len(self.lights),
            abi.ptr(self.sky) if self.sky is not None else None), "gz_gpu_set_mis_scene")

    def bounce(self, paths_t, count: int | None = None, stream=None) -> None:
        """One bounce of every active path (device tensor of GzPath records, in place)."""
        n = int(paths_t.numel() * paths_t.element_size() // abi.GZ_PATH.itemsize) if count is None else count
        abi.check(abi.load().gz_gpu_mis_bounce_device(self.bh.handle, abi.ptr(paths_t), n,
                                                       ctypes.byref(self._params), abi.stream_ptr(stream)),
                  "gz_gpu_mis_bounce_device")

    def set_launch(self, queued: int = 1, block: int = 256, grid_scale: int = 1, chunk: int = 256,
                   drain_level: int = 0) -> None:
        """Launch configuration of trace/trace_stats and the adjoint on this scene handle (the
        results never depend on it)."""
        c = abi.GzMisLaunch(queued, block, grid_scale, chunk, drain_level)
        abi.check(abi.load().gz_gpu_set_mis_launch(self.bh.handle, ctypes.byref(c)), "gz_gpu_set_mis_launch")

    def launch(self) -> tuple[int, int, int, int, int]:
        """(queued, block, grid_scale, chunk, drain_level) of this scene handle."""
        c = abi.GzMisLaunch()
        abi.check(abi.load().gz_gpu_get_mis_launch(self.bh.handle, ctypes.byref(c)), "gz_gpu_get_mis_launch")
        return c.queued, c.block, c.gridScale, c.chunk, c.drainLevel

    def release_queues(self) -> None:
        """Frees the path-queue storage kept on this scene handle for its largest launch; the
        next trace allocates it again."""
        abi.check(abi.load().gz_gpu_release_mis_queues(self.bh.handle), "gz_gpu_release_mis_queues")

    def trace(self, paths_t, count: int | None = None, stream=None) -> None:
        """All bounces of every active path (maxDepth + 1 levels), in place."""
        n = int(paths_t.numel() * paths_t.element_size() // abi.GZ_PATH.itemsize) if count is None else count
        abi.check(abi.load().gz_gpu_mis_trace_device(self.bh.handle, abi.ptr(paths_t), n,
                                                      ctypes.byref(self._params), abi.stream_ptr(stream)),
                  "gz_gpu_mis_trace_device")

    def trace_stats(self, paths_t, stats_t, count: int | None = None, stream=None) -> None:
        """``trace`` that also adds each path's traversal work to stats_t (abi.PATH_STATS records,
        zeroed by the caller)."""
        n = int(paths_t.numel() * paths_t.element_size() // abi.GZ_PATH.itemsize) if count is None else count
        abi.check(abi.load().gz_gpu_mis_trace_stats_device(self.bh.handle, abi.ptr(paths_t), n,
                                                            ctypes.byref(self._params), abi.ptr(stats_t),
                                                            abi.stream_ptr(stream)),
                  "gz_gpu_mis_trace_stats_device")


def pinhole_rays(camera: dict, width: int, height: int, jitter: np.ndarray | None = None) -> tuple:
    """Pinhole camera rays through pixel centres (or jittered positions), one per pixel, row-major.
    Returns (rays, px, py)."""
    eye = np.asarray(camera["eye"], np.float64)
    look = np.asarray(camera["look"], np.float64)
    up = np.asarray(camera["up"], np.float64)
    fwd = look - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    true_up = np.cross(right, fwd)
    scale = np.tan(np.radians(camera["fov"]) / 2)
    py, px = np.mgrid[0:height, 0:width]
    px = px.ravel().astype(np.int32)
    py = py.ravel().astype(np.int32)
    jx, jy = (0.5, 0.5) if jitter is None else (jitter[:, 0], jitter[:, 1])
    aspect = width / height
    sx = (2 * (px + jx) / width - 1) * scale * (aspect if aspect > 1 else 1)
    sy = (1 - 2 * (py + jy) / height) * scale / (aspect if aspect < 1 else 1)
    d = fwd[None, :] + sx[:, None] * right[None, :] + sy[:, None] * true_up[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(len(px), abi.RAY)
    rays["orgX"], rays["orgY"], rays["orgZ"] = eye
    rays["dirX"], rays["dirY"], rays["dirZ"] = d[:, 0], d[:, 1], d[:, 2]
    return rays, px, py
