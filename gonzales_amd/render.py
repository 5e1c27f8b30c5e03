"""Frame rendering on the device: the host loop that replaces TileRenderer/Tile.render
(Sources/libgonzales/Core/TileRenderer.swift, Core/Tile.swift:19-219) for the MIS integrator.

A frame is cut into batches of whole pixels (pixel-major path numbering, see GzCamera); each
batch is: camera paths -> all MIS bounces -> film sums, every step a kernel launch on one stream,
with the path records and the film resident in HBM. Nothing returns to the host until resolve.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import abi
from .camera import DeviceFilm, PerspectiveCamera
from .integrator import VolumePathIntegrator


@dataclass
class FrameRenderer:
    camera: PerspectiveCamera
    integrator: VolumePathIntegrator
    film: DeviceFilm
    batch_paths: int = 1 << 23

    def __post_init__(self):
        import torch
        spp = self.camera.spp
        per = max(spp, (min(self.batch_paths, self.camera.path_count) // spp) * spp)
        self.batch = per
        self.d_paths = torch.empty(per * abi.GZ_PATH.itemsize, dtype=torch.uint8, device=self.film.light.device)

    def render(self, stream=None) -> None:
        """Adds one frame (every path sample of the crop) to the film."""
        n = self.camera.path_count
        first = 0
        while first < n:
            count = min(self.batch, n - first)
            self.camera.paths_device(self.d_paths, first, count, stream)
            self.integrator.trace(self.d_paths, count, stream)
            self.film.add(self.d_paths, count, stream)
            first += count
