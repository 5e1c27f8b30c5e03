"""Many scenes rendered in one device launch sequence (``gz_batch_*`` in include/gonzales_amd.h).

The reference renders one scene per process, tile by tile (Core/TileRenderer.swift:77-107,
Core/Tile.swift:19-219). A ``SceneBatch`` takes several uploaded scenes, each with its own MIS
integrator (materials, lights, sky, parameters), camera and device film, and renders every path
of every scene with one fixed launch sequence per round of path slots: camera rays, the bounce
levels over device path queues, film. The sequence is captured once as a HIP graph and replayed
per round; the films match what ``FrameRenderer`` gives for each scene alone, bit for bit.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import abi
from .camera import DeviceFilm, PerspectiveCamera
from .integrator import VolumePathIntegrator


@dataclass
class BatchEntry:
    integrator: VolumePathIntegrator
    camera: PerspectiveCamera
    film: DeviceFilm


class SceneBatch:
    """Scenes rendered together (gz_batch_*): the scenes with at most `solo_paths` paths are
    packed into rounds of at most `round_paths` path slots, each round one launch sequence (a
    replayed HIP graph); a scene with more paths fills the GPU alone and renders with the
    single-scene launch sequence, in four lanes (after the rounds on the render's stream, and
    beside them on the batch's three side streams), so that one scene's level tails overlap the
    other lanes' work. solo_paths = 0 puts every scene in the rounds."""

    def __init__(self, entries: list[BatchEntry], round_paths: int = 1 << 25, solo_paths: int = abi.GZ_BATCH_SOLO_PATHS):
        if not entries:
            raise ValueError("a batch needs at least one scene")
        self.entries = list(entries)
        arr = (abi.GzBatchScene * len(entries))()
        for k, e in enumerate(entries):
            f = e.camera.film
            arr[k].handle = e.integrator.bh.handle
            arr[k].camera = e.camera.c
            arr[k].params = e.integrator.options.params()
            arr[k].iso = f.iso
            arr[k].maxComponentValue = f.max_component_value
            arr[k].d_light = abi.ptr(e.film.light)
            arr[k].d_weight = abi.ptr(e.film.weight)
            arr[k].d_albedo = abi.ptr(e.film.albedo)
            arr[k].d_normal = abi.ptr(e.film.normal)
        self._scenes = arr
        h = abi.load().gz_batch_create(arr, len(entries), int(round_paths), int(solo_paths))
        if not h:
            raise RuntimeError(f"gz_batch_create: {abi.last_error()}")
        self.handle = h

    def info(self) -> abi.GzBatchInfo:
        out = abi.GzBatchInfo()
        abi.check(abi.load().gz_batch_info(self.handle, ctypes.byref(out)), "gz_batch_info")
        return out

    def set_launch(self, block: int = 256, grid_scale: int = 1, chunk: int = 256, drain_level: int = 3,
                   queued: int = 1) -> None:
        """Launch shape of the batch's bounce levels: queued 1 = one persistent bounce kernel per
        level, 3 = the full wavefront per level (closest hits sorted by BSDF class, shade kernels,
        shadow and BSDF rays traced by their own kernels, finish kernels); drain_level > 0 runs
        the bounce levels from that one on in a single launch (a lane per path)."""
        c = abi.GzMisLaunch(queued, block, grid_scale, chunk, drain_level)
        abi.check(abi.load().gz_batch_set_launch(self.handle, ctypes.byref(c)), "gz_batch_set_launch")

    def clear_films(self) -> None:
        for e in self.entries:
            e.film.clear()

    def render(self, use_graph: bool = True, stream=None) -> None:
        """Adds one frame of every scene to its film (asynchronous on `stream`)."""
        abi.check(abi.load().gz_batch_render_device(self.handle, 1 if use_graph else 0, abi.stream_ptr(stream)),
                  "gz_batch_render_device")

    def check(self, stream=None) -> None:
        """Waits for `stream`; raises if a traversal of the batch's paths overflowed its stack since
        the last check (gz_batch_check)."""
        abi.check(abi.load().gz_batch_check(self.handle, abi.stream_ptr(stream)), "gz_batch_check")

    def close(self) -> None:
        if getattr(self, "handle", None):
            abi.load().gz_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass(frozen=True)
class BatchLaunch:
    block: int = 256
    grid_scale: int = 1
    chunk: int = 256
    drain_level: int = 3
    queued: int = 1


BATCH_CANDIDATES = (
    BatchLaunch(256, 1, 64, 3, 3),
    BatchLaunch(256, 1, 256, 3, 3),
    BatchLaunch(256, 1, 64, 2, 3),
    BatchLaunch(256, 1, 64),
    BatchLaunch(256, 1, 256),
    BatchLaunch(256, 1, 1024),
    BatchLaunch(256, 2, 256),
    BatchLaunch(128, 1, 256),
    BatchLaunch(256, 1, 256, 0),
)


def autotune_batch(batch: SceneBatch, candidates=BATCH_CANDIDATES, reps: int = 1) -> tuple[BatchLaunch, dict]:
    """Times whole renders of the batch (graph mode) under each launch shape, keeps the fastest
    and clears the films afterwards. Returns (best, {shape: ms})."""
    import torch

    s = torch.cuda.current_stream()
    timings = {}
    for cfg in candidates:
        batch.set_launch(cfg.block, cfg.grid_scale, cfg.chunk, cfg.drain_level, cfg.queued)
        batch.render(True)  # captures the graph for this shape
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(reps):
            batch.render(True)
        b.record(s)
        torch.cuda.synchronize()
        timings[cfg] = a.elapsed_time(b) / reps
    best = min(timings, key=timings.get)
    batch.set_launch(best.block, best.grid_scale, best.chunk, best.drain_level, best.queued)
    batch.clear_films()
    return best, timings


__all__ = ["BatchEntry", "SceneBatch", "BatchLaunch", "BATCH_CANDIDATES", "autotune_batch"]
