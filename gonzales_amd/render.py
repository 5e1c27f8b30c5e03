"""Frame rendering on the device: the host loop that replaces TileRenderer/Tile.render
(Sources/libgonzales/Core/TileRenderer.swift, Core/Tile.swift:19-219) for the MIS integrator.

A frame is cut into batches of whole pixels (pixel-major path numbering, see GzCamera); each
batch is: camera paths -> all MIS bounces -> film sums, every step a kernel launch on its lane's
stream, with the path records and the film resident in HBM. Nothing returns to the host until
resolve.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import abi
from .camera import DeviceFilm, PerspectiveCamera
from .integrator import VolumePathIntegrator


def lane_plan(path_count: int, spp: int, lanes: int = 0, batch_paths: int = 1 << 25) -> tuple[list[int], int]:
    """FrameRenderer's cut of a frame's paths (pixel-major, spp per pixel) into lanes: the lane
    bounds (lane l renders paths [bounds[l], bounds[l + 1]), whole pixels, equal pixel counts up to
    one) and the paths per launch sequence. lanes = 0: two lanes from 4M paths (2^22) on, else
    one; never more lanes than pixels."""
    pixels = path_count // spp
    n = int(lanes) if lanes > 0 else (2 if path_count >= 1 << 22 else 1)
    n = max(1, min(n, pixels))
    bounds = [(pixels * l // n) * spp for l in range(n + 1)]
    widest = max(b - a for a, b in zip(bounds, bounds[1:])) if path_count else spp
    return bounds, max(spp, (min(batch_paths // n, widest) // spp) * spp)


@dataclass
class FrameRenderer:
    """One frame through the MIS integrator in launch sequences of whole pixels. Each sequence
    holds about 285 B of device state per path (112 B here, the rest in the scene handle's queue
    storage, one per stream); larger sequences are faster, since every sequence pays each bounce
    level's tail (the chip draining the level's last waves): in one lane the 5M-triangle 1080p x
    64 spp frame takes 259 ms in 32M-path sequences and 241 ms as one sequence of 132.7M paths
    (38 GB).

    The frame's pixels are cut into `lanes` contiguous ranges, each rendered on its own stream
    (lane 0 on the caller's stream, the others on streams of the renderer's own that start after
    the work queued on the caller's stream and that the caller's stream waits for at the end), in
    sequences of at most batch_paths / lanes paths with a path buffer per lane: one range's level
    tails overlap the other's work. The 5M-triangle frame with the same device state: 229-230 ms
    as two 66M-path lanes against 241 ms as one sequence, 246 ms as two lanes of 16M-path
    sequences against 259 ms as one lane of 32M-path sequences; the Cornell box at 512^2 x 16 spp
    (4.2M paths, the queued megakernel) 3.64 against 3.84 ms. lanes = 0 (the default) takes two
    lanes for frames of at least 4M paths and one below: smaller frames do not fill the chip with
    one lane's kernels, and their persistent kernels only queue behind each other (the Cornell box
    at 256^2 x 16 spp: 2.41 ms in two lanes against 1.52 in one). The lanes add to disjoint pixels of the film, and every path's
    result is independent of its sequence, so the film is bit-identical for any lane count and
    batch size. The lanes take equal pixel counts: cutting them so that they end together (by
    measured lane times) made the 5M-triangle frame slower (235 against 228-230 ms), as the two
    lanes' level tails then fall together instead of overlapping the other lane's work."""
    camera: PerspectiveCamera
    integrator: VolumePathIntegrator
    film: DeviceFilm
    batch_paths: int = 1 << 25
    lanes: int = 0

    def __post_init__(self):
        import torch
        self.bounds, self.batch = lane_plan(self.camera.path_count, self.camera.spp, self.lanes, self.batch_paths)
        lanes = self.lanes = len(self.bounds) - 1
        self.side = [torch.cuda.Stream(device=self.film.light.device) for _ in range(lanes - 1)]
        self.lane_paths = [torch.empty(self.batch * abi.GZ_PATH.itemsize, dtype=torch.uint8,
                                       device=self.film.light.device) for _ in range(lanes)]
        self.d_paths = self.lane_paths[0]

    def render(self, stream=None) -> None:
        """Adds one frame (every path sample of the crop) to the film."""
        import torch
        main = stream if stream is not None else torch.cuda.current_stream()
        if self.side:
            fork = torch.cuda.Event()
            fork.record(main)
            for s in self.side:
                s.wait_event(fork)
        for lane in range(self.lanes):
            s = main if lane == 0 else self.side[lane - 1]
            buf = self.lane_paths[lane]
            first, end = self.bounds[lane], self.bounds[lane + 1]
            while first < end:
                count = min(self.batch, end - first)
                self.camera.paths_device(buf, first, count, s)
                self.integrator.trace(buf, count, s)
                self.film.add(buf, count, s)
                first += count
        for s in self.side:
            join = torch.cuda.Event()
            join.record(s)
            main.wait_event(join)

    def check(self, stream=None) -> None:
        """gz_gpu_check_traversal for every lane's stream (each stream has its own flag)."""
        self.integrator.bh.check_traversal(stream)
        for s in self.side:
            self.integrator.bh.check_traversal(s)


class TileRenderer:
    """The frame loop the reference renders with (Core/Tile.swift:19-219, "mode A"): camera paths
    with their filter weight, then up to maxDepth + 1 rounds of closest hit + the wavefront diffuse
    shade (gz_gpu_shade_device: emission ends a path, a diffuse hit bounces it, anything else
    stops it; no next-event estimation, no Russian roulette), each path's estimate into the film.
    The reference draws each path's PCG stream seeds (pcgState, pcgInc) from its xoshiro generator
    before every shade; here they come from a seeded torch generator on the device.

    Everything stays on the device: GzPath records from the camera kernel, PathState_C records
    for the shade, rays gathered for the traversal, Intersection_C records for the shade, and
    the estimates copied back into the GzPath records for the film."""

    def __init__(self, camera: PerspectiveCamera, bh, film: DeviceFilm, max_depth: int, batch_paths: int = 1 << 22,
                 seed: int = 0):
        import torch
        self.camera, self.bh, self.film, self.max_depth = camera, bh, film, int(max_depth)
        spp = camera.spp
        self.batch = max(spp, (min(batch_paths, camera.path_count) // spp) * spp)
        dev = film.light.device
        n = self.batch
        self.d_gz = torch.empty(n * abi.GZ_PATH.itemsize, dtype=torch.uint8, device=dev)
        self.d_ps = torch.zeros(n * abi.PATH_STATE.itemsize, dtype=torch.uint8, device=dev)
        self.d_rays = torch.empty(n * abi.RAY.itemsize, dtype=torch.uint8, device=dev)
        self.d_tmax = torch.full((n,), float("inf"), dtype=torch.float32, device=dev)
        self.d_hits = torch.empty(n * abi.GZ_HIT.itemsize, dtype=torch.uint8, device=dev)
        self.d_isect = torch.empty(n * abi.INTERSECTION.itemsize, dtype=torch.uint8, device=dev)
        self.gen = torch.Generator(device=dev)
        self.gen.manual_seed(seed)

    def render(self, stream=None) -> None:
        """Adds one frame (every path sample of the crop) to the film."""
        import torch
        lib = abi.load()
        st = abi.stream_ptr(stream)
        gz_size, ps_size = abi.GZ_PATH.itemsize, abi.PATH_STATE.itemsize
        off_ray = abi.GZ_PATH.fields["ray"][1]
        off_est = abi.GZ_PATH.fields["estimate"][1]
        p_thr = abi.PATH_STATE.fields["throughput"][1]
        p_est = abi.PATH_STATE.fields["estimate"][1]
        p_pcg = abi.PATH_STATE.fields["pcgState"][1]
        p_act = abi.PATH_STATE.fields["active"][1]
        one = torch.tensor([1.0, 1.0, 1.0], dtype=torch.float32, device=self.d_ps.device).view(torch.uint8)
        n_all = self.camera.path_count
        first = 0
        while first < n_all:
            n = min(self.batch, n_all - first)
            self.camera.paths_device(self.d_gz, first, n, stream)
            gz = self.d_gz[:n * gz_size].view(n, gz_size)
            ps = self.d_ps[:n * ps_size].view(n, ps_size)
            ps.zero_()
            ps[:, 0:24] = gz[:, off_ray:off_ray + 24]
            ps[:, p_thr:p_thr + 12] = one
            ps[:, p_act] = 1
            rays = self.d_rays[:n * 24].view(n, 24)
            for _ in range(self.max_depth + 1):
                rays.copy_(ps[:, 0:24])
                self.bh.closest_device(self.d_rays[:n * 24], self.d_tmax[:n], self.d_hits[:n * 16], stream=stream)
                abi.check(lib.gz_gpu_expand_hits_device(self.bh.handle, abi.ptr(self.d_hits), abi.ptr(self.d_tmax), n,
                                                         abi.ptr(self.d_isect), st), "gz_gpu_expand_hits_device")
                seeds = torch.randint(-(1 << 63), (1 << 63) - 1, (n, 2), dtype=torch.int64,
                                      device=self.d_ps.device, generator=self.gen)
                ps[:, p_pcg:p_pcg + 16] = seeds.view(torch.uint8).view(n, 16)
                abi.check(lib.gz_gpu_shade_device(self.bh.handle, abi.ptr(self.d_ps), n, abi.ptr(self.d_isect), st),
                          "gz_gpu_shade_device")
            gz[:, off_est:off_est + 12] = ps[:, p_est:p_est + 12]
            self.film.add(self.d_gz, n, stream)
            first += n
