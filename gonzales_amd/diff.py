"""Differentiable rendering: a torch.autograd.Function over the GPU MIS integrator whose backward
pass is the path-replay adjoint kernel (gz_gpu_mis_adjoint_device).

By default the forward pass also keeps a replay record in HBM (gz_gpu_mis_trace_record_device: each
path's closest hit, BSDF-ray hit and shadow visibility for the wavefront levels, 33 B per path and
level), and the backward pass replays the paths from it instead of tracing their rays again
(gz_gpu_mis_adjoint_record_device); the gradients are the same.

Parameters are the reflectance of diffuse materials and the (isotropic) roughness of conductor
materials of the scene's GzMaterial table. forward renders the image (camera -> bounces -> film)
and keeps the paths' starting records and radiance; backward turns dLoss/dImage into each path
sample's radiance adjoint (the film divides the sum of a pixel's samples, scaled by iso/100, by
the sum of their filter weights) and replays every path to accumulate dLoss/dtheta in float64.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .camera import DeviceFilm, PerspectiveCamera
from .integrator import MisOptions, VolumePathIntegrator


class DiffRenderer:
    """Renders a full-frame camera (no crop) with the parameters written into the material table.
    record (default): keep the forward pass's replay record and replay from it; this sets the scene
    handle's launch to the full wavefront without a drain level (every level recorded)."""

    def __init__(self, bh, camera: PerspectiveCamera, materials: np.ndarray, options: MisOptions, sky=None,
                 device="cuda", record: bool = True):
        import torch
        c = camera.c
        if (c.x0, c.y0, c.x1, c.y1) != (0, 0, c.resX, c.resY):
            raise ValueError("DiffRenderer renders whole frames (no crop)")
        self.bh, self.camera, self.options, self.sky = bh, camera, options, sky
        self.materials = np.ascontiguousarray(materials, abi.GZ_MATERIAL).copy()
        self.integrator = VolumePathIntegrator(bh, options, sky=sky, materials=self.materials)
        self.film = DeviceFilm(camera, device)
        n = camera.path_count
        self.n = n
        self.d_start = torch.empty(n * abi.GZ_PATH.itemsize, dtype=torch.uint8, device=device)
        self.d_paths = torch.empty_like(self.d_start)
        self.device = device
        self.record = record
        self.d_record = None
        self.recorded_levels = 0  # levels the last render recorded (0: none; backward traces every ray)
        if record:
            # the full wavefront for every level (no drain), so that the record holds every level
            # and the backward traces nothing: on the 512^2 x 32 spp inverse-rendering frame the
            # forward takes 18.1 ms against 17.8 with the default drain from level 3, the backward
            # 10.6 against 12.6 ms
            self.integrator.set_launch(3, 256, 1, 64, 0)

    def set_parameters(self, albedo=None, roughness=None) -> None:
        """albedo: (M, 3) reflectances (used where the material is diffuse); roughness: (M,)
        (used where the material is a conductor)."""
        m = self.materials
        if albedo is not None:
            a = np.asarray(albedo, np.float32).reshape(len(m), 3)
            sel = m["type"] == abi.GZ_MAT_DIFFUSE
            m["reflectance"][sel] = a[sel]
        if roughness is not None:
            r = np.asarray(roughness, np.float32).reshape(len(m))
            sel = m["type"] == abi.GZ_MAT_CONDUCTOR
            m["roughness"][sel] = np.stack([r[sel], r[sel]], 1)
        self.integrator = VolumePathIntegrator(self.bh, self.options, sky=self.sky, materials=m)
        self.recorded_levels = 0  # a record of the old parameters would not match the replay

    def render(self):
        """Forward pass: returns the (H, W, 3) image (a torch tensor on the device)."""
        import torch
        self.camera.paths_device(self.d_start)
        self.d_paths.copy_(self.d_start)
        levels = self._record_levels() if self.record else 0
        if levels > 0:
            lib = abi.load()
            need = int(lib.gz_mis_record_bytes(self.n, levels))
            if self.d_record is None or self.d_record.numel() < need:
                self.d_record = None
                self.d_record = torch.empty(need, dtype=torch.uint8, device=self.device)
            abi.check(lib.gz_gpu_mis_trace_record_device(self.bh.handle, abi.ptr(self.d_paths), self.n,
                                                         ctypes.byref(self.options.params()), abi.ptr(self.d_record),
                                                         levels, abi.stream_ptr(None)),
                      "gz_gpu_mis_trace_record_device")
        else:
            self.integrator.trace(self.d_paths)
        self.recorded_levels = levels
        self.film.clear()
        self.film.add(self.d_paths, self.n)
        img = self.film.resolve()
        c = self.camera.c
        return img.view(c.resY, c.resX, 3).clone()

    def _record_levels(self) -> int:
        """Bounce levels the forward pass can record: those the full wavefront launch runs level by
        level (below its drain level); none for the other launch shapes."""
        queued, _, _, _, drain = self.integrator.launch()
        if queued != 3:
            return 0
        levels = self.options.max_depth + 1
        return min(levels, drain) if drain > 0 else levels

    def path_radiance(self):
        """(n, 3) radiance of each path sample of the last render (a strided view of the records)."""
        import torch
        rec = self.d_paths.view(self.n, abi.GZ_PATH.itemsize)
        off = abi.GZ_PATH.fields["estimate"][1]
        return rec[:, off:off + 12].contiguous().view(torch.float32).view(self.n, 3)

    def backward(self, grad_image):
        """dLoss/dtheta for dLoss/dImage (H, W, 3) of the last render: (grad_albedo (M, 3),
        grad_roughness (M,)) as float64 torch tensors."""
        import torch
        spp = self.camera.spp
        k = self.camera.film.iso / 100.0
        g = grad_image.reshape(-1, 3).to(torch.float32)
        w = self.film.weight
        per_pixel = g * (k / w)[:, None]
        adj = torch.empty(self.n, 6, dtype=torch.float32, device=self.device)
        adj[:, :3] = self.path_radiance()
        adj[:, 3:] = per_pixel.repeat_interleave(spp, dim=0)
        grad = torch.zeros(len(self.materials) * 4, dtype=torch.float64, device=self.device)
        self.d_paths.copy_(self.d_start)
        prm = self.options.params()
        if self.recorded_levels > 0:
            abi.check(abi.load().gz_gpu_mis_adjoint_record_device(
                self.bh.handle, abi.ptr(self.d_paths), abi.ptr(adj), self.n, ctypes.byref(prm), abi.ptr(grad),
                abi.ptr(self.d_record), self.recorded_levels, None), "gz_gpu_mis_adjoint_record_device")
        else:
            abi.check(abi.load().gz_gpu_mis_adjoint_device(self.bh.handle, abi.ptr(self.d_paths), abi.ptr(adj),
                                                            self.n, ctypes.byref(prm), abi.ptr(grad), None),
                      "gz_gpu_mis_adjoint_device")
        grad = grad.view(-1, 4)
        return grad[:, :3].clone(), grad[:, 3].clone()


def central_difference_check(make_renderer, albedo, roughness, target, rough_ids, albedo_ids, seeds, h=1e-3):
    """Adjoint gradients against central differences (step h, common random numbers: the +h and -h
    renders reuse every Z-Sobol index and roulette and BSDF stream) of the float64 L2 loss against
    `target`, for renderers make_renderer(seed) over `seeds`. Parameters: roughness of the materials
    rough_ids, reflectance channels albedo_ids [(material, channel)]. Returns per-seed arrays
    (adjoint, fd), shape (len(seeds), len(rough_ids) + len(albedo_ids)), roughness first."""
    import torch

    def loss(img):
        d = img.double() - target
        return float((d * d).sum())

    adj, fd = [], []
    for seed in seeds:
        r = make_renderer(seed)
        r.set_parameters(albedo, roughness)
        img = r.render()
        ga, gr = r.backward(2 * (img.double() - target))
        ga, gr = ga.cpu().numpy(), gr.cpu().numpy()
        adj.append(np.concatenate([gr[list(rough_ids)], [ga[k, c] for k, c in albedo_ids]]))
        row = []
        for k in rough_ids:
            v = []
            for s in (1, -1):
                rr = np.array(roughness, np.float32, copy=True)
                rr[k] += s * h
                r.set_parameters(albedo, rr)
                v.append(loss(r.render()))
            row.append((v[0] - v[1]) / (2 * h))
        for k, c in albedo_ids:
            v = []
            for s in (1, -1):
                aa = np.array(albedo, np.float32, copy=True)
                aa[k, c] += s * h
                r.set_parameters(aa, roughness)
                v.append(loss(r.render()))
            row.append((v[0] - v[1]) / (2 * h))
        fd.append(np.array(row))
        del r
        torch.cuda.empty_cache()
    return np.array(adj), np.array(fd)


def render_function():
    """torch.autograd.Function: image = f(albedo (M, 3), roughness (M,)) for a DiffRenderer."""
    import torch

    class RenderFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, albedo, roughness, renderer: DiffRenderer):
            renderer.set_parameters(albedo.detach().cpu().numpy(), roughness.detach().cpu().numpy())
            ctx.renderer = renderer
            return renderer.render()

        @staticmethod
        def backward(ctx, grad_image):
            ga, gr = ctx.renderer.backward(grad_image)
            return ga.to(torch.float32), gr.to(torch.float32), None

    return RenderFunction
