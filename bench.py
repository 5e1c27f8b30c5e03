#!/usr/bin/env python3
"""Benchmark: incoherent closest-hit Mrays/s over 1M triangles on one MI355X.

Workload (one "step" = one launch of the closest-hit kernel over the whole batch):
  scene  1,048,576-triangle soup, centroids U[0,1]^3, log-normal edges (median 2e-3, sigma 0.5),
         random orientations (seed 1), one SAH BVH2 built by the host builder;
  rays   16,777,216 incoherent rays: origins U[0,1]^3 (seed 2), directions uniform on the sphere
         (seed 3), tMax U[0.1, 2] (seed 4); inputs resident in HBM before timing starts.
Prints ONE JSON line. ``--gpus N`` with N > 1 is not supported (one GPU per box).

value = rays / the whole call's time (binning the batch by origin cell + the traversal kernel).
roofline.achieved = algorithmic bytes of the traversal kernel / its average duration (HIP events
the library records around it), with the per-ray figure B = 32 (the sorted ray record; 28 for
ray + tMax when not binned) + 16 (hit record) + 64 x interior visits + 40 x triangle tests, the
visit and test counts taken from the counting variant of the same kernel over the same rays
(checked against the CPU oracle on the baseline sample). cpu_baseline = the CPU oracle (a port of
the reference's mojo_cpu_traverse_batch) on the host cores, on a bounded sample of the same rays.

The other metrics of the workload set are reported under "extra", each with its own roofline
and CPU baseline: path samples/s of whole frames on the device for the Cornell box (512^2 x 16
spp, "path_trace") and for the 5M-triangle displaced icosphere (1920x1080 x 64 spp,
"path_trace_full"), 64 scenes of uneven size in one launch sequence ("batch"), and the path-replay
adjoint's time over the forward render ("adjoint").
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "incoherent Mrays/s @1M tris"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


PMC_SUMMARY = os.path.join(ROOT, "profiles", "pmc_k_trace_closest_incoherent_1M.md")


def measured_traffic() -> float | None:
    """Bytes per launch of the traversal kernel that leave L2 for the memory side (the MALL or HBM),
    from the committed rocprofv3 --pmc summary: 128/64/32 B x the sized read requests
    (TCC_EA0_RDREQ_128B/64B/32B) plus the write requests (TCC_EA0_WRREQ, 64B); None if absent."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def cpu_host() -> dict:
    """The host CPU the baseline runs on: logical CPUs of the machine, those this process may use,
    the share per GPU the box grants (OMP_NUM_THREADS, which the box sets to its CPU share), and
    the CPU model from /proc/cpuinfo. The baseline runs at share - 1 threads, the reference's rule
    of one worker less than the cores it sees (TileRenderer.swift:77), and at share threads."""
    nproc = os.cpu_count() or 1
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "") or aff)
    except ValueError:
        share = aff
    share = max(1, min(share, aff))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "affinity": aff, "share": share, "model": model,
            "threads": [max(1, share - 1), share]}


CPU_BUILD = {"flags": None}


def cpu_build() -> str:
    """Builds the oracle with -march=native on this machine once (the CPU baseline's build)."""
    if CPU_BUILD["flags"] is None:
        import oracle
        CPU_BUILD["flags"] = oracle.use_native()
        log(f"cpu baseline build: {CPU_BUILD['flags']}")
    return CPU_BUILD["flags"]


def cpu_baseline_record(unit: str, rates: dict, host: dict, kind: str, sample: str) -> dict:
    """cpu_baseline object: value at share - 1 threads (the reference's worker rule), the rate at
    all share threads beside it, the host, the build flags and the sample."""
    n1, n = host["threads"]
    return {"value": rates[n1], "unit": unit, "cores": n1, "kind": kind, "value_at_share_threads": rates[n],
            "share_threads": n, "host_nproc": host["nproc"], "host_affinity": host["affinity"],
            "cpu_model": host["model"], "build": CPU_BUILD["flags"], "sample": sample}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1 << 20)
    ap.add_argument("--rays", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-path", action="store_true", help="skip the MIS path-tracing leg")
    ap.add_argument("--path-res", type=int, default=512)
    ap.add_argument("--path-spp", type=int, default=16)
    ap.add_argument("--path-steps", type=int, default=5)
    ap.add_argument("--no-full-path", action="store_true", help="skip the 5M-triangle 1080p path-tracing leg")
    ap.add_argument("--full-res", type=lambda s: tuple(int(x) for x in s.split("x")), default=(1920, 1080))
    ap.add_argument("--full-spp", type=int, default=64)
    ap.add_argument("--full-subdiv", type=int, default=9)
    ap.add_argument("--full-frames", type=int, default=1)
    ap.add_argument("--full-batch-paths", type=int, default=1 << 28,
                    help="paths per launch sequence of the frame, over all lanes (default: the whole frame)")
    ap.add_argument("--path-lanes", type=int, default=0,
                    help="FrameRenderer lanes for the Cornell frame (0: automatic, two at 4.2M paths)")
    ap.add_argument("--full-lanes", type=int, default=0,
                    help="FrameRenderer lanes for the frame (0: automatic, two for the 5M-triangle frame)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1)
    ap.add_argument("--batch-round-paths", type=int, default=1 << 27,
                    help="slots per round of the batch, and paths per launch sequence of the one-at-a-time loop")
    ap.add_argument("--batch-solo-paths", type=int, default=None,
                    help="scenes of the batch with more paths render alone (default GZ_BATCH_SOLO_PATHS; 0: none)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the inverse-rendering adjoint leg")
    ap.add_argument("--adj-res", type=int, default=512)
    ap.add_argument("--adj-spp", type=int, default=32)
    ap.add_argument("--adj-target-spp", type=int, default=256)
    ap.add_argument("--adj-steps", type=int, default=3)
    ap.add_argument("--adj-check-seeds", type=int, default=8,
                    help="Z-Sobol seeds of the adjoint's central-difference check (0: skip it)")
    ap.add_argument("--trace-config", type=lambda t: tuple(int(x) for x in t.split(",")), default=None,
                    help="persistent,lds_stack,refill_lanes,grid_scale,chunk,bin_min_rays for the traversal leg "
                         "instead of the autotuner's choice (for profiler runs, where the timings it ranks by "
                         "are distorted)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device copy (reference dispatch for HBM byte counters)")
    ap.add_argument("--only", choices=["path", "full", "batch", "adjoint"], default=None,
                    help="run one of the extra legs alone and print its JSON (for profiling)")
    return ap.parse_args()


def spread_sample(cam, n_paths: int, table, block_pixels: int = 16):
    """CPU-baseline sample: about n_paths path samples in blocks of whole pixels spread evenly
    over the crop (so the sample sees sky, walls and objects in the image's proportions)."""
    import oracle
    from gonzales_amd import abi
    spp = cam.spp
    npix = cam.path_count // spp
    blocks = max(1, min(npix // block_pixels, n_paths // (block_pixels * spp)))
    stride = npix // blocks
    parts = [oracle.camera_paths(cam.c, j * stride * spp, block_pixels * spp, abi.GZ_PATH, table)
             for j in range(blocks)]
    return abi.concat(parts, abi.GZ_PATH), [(j * stride * spp, block_pixels * spp) for j in range(blocks)]


def gpu_paths_of_ranges(cam, integ, ranges, dev):
    """Device records of the path samples in `ranges` ((first, count) of the frame's path
    numbering): camera paths of each range into one buffer, then one trace of them all. The
    records do not depend on the launch that traces them (tests/test_gpu_mis.py), so they equal
    the timed frame's records of the same path samples."""
    import torch

    from gonzales_amd import abi
    k = sum(c for _, c in ranges)
    rec = abi.GZ_PATH.itemsize
    d = torch.empty(k * rec, dtype=torch.uint8, device=dev)
    off = 0
    for first, count in ranges:
        cam.paths_device(d[off * rec:(off + count) * rec], first, count)
        off += count
    integ.trace(d, k)
    torch.cuda.synchronize()
    integ.bh.check_traversal()
    return d.cpu().numpy().view(abi.GZ_PATH)


def paths_agreeing(got, ref) -> float:
    """Fraction of paths with the same bounce count and flags and estimate within 1e-4 relative."""
    e_g, e_r = got["estimate"].astype(np.float64), ref["estimate"].astype(np.float64)
    ok = ((np.abs(e_g - e_r) <= 1e-4 * np.abs(e_r) + 1e-6).all(1) & (got["bounce"] == ref["bounce"])
          & (got["flags"] == ref["flags"]))
    return float(ok.mean())


def path_cpu_leg(args, bh, integ, cam, sky, host: dict, dev, label: str, seconds: float) -> dict:
    """The oracle's whole-frame chain (camera paths -> every MIS bounce, each ray through the
    oracle's traversal -> film sums) on blocks of 16 pixels spread over the frame, timed at
    share - 1 and share threads (about `seconds` in all), and the fraction of those path samples
    whose device records agree with the oracle's."""
    import ctypes

    import oracle
    from gonzales_amd import abi
    cpu_build()
    opt = integ.options
    desc, keep = bh.descriptor()
    table = oracle.sobol_matrices()
    skyv = (0.0, 0.0, 0.0) if sky is None else tuple(sky)
    prm = oracle.MisParams(opt.max_depth, 0, opt.zsobol_seed, opt.log2_spp, opt.base4_digits, opt.rr_seed,
                           (ctypes.c_float * 3)(*skyv), 0 if sky is None else 1)
    n1, nn = host["threads"]
    calib, _ = spread_sample(cam, 1 << 13, table)
    t1 = time.perf_counter()
    oracle.mis_trace(ctypes.addressof(desc), len(bh.nodes), integ.materials, integ.lights, prm, calib, nn, table)
    rate = len(calib) / (time.perf_counter() - t1)
    start, ranges = spread_sample(cam, int(max(1 << 13, rate * seconds / 2)), table)
    c = cam.c
    px = c.resX * c.resY
    rates, sample = {}, None
    for th in sorted({n1, nn}):
        sample = start.copy()
        light, albedo, normal = (np.zeros(3 * px, np.float32) for _ in range(3))
        weight = np.zeros(px, np.float32)
        t1 = time.perf_counter()
        oracle.mis_trace(ctypes.addressof(desc), len(bh.nodes), integ.materials, integ.lights, prm, sample, th,
                         table)
        oracle.film_add(c, sample, cam.film.iso, cam.film.max_component_value, light, weight, albedo, normal)
        rates[th] = round(len(sample) / (time.perf_counter() - t1), 1)
    rates.setdefault(n1, rates[nn])
    rates.setdefault(nn, rates[n1])
    agree = paths_agreeing(gpu_paths_of_ranges(cam, integ, ranges, dev), sample)
    m = len(sample)
    rec = cpu_baseline_record("path samples/s", rates, host, "port",
                              f"{m} of the {cam.path_count} path samples ({len(ranges)} blocks of 16 pixels spread "
                              f"over the {label}): camera + every bounce (the oracle's traversal per ray) + film")
    log(f"{label} cpu oracle {rates[n1]:.0f} / {rates[nn]:.0f} path samples/s on {n1} / {nn} threads "
        f"({m} paths), agreeing {agree:.5f}")
    return {"cpu_baseline": rec, "cpu_sample_paths_agreeing": round(agree, 5)}


PATH_RECORD_BYTES = 224  # GzPath read + write per bounce
RECORD_BYTES = 33  # replay record per path and level: path-ray hit, BSDF-ray hit, shadow visibility
SHADING_FETCH_BYTES = 160  # PrimId 32 + mesh view 40 + 3 vertex indices 24 + 3 points 48 + material 16


def path_leg(args, dev, stream, host: dict, run_cpu: bool) -> dict:
    """Cornell box, 512 x 512 x 16 spp, max depth 5: a whole frame on the device.

    One step = one frame: camera paths for all 4,194,304 path samples (Z-Sobol camera sample,
    Gaussian filter r 1.5 sigma 0.5, pinhole rays) + every MIS bounce of every path (shadow rays
    and MIS BSDF rays included) + the film sums. Algorithmic bytes per path = 112 (record written
    by the camera) + sum over its bounces of (224 path record + 160 shading fetch) + 64 x interior
    visits + 40 x triangle tests of all its rays + 112 (record read by the film), the work
    counts from the counting variant of the same kernel."""
    import ctypes

    import torch

    from gonzales_amd import abi, scenes
    from gonzales_amd.camera import DeviceFilm, cornell_camera
    from gonzales_amd.integrator import MisOptions, VolumePathIntegrator

    res, spp = args.path_res, args.path_spp
    cam = cornell_camera(res=res, spp=spp, seed=0)
    bh, _ = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    opt = MisOptions(max_depth=5, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    integ = VolumePathIntegrator(bh, opt)
    n = cam.path_count
    d_paths = torch.empty(n * abi.GZ_PATH.itemsize, dtype=torch.uint8, device=dev)
    d_stats = torch.zeros(n * 3, dtype=torch.int32, device=dev)
    film = DeviceFilm(cam, dev)
    cam.paths_device(d_paths)
    integ.trace_stats(d_paths, d_stats)
    bh.check_traversal()
    st = d_stats.view(n, 3).to(torch.int64).sum(0).tolist()
    algo = 2 * abi.GZ_PATH.itemsize * n + (PATH_RECORD_BYTES + SHADING_FETCH_BYTES) * st[2] + 64 * st[0] + 40 * st[1]
    from gonzales_amd.autotune import autotune_mis, describe
    cam.paths_device(d_paths)
    best, timings = autotune_mis(integ, d_paths.clone(), n, reps=2)
    log(f"path leg launch {best}")

    from gonzales_amd.render import FrameRenderer
    rnd = FrameRenderer(cam, integ, film, batch_paths=n, lanes=args.path_lanes)
    log(f"path leg lanes {rnd.lanes}, bounds {rnd.bounds}")

    def step():
        rnd.render(stream)

    film.clear()
    for _ in range(2):
        step()
    torch.cuda.synchronize()
    film.clear()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    steps = args.path_steps
    a.record(stream)
    for _ in range(steps):
        step()
    b.record(stream)
    torch.cuda.synchronize()
    rnd.check(stream)
    ms = a.elapsed_time(b) / steps
    img = film.resolve().cpu().numpy().reshape(res, res, 3)
    res_d = {"workload": f"Cornell box (34 triangles, 2 area-light triangles), {res}x{res}, {spp} spp, max depth 5, "
                         "Z-Sobol seed 0, RR after bounce 1, diffuse, Gaussian filter r 1.5 sigma 0.5; "
                         "camera + bounces + film per frame",
             "path_samples_per_s": round(n / (ms * 1e-3), 1), "ms_per_frame": round(ms, 3),
             "bounces_per_path": round(st[2] / n, 3), "interior_per_path": round(st[0] / n, 2),
             "tris_per_path": round(st[1] / n, 2), "bytes_per_path": round(algo / n, 1),
             "roofline": {"bound": "hbm", "achieved": round(algo / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
             "image_mean_rgb": [round(float(x), 5) for x in img.reshape(-1, 3).mean(0)],
             "film_zero_weight_pixels": int(film.bad.item()), "lanes": rnd.lanes,
             "launch": describe(timings)[0], "launch_ms": describe(timings)}
    log(f"path leg: {res_d['path_samples_per_s'] / 1e6:.2f} M path samples/s, {ms:.2f} ms/frame, "
        f"{res_d['bounces_per_path']} bounces/path, {res_d['bytes_per_path']} B/path")
    if run_cpu:
        res_d.update(path_cpu_leg(args, bh, integ, cam, None, host, dev, "Cornell image", args.cpu_seconds / 2))
    bh.free()
    return res_d


def full_path_leg(args, dev, stream, host: dict, run_cpu: bool) -> dict:
    """Displaced icosphere (20 x 4^9 = 5,242,880 triangles, fBm amplitude 0.05) on a ground plane,
    diffuse / dielectric / coated-conductor patches of 64 faces, one quad area light and a
    constant sky; 1920 x 1080, 64 spp, max depth 8, RR after bounce 1. One step = one frame of
    132,710,400 path samples rendered in batches of whole pixels (camera -> bounces -> film).
    Work counts (bytes per path) come from the counting kernel over the first batch."""
    import ctypes

    import torch

    from gonzales_amd import abi, scenes
    from gonzales_amd.camera import DeviceFilm, Film, PerspectiveCamera, look_at
    from gonzales_amd.integrator import MisOptions, VolumePathIntegrator
    from gonzales_amd.render import FrameRenderer

    t0 = time.perf_counter()
    bh, gz = scenes.displaced_sphere_scene(args.full_subdiv)
    t_build = time.perf_counter() - t0
    bh.upload_to_gpu()
    w, h = args.full_res
    cam = PerspectiveCamera(look_at((0.0, 0.45, 3.4), (0.0, -0.15, 0.0), (0, 1, 0)), 40.0, Film(w, h),
                            args.full_spp, seed=0)
    opt = MisOptions(max_depth=8, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    sky = (0.4, 0.45, 0.5)
    integ = VolumePathIntegrator(bh, opt, sky=sky, materials=gz)
    film = DeviceFilm(cam, dev)
    rnd = FrameRenderer(cam, integ, film, batch_paths=args.full_batch_paths, lanes=args.full_lanes)
    n = cam.path_count
    # counting pass over a whole frame (untimed; it also warms the kernels up)
    m = rnd.batch
    d_stats = torch.zeros(m * 3, dtype=torch.int32, device=dev)
    sums = torch.zeros(3, dtype=torch.int64, device=dev)
    first = 0
    while first < n:
        c = min(m, n - first)
        cam.paths_device(rnd.d_paths, first, c)
        d_stats.zero_()
        integ.trace_stats(rnd.d_paths, d_stats, c)
        sums += d_stats[:3 * c].view(c, 3).to(torch.int64).sum(0)
        first += c
    bh.check_traversal()
    st = sums.tolist()
    del d_stats
    per_path = (2 * abi.GZ_PATH.itemsize * n + (PATH_RECORD_BYTES + SHADING_FETCH_BYTES) * st[2] + 64 * st[0]
                + 40 * st[1]) / n
    from gonzales_amd.autotune import autotune_mis, describe
    tune_n = min(rnd.batch, n)  # one whole batch: the tail of each bounce level depends on its size
    cam.paths_device(rnd.d_paths, n // 2 - tune_n // 2, tune_n)  # centre of the frame
    best, timings = autotune_mis(integ, rnd.d_paths[:tune_n * abi.GZ_PATH.itemsize].clone(), tune_n, reps=1)
    log(f"full path leg launch {best}")
    log(f"full path leg lanes {rnd.lanes}, bounds {rnd.bounds}")
    torch.cuda.synchronize()
    film.clear()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(args.full_frames):
        rnd.render(stream)
    b.record(stream)
    torch.cuda.synchronize()
    rnd.check(stream)
    ms = a.elapsed_time(b) / args.full_frames
    img = film.resolve().cpu().numpy().reshape(h, w, 3)
    achieved = per_path * n / (ms * 1e-3) / 1e9
    res = {"workload": f"displaced icosphere {len(bh.prim_ids) - 4} triangles + ground plane, patch materials "
                       f"(diffuse / dielectric / coated conductor, GGX roughness U[0.05,0.6]), quad area light + "
                       f"constant sky, {w}x{h}, {args.full_spp} spp, max depth 8; camera + bounces + film",
           "path_samples_per_s": round(n / (ms * 1e-3), 1), "ms_per_frame": round(ms, 2),
           "host_scene_build_s": round(t_build, 2),
           "bounces_per_path": round(st[2] / n, 3), "interior_per_path": round(st[0] / n, 2),
           "tris_per_path": round(st[1] / n, 2), "bytes_per_path": round(per_path, 1),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4)},
           "image_mean_rgb": [round(float(x), 5) for x in img.reshape(-1, 3).mean(0)],
           "film_zero_weight_pixels": int(film.bad.item()),
           "lanes": rnd.lanes, "lane_bounds": rnd.bounds,
           "launch": describe(timings)[0], "launch_ms": describe(timings)}
    log(f"full path leg: {res['path_samples_per_s'] / 1e6:.2f} M path samples/s, {ms:.1f} ms/frame, "
        f"{res['bounces_per_path']} bounces/path")
    if run_cpu:
        res.update(path_cpu_leg(args, bh, integ, cam, sky, host, dev, "1080p frame", args.cpu_seconds / 2))
    bh.free()
    return res


def device_build_leg(args, host_bh, d_rays, d_tmax, n, stream, host_mrays) -> dict:
    """The same 1M-triangle soup built on the device against the host SAH build, two ways: the
    linear BVH (Morton codes, radix sort, Karras hierarchy, bottom-up boxes, leaves of <= 4 or
    <= 2) and the binned SAH with the host builder's decisions (the host's tree, built on the
    device). Build time (upload of primIds and meshes included), SAH cost of the exported tree in
    the host builder's model, and incoherent closest-hit Mrays/s over the same rays with its own
    autotuned launch configuration; the hit distances of both trees are compared bit for bit."""
    import torch

    from gonzales_amd import scenes
    from gonzales_amd.accelerator import build_bounding_hierarchy_gpu, sah_cost
    from gonzales_amd.autotune import autotune

    scene, prims = scenes.soup_scene(args.tris, 1)
    out = {}
    for max_leaf, sah in ((4, False), (2, False), (4, True)):
        build_bounding_hierarchy_gpu(scene, prims, max_leaf, sah).free()  # warm-up (module load, allocator)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dbh = build_bounding_hierarchy_gpu(scene, prims, max_leaf, sah)
        t_dev = time.perf_counter() - t0
        best, _ = autotune(dbh, d_rays, d_tmax, any_hit=False, sample=n, reps=2)
        d_hits = torch.empty(n * 16, dtype=torch.uint8, device=d_rays.device)
        for _ in range(2):
            dbh.closest_device(d_rays, d_tmax, d_hits)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(5):
            dbh.closest_device(d_rays, d_tmax, d_hits)
        b.record(stream)
        torch.cuda.synchronize()
        dbh.check_traversal()
        ms = a.elapsed_time(b) / 5
        h_hits = torch.empty_like(d_hits)
        host_bh.closest_device(d_rays, d_tmax, h_hits)
        torch.cuda.synchronize()
        t_same = bool(torch.equal(d_hits.view(torch.int32).view(n, 4)[:, 1], h_hits.view(torch.int32).view(n, 4)[:, 1]))
        info = dbh.info()
        key = "sah" if sah else f"max_leaf_{max_leaf}"
        out[key] = {
            "build_s_including_upload": round(t_dev, 3), "sah_cost": round(sah_cost(dbh.nodes), 3),
            "closest_mrays_per_s": round(n / (ms * 1e-3) / 1e6, 2), "depth": info.maxDepth,
            "hit_distances_equal_host_build": t_same}
        what = "binned SAH" if sah else f"linear BVH, leaves <= {max_leaf}"
        log(f"device build ({what}): {t_dev:.3f}s, SAH {out[key]['sah_cost']}, "
            f"{n / (ms * 1e-3) / 1e6:.1f} Mrays/s, distances equal: {t_same}")
        dbh.free()
    out["host_build"] = {"sah_cost": round(sah_cost(host_bh.nodes), 3), "closest_mrays_per_s": round(host_mrays, 2)}
    return out


def batch_leg(args, dev, stream, host: dict, run_cpu: bool) -> dict:
    """64 scenes of uneven size in one launch sequence: triangle counts log-uniform in [1e4, 2e6]
    (displaced spheres and triangle soups), resolutions 256^2 / 512^2 / 1024^2, 4 / 16 / 64 spp,
    patches of diffuse / dielectric / GGX conductor materials, an area light and a sky, max depth 6
    (seed 7). One step = every path sample of every scene: the scenes of at most 4M paths in rounds
    of 128M path slots, each round one replay of a captured HIP graph (camera rays, bounce levels
    over device path queues, film), then each larger scene alone with the single-scene launch
    sequence (one queue storage for all of them).
    The launch shape is autotuned on 128x128 crops of the same scenes. For comparison, the same
    frames rendered one scene at a time (the single-scene loop) are timed too."""
    from concurrent.futures import ThreadPoolExecutor

    import torch

    from gonzales_amd import scenes
    from gonzales_amd.batch import BatchEntry, SceneBatch, autotune_batch, BATCH_CANDIDATES
    from gonzales_amd.autotune import describe
    from gonzales_amd.camera import DeviceFilm, Film, PerspectiveCamera, look_at
    from gonzales_amd.integrator import MisOptions, VolumePathIntegrator
    from gonzales_amd.render import FrameRenderer

    specs = scenes.batch_scene_specs(args.batch_scenes, seed=7)
    t0 = time.perf_counter()
    workers = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8))
    with ThreadPoolExecutor(workers) as ex:
        built = list(ex.map(scenes.batch_scene, sorted(specs, key=lambda s: -s["tris"])))
    built = {sp["index"]: b for sp, b in zip(sorted(specs, key=lambda s: -s["tris"]), built)}
    t_build = time.perf_counter() - t0
    log(f"batch leg: {len(specs)} scenes, {sum(s['tris'] for s in specs)} triangles built in {t_build:.1f}s "
        f"on {workers} threads")
    sky = (0.4, 0.45, 0.5)
    entries, tune_entries, hs = [], [], []
    for sp in specs:
        bh, gz, (eye, at, up, fov) = built[sp["index"]]
        bh.upload_to_gpu()
        hs.append(bh)
        r = sp["res"]
        cam = PerspectiveCamera(look_at(eye, at, up), fov, Film(r, r), sp["spp"], seed=sp["index"])
        opt = MisOptions(max_depth=sp["max_depth"], zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                         base4_digits=cam.c.zsBase4Digits)
        integ = VolumePathIntegrator(bh, opt, sky=sky, materials=gz)
        entries.append(BatchEntry(integ, cam, DeviceFilm(cam, dev)))
        c0 = r // 2 - 64
        tcam = PerspectiveCamera(look_at(eye, at, up), fov, Film(r, r), sp["spp"], seed=sp["index"],
                                 crop=(c0, c0 + 128, c0, c0 + 128))
        tune_entries.append(BatchEntry(integ, tcam, DeviceFilm(tcam, dev)))
    del built
    tune = SceneBatch(tune_entries, round_paths=1 << 23)
    best, timings = autotune_batch(tune, BATCH_CANDIDATES, reps=1)
    tune.close()
    from gonzales_amd import abi
    solo = abi.GZ_BATCH_SOLO_PATHS if args.batch_solo_paths is None else args.batch_solo_paths
    batch = SceneBatch(entries, round_paths=args.batch_round_paths, solo_paths=solo)
    batch.set_launch(best.block, best.grid_scale, best.chunk, best.drain_level, best.queued)
    info = batch.info()
    n = info.totalPaths
    batch.render(True)  # capture + warm-up
    torch.cuda.synchronize()
    batch.clear_films()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(args.batch_steps):
        batch.render(True, stream)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.batch_steps
    batch.check(stream)
    for bh in hs:
        bh.check_traversal()
    # the same frames, one scene at a time through the single-scene loop (queued trace), and
    # (untimed) each scene's work counts for the algorithmic bytes
    seq_t, sums = [], torch.zeros(3, dtype=torch.int64, device=dev)
    d_st = None
    for e in entries:
        rnd = FrameRenderer(e.camera, e.integrator, e.film, batch_paths=args.batch_round_paths)
        rnd.render(stream)  # untimed: allocates the scene's queue storage
        a.record(stream)
        rnd.render(stream)
        b.record(stream)
        torch.cuda.synchronize()
        seq_t.append(a.elapsed_time(b))
        if d_st is None or d_st.numel() < 3 * rnd.batch:
            d_st = torch.empty(3 * rnd.batch, dtype=torch.int32, device=dev)
        first, cn = 0, e.camera.path_count
        while first < cn:
            c = min(rnd.batch, cn - first)
            e.camera.paths_device(rnd.d_paths, first, c, stream)
            d_st[:3 * c].zero_()
            e.integrator.trace_stats(rnd.d_paths, d_st, c, stream)
            sums += d_st[:3 * c].view(c, 3).to(torch.int64).sum(0)
            first += c
        del rnd
        e.integrator.release_queues()  # up to 170 B per path: 64 scenes' worth would not fit
    del d_st
    torch.cuda.synchronize()
    for bh in hs:
        bh.check_traversal()
    seq_ms = float(sum(seq_t))
    st = sums.tolist()
    algo = 2 * abi.GZ_PATH.itemsize * n + (PATH_RECORD_BYTES + SHADING_FETCH_BYTES) * st[2] + 64 * st[0] + 40 * st[1]
    achieved = algo / (ms * 1e-3) / 1e9
    rates = np.array([e.camera.path_count / (t * 1e-3) for e, t in zip(entries, seq_t)])
    # each size class alone in one batch (same launch shape) against the loop over its scenes
    counts = np.array([e.camera.path_count for e in entries])
    by_class = {}
    for name, sel in (("small: <= 1M paths (256^2 x 4-16 spp, 512^2 x 4 spp)", counts <= (1 << 20)),
                      ("medium: 1M-4M paths", (counts > (1 << 20)) & (counts <= (1 << 22))),
                      ("large: > 4M paths", counts > (1 << 22))):
        idx = np.nonzero(sel)[0]
        if len(idx) == 0:
            continue
        sub = SceneBatch([entries[k] for k in idx], round_paths=args.batch_round_paths, solo_paths=solo)
        sub.set_launch(best.block, best.grid_scale, best.chunk, best.drain_level, best.queued)
        sub.render(True)
        torch.cuda.synchronize()
        a.record(stream)
        sub.render(True, stream)
        b.record(stream)
        torch.cuda.synchronize()
        t_b = a.elapsed_time(b)
        sub.close()
        t_l = float(sum(seq_t[k] for k in idx))
        by_class[name] = {"scenes": int(len(idx)), "path_samples": int(counts[idx].sum()),
                          "batch_ms": round(t_b, 2), "one_scene_at_a_time_ms": round(t_l, 2),
                          "batch_speedup": round(t_l / t_b, 3)}
        log(f"batch class {name}: {len(idx)} scenes, batch {t_b:.1f} ms vs loop {t_l:.1f} ms")
    res = {"workload": f"{len(specs)} scenes (seed 7): {sum(s['tris'] for s in specs)} triangles in total, "
                       "log-uniform 1e4..2e6 per scene, displaced spheres and soups, 256^2/512^2/1024^2, "
                       "4/16/64 spp, diffuse/dielectric/GGX conductor patches, area light + sky, max depth 6",
           "path_samples": n, "rounds": info.rounds, "round_paths": info.roundPaths,
           "solo_paths": solo, "solo_scenes": info.soloScenes,
           "path_samples_per_s": round(n / (ms * 1e-3), 1), "ms_per_batch": round(ms, 2),
           "one_scene_at_a_time_ms": round(seq_ms, 2), "batch_speedup": round(seq_ms / ms, 3),
           "by_size_class": by_class,
           "per_scene_path_samples_per_s_alone": {"min": round(float(rates.min()), 1),
                                                   "median": round(float(np.median(rates)), 1),
                                                   "max": round(float(rates.max()), 1)},
           "bounces_per_path": round(st[2] / n, 3), "interior_per_path": round(st[0] / n, 2),
           "tris_per_path": round(st[1] / n, 2), "bytes_per_path": round(algo / n, 1),
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4)},
           "host_scene_build_s": round(t_build, 2), "launch": describe(timings)[0], "launch_ms": describe(timings)}
    log(f"batch leg: {n / (ms * 1e-3) / 1e6:.1f} M path samples/s over {n} samples, {ms:.0f} ms "
        f"(one scene at a time {seq_ms:.0f} ms), {info.rounds} rounds, {algo / n:.0f} B/path")
    batch.close()
    if run_cpu:
        res.update(batch_cpu_leg(args, entries, host, dev))
    for e in entries:
        e.film = None
    for bh in hs:
        bh.free()
    return res


def batch_cpu_leg(args, entries, host: dict, dev) -> dict:
    """The oracle's chain on blocks of 16 pixels spread over every scene of the batch, timed at
    share - 1 and share threads (about --cpu-seconds / 2 in all), and the fraction of those path
    samples whose device records agree with the oracle's."""
    import ctypes

    import oracle
    cpu_build()
    table = oracle.sobol_matrices()
    n1, nn = host["threads"]
    views = []
    for e in entries:
        integ, opt = e.integrator, e.integrator.options
        sky = (0.0, 0.0, 0.0) if integ.sky is None else tuple(float(x) for x in integ.sky)
        prm = oracle.MisParams(opt.max_depth, 0, opt.zsobol_seed, opt.log2_spp, opt.base4_digits, opt.rr_seed,
                               (ctypes.c_float * 3)(*sky), 0 if integ.sky is None else 1)
        desc, keep = integ.bh.descriptor()
        views.append((e, prm, desc, keep))

    def run(samples, th):
        t1 = time.perf_counter()
        for (e, prm, desc, _), smp in zip(views, samples):
            oracle.mis_trace(ctypes.addressof(desc), len(e.integrator.bh.nodes), e.integrator.materials,
                             e.integrator.lights, prm, smp, th, table)
        return time.perf_counter() - t1

    calib = [spread_sample(e.camera, 256, table)[0] for e in entries]
    rate = sum(len(c) for c in calib) / run(calib, nn)
    per_scene = int(max(256, rate * args.cpu_seconds / 4 / len(entries)))
    picks = [spread_sample(e.camera, per_scene, table) for e in entries]
    m = sum(len(p[0]) for p in picks)
    rates, last = {}, None
    for th in sorted({n1, nn}):
        last = [p[0].copy() for p in picks]
        rates[th] = round(m / run(last, th), 1)
    rates.setdefault(n1, rates[nn])
    rates.setdefault(nn, rates[n1])
    agree = [paths_agreeing(gpu_paths_of_ranges(e.camera, e.integrator, p[1], dev), smp)
             for e, p, smp in zip(entries, picks, last)]
    w = np.array([len(s) for s in last], np.float64)
    frac = float((np.array(agree) * w).sum() / w.sum())
    for e in entries:
        e.integrator.release_queues()
    rec = cpu_baseline_record("path samples/s", rates, host, "port",
                              f"{m} path samples of the {len(entries)} scenes (blocks of 16 pixels spread over "
                              "each frame): every MIS bounce, the oracle's traversal per ray")
    log(f"batch cpu oracle {rates[n1]:.0f} / {rates[nn]:.0f} path samples/s on {n1} / {nn} threads ({m} paths), "
        f"agreeing {frac:.5f} (worst scene {min(agree):.5f})")
    return {"cpu_baseline": rec, "cpu_sample_paths_agreeing": round(frac, 5),
            "cpu_sample_paths_agreeing_worst_scene": round(float(min(agree)), 5)}


def adjoint_gradient_check(args, bh, gz, ta, tr, albedo, rough, dev) -> dict:
    """The adjoint's gradient error on the inverse-rendering scene: against central differences
    (h = 1e-3, common random numbers, float64 L2 loss) against a target rendered at the hidden
    parameters, for 8 roughness and 8 reflectance parameters, summed over independent Z-Sobol
    seeds at 128^2 x 1024 spp and max depth 2 (the finite difference crosses visibility edges,
    whose jumps of about dL / 2h make its noise fall as 1 / sqrt(seeds); at max depth 5 roulette
    flips add heavier-tailed jumps of the same kind). Also the adjoint/forward time at that depth on
    the timed frame's size."""
    import torch

    from gonzales_amd.camera import cornell_camera
    from gonzales_amd.diff import DiffRenderer, central_difference_check
    from gonzales_amd.integrator import MisOptions

    n_params = len(ta)

    def make(res, spp, seed, depth):
        cam = cornell_camera(res=res, spp=spp, seed=seed)
        opt = MisOptions(max_depth=depth, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                         base4_digits=cam.c.zsBase4Digits)
        return DiffRenderer(bh, cam, gz, opt, device=dev)

    a_t, r_t = albedo.copy(), rough.copy()
    a_t[3:3 + n_params] = ta
    r_t[3 + n_params:] = tr
    rt = make(128, 256, 99, 2)
    rt.set_parameters(a_t, r_t)
    target = rt.render().double()
    del rt
    rough_ids = [3 + n_params + k for k in range(0, n_params, 4)]
    albedo_ids = [(3 + k, (k // 4) % 3) for k in range(0, n_params, 4)]
    t0 = time.perf_counter()
    adj, fd = central_difference_check(lambda seed: make(128, 1024, seed, 2), albedo, rough, target, rough_ids,
                                       albedo_ids, range(args.adj_check_seeds))
    secs = time.perf_counter() - t0
    a, f = adj.sum(0), fd.sum(0)
    k = len(rough_ids)
    noise = (adj - fd).std(0, ddof=1) * np.sqrt(len(adj)) if len(adj) > 1 else np.zeros_like(a)

    def vec(sl):
        return round(float(np.linalg.norm(a[sl] - f[sl]) / np.linalg.norm(f[sl])), 5)

    rel = np.abs(a - f) / np.abs(f)
    # the timed frame's size at the checked depth
    r2 = make(args.adj_res, args.adj_spp, 0, 2)
    r2.set_parameters(albedo, rough)
    tgt2 = torch.zeros(args.adj_res, args.adj_res, 3, dtype=torch.float64, device=dev)
    fw, bw = [], []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    for i in range(4):
        ev[0].record()
        img = r2.render()
        ev[1].record()
        r2.backward(2 * (img.double() - tgt2))
        ev[2].record()
        torch.cuda.synchronize()
        if i:
            fw.append(ev[0].elapsed_time(ev[1]))
            bw.append(ev[1].elapsed_time(ev[2]))
    del r2
    torch.cuda.empty_cache()
    log(f"adjoint gradient check: roughness vector error {vec(slice(0, k))}, reflectance {vec(slice(k, None))} "
        f"({secs:.1f} s)")
    return {"method": "central differences h=1e-3, common random numbers, float64 L2 loss vs a target rendered at "
                      "the hidden parameters (128^2 x 256 spp), parameters at 0.5; device renderer, whose finite "
                      "differences equal the oracle's (tests/test_gpu_adjoint.py)",
            "sample": f"128^2 x 1024 spp x {len(adj)} Z-Sobol seeds ({128 * 128 * 1024 * len(adj)} paths per "
                      "render), max depth 2",
            "params": f"roughness of materials {rough_ids}, reflectance (material, channel) {albedo_ids}",
            "roughness_vector_error": vec(slice(0, k)), "reflectance_vector_error": vec(slice(k, None)),
            "max_relative_error_per_parameter": round(float(rel.max()), 5),
            "noise_floor_relative": [round(float(x), 5) for x in noise / np.abs(f)],
            "adjoint_over_forward_depth2": round(float(np.median(bw) / np.median(fw)), 3),
            "seconds": round(secs, 1)}


def adjoint_leg(args, dev, stream) -> dict:
    """Inverse rendering: the Cornell box with three displaced occluders (200,320 triangles) whose
    patches carry 32 diffuse (reflectance) and 32 conductor (roughness) parameters; 512 x 512 x 32
    spp, max depth 5. Forward = one frame (camera + bounces + film; the full wavefront for every
    level), keeping the replay record of every level; backward = the per-path adjoints from dLoss/dImage of an L2 loss against a
    256-spp target image + the path-replay adjoint of every path from that record (the rays of the
    levels past the record traced again). The backward that traces every ray again is timed too."""
    import torch

    from gonzales_amd import scenes
    from gonzales_amd.camera import cornell_camera
    from gonzales_amd.diff import DiffRenderer
    from gonzales_amd.integrator import MisOptions

    bh, gz, ta, tr = scenes.inverse_rendering_scene(subdivisions=6, n_params=32)
    bh.upload_to_gpu()

    def make(spp):
        cam = cornell_camera(res=args.adj_res, spp=spp, seed=0)
        opt = MisOptions(max_depth=5, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                         base4_digits=cam.c.zsBase4Digits)
        return DiffRenderer(bh, cam, gz, opt, device=dev)

    target = make(args.adj_target_spp).render().double()
    r = make(args.adj_spp)
    m = len(gz)
    albedo = np.full((m, 3), 0.5, np.float32)
    rough = np.full(m, 0.5, np.float32)
    albedo[:3] = gz["reflectance"][:3]
    r.set_parameters(albedo, rough)
    for _ in range(2):
        img = r.render()
        r.backward(2 * (img.double() - target))
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd, bwd = [], []
    for _ in range(args.adj_steps):
        ev[0].record(stream)
        img = r.render()
        ev[1].record(stream)
        g_img = 2 * (img.double() - target)
        ga, gr = r.backward(g_img)
        ev[2].record(stream)
        torch.cuda.synchronize()
        fwd.append(ev[0].elapsed_time(ev[1]))
        bwd.append(ev[1].elapsed_time(ev[2]))
    loss = float(((img.double() - target) ** 2).sum())
    f, b = float(np.median(fwd)), float(np.median(bwd))
    rec_levels = r.recorded_levels
    # the same backward without the record: every ray traced again
    r.record = False
    img = r.render()
    g_img = 2 * (img.double() - target)
    bwd_traced = []
    for _ in range(args.adj_steps):
        ev[1].record(stream)
        r.backward(g_img)
        ev[2].record(stream)
        torch.cuda.synchronize()
        bwd_traced.append(ev[1].elapsed_time(ev[2]))
    b_traced = float(np.median(bwd_traced))
    r.record = True
    grad_check = adjoint_gradient_check(args, bh, gz, ta, tr, albedo, rough, dev) if args.adj_check_seeds > 0 else None
    # work counts of the frame's paths (untimed): the replay traces the same rays and reads and
    # writes the same records as the forward trace
    from gonzales_amd import abi
    n = r.n
    d_st = torch.zeros(3 * n, dtype=torch.int32, device=dev)
    r.d_paths.copy_(r.d_start)
    r.integrator.trace_stats(r.d_paths, d_st)
    st = d_st.view(n, 3).to(torch.int64).sum(0).tolist()
    # the levels past the record: the same paths advanced by the recorded levels' bounces first
    r.d_paths.copy_(r.d_start)
    for _ in range(rec_levels):
        r.integrator.bounce(r.d_paths)
    d_st.zero_()
    r.integrator.trace_stats(r.d_paths, d_st)
    tail = d_st.view(n, 3).to(torch.int64).sum(0).tolist()
    bh.check_traversal()
    shade_bytes = (PATH_RECORD_BYTES + SHADING_FETCH_BYTES) * st[2]
    trace_bytes = shade_bytes + 64 * st[0] + 40 * st[1]
    # forward: camera records written + trace + film read + the record (33 B per path and recorded
    # level, cleared and written); backward: the start records copied, the per-path adjoint input
    # (radiance + pixel adjoint, 24 B written and read), the replay's shading, the record read for
    # the recorded bounces and the rays of the later levels traced again
    fwd_bytes = 2 * abi.GZ_PATH.itemsize * n + trace_bytes + 2 * RECORD_BYTES * n * rec_levels
    bwd_bytes = (2 * abi.GZ_PATH.itemsize * n + 48 * n + shade_bytes + RECORD_BYTES * (st[2] - tail[2]) +
                 64 * tail[0] + 40 * tail[1])
    bwd_traced_bytes = 2 * abi.GZ_PATH.itemsize * n + 48 * n + trace_bytes
    res = {"workload": f"inverse rendering: Cornell box + 2 displaced icospheres ({len(bh.prim_ids)} triangles), "
                       f"32 diffuse + 32 conductor parameter materials, {args.adj_res}x{args.adj_res} x "
                       f"{args.adj_spp} spp, max depth 5, L2 loss vs a {args.adj_target_spp}-spp target",
           "forward_ms": round(f, 3), "backward_ms": round(b, 3), "adjoint_over_forward": round(b / f, 3),
           "recorded_levels": rec_levels, "record_bytes": RECORD_BYTES * n * rec_levels,
           "backward_traced_ms": round(b_traced, 3), "adjoint_over_forward_traced": round(b_traced / f, 3),
           "backward_traced_bytes_per_path": round(bwd_traced_bytes / n, 1),
           "loss": loss, "grad_albedo_norm": float(ga.norm()), "grad_roughness_norm": float(gr.norm()),
           "bounces_per_path": round(st[2] / n, 3), "forward_bytes_per_path": round(fwd_bytes / n, 1),
           "backward_bytes_per_path": round(bwd_bytes / n, 1),
           "roofline": {"bound": "hbm", "kernel": "backward (path replay)",
                        "achieved": round(bwd_bytes / (b * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(bwd_bytes / (b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "roofline_forward": {"bound": "hbm", "achieved": round(fwd_bytes / (f * 1e-3) / 1e9, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(fwd_bytes / (f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "gradient_check": grad_check}
    log(f"adjoint leg: forward {f:.2f} ms, backward {b:.2f} ms, ratio {b / f:.2f}")
    bh.free()
    return res


def main() -> int:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the traversal shards by rays with a replicated scene, which this build does not launch",
              file=sys.stderr)
        return 2

    import torch

    from gonzales_amd import abi, scenes

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    lib = abi.load()
    if not lib.gz_gpu_available():
        print("no gfx950 device", file=sys.stderr)
        return 2
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        # A plain copy of exactly 1 GiB: reads 1 GiB and writes 1 GiB of HBM, the yardstick that
        # scales FETCH_SIZE / WRITE_SIZE counter readings of the traversal dispatches.
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.only:
        stream = torch.cuda.current_stream()
        host = cpu_host()
        run_cpu = not args.no_cpu
        leg = {"path": lambda: path_leg(args, dev, stream, host, run_cpu),
               "full": lambda: full_path_leg(args, dev, stream, host, run_cpu),
               "batch": lambda: batch_leg(args, dev, stream, host, run_cpu),
               "adjoint": lambda: adjoint_leg(args, dev, stream)}[args.only]()
        print(json.dumps({"leg": args.only, **leg}), flush=True)
        return 0

    t0 = time.perf_counter()
    bh = scenes.soup_hierarchy(args.tris, 1)
    t_build = time.perf_counter() - t0
    log(f"scene {args.tris} tris, {len(bh.nodes)} nodes, host build {t_build:.2f}s")
    t0 = time.perf_counter()
    bh.upload_to_gpu()
    t_upload = time.perf_counter() - t0
    info = bh.info()
    log(f"upload {t_upload:.2f}s, device bytes {info.deviceBytes / 2**20:.1f} MiB, depth {info.maxDepth}")

    rays, tmax = scenes.incoherent_rays(args.rays)
    d_rays = torch.from_numpy(rays.view(np.uint8)).to(dev)
    d_tmax = torch.from_numpy(tmax).to(dev)
    n = args.rays
    d_hits = torch.empty(n * 16, dtype=torch.uint8, device=dev)
    d_stats = torch.empty(n * 8, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    # Launch configuration picked by the autotuner on this batch (untimed).
    from gonzales_amd.autotune import autotune, describe

    if args.trace_config is not None:
        from gonzales_amd.autotune import TraceConfig
        best, timings = TraceConfig(*args.trace_config), {}
        log(f"launch config given: {best}")
    else:
        best, timings = autotune(bh, d_rays, d_tmax, any_hit=False, sample=n, reps=2)
        log(f"autotuned launch config: {best}")
    binned = best.bin_min_rays > 0 and n >= best.bin_min_rays
    bh.set_trace_config(best.persistent, best.lds_stack, best.refill_lanes, best.grid_scale, best.chunk,
                        best.bin_min_rays, record_times=1)

    # Work counts for the algorithmic bytes (untimed counting pass).
    bh.closest_device(d_rays, d_tmax, d_hits, d_stats)
    bh.check_traversal()
    st = d_stats.view(torch.int32).view(n, 2).to(torch.int64)
    sum_interior = int(st[:, 0].sum())
    sum_tris = int(st[:, 1].sum())
    # per ray: its 24-byte Ray_C and 4-byte tMax (as a 32-byte sorted record when binned) read, the
    # 16-byte hit written, 64 B per interior visit, 40 B per triangle test
    ray_bytes = (32 if binned else 28) + 16
    algo_bytes = ray_bytes * n + 64 * sum_interior + 40 * sum_tris
    hit_rate = float((d_hits.view(torch.int32).view(n, 4)[:, 0] >= 0).float().mean())
    log(f"interior/ray {sum_interior / n:.1f}, tris/ray {sum_tris / n:.2f}, bytes/ray {algo_bytes / n:.0f}, "
        f"hit rate {hit_rate:.3f}")

    for _ in range(args.warmup):
        bh.closest_device(d_rays, d_tmax, d_hits)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        bh.closest_device(d_rays, d_tmax, d_hits)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    bh.check_traversal()
    ms = ev0.elapsed_time(ev1) / args.steps
    mrays = n / (ms * 1e-3) / 1e6
    # the traversal kernel alone (the roofline's kernel): HIP events the library records around the
    # binning and the traversal launch on this stream, read after each of args.steps more launches
    kbin, ktrace = [], []
    for _ in range(args.steps):
        bh.closest_device(d_rays, d_tmax, d_hits)
        b_ms, t_ms = bh.trace_times()
        kbin.append(b_ms)
        ktrace.append(t_ms)
    bh.check_traversal()
    trace_ms = float(np.mean(ktrace))
    bin_ms = float(np.mean(kbin))
    achieved = algo_bytes / (trace_ms * 1e-3) / 1e9
    log(f"closest: {ms:.3f} ms/launch, {mrays:.1f} Mrays/s (wall {n * args.steps / wall / 1e6:.1f}); "
        f"binning {bin_ms:.3f} ms + traversal kernel {trace_ms:.3f} ms")

    extra = {"host_build_s": round(t_build, 3), "upload_s": round(t_upload, 3),
             "bytes_per_ray": round(algo_bytes / n, 1), "interior_per_ray": round(sum_interior / n, 2),
             "tris_per_ray": round(sum_tris / n, 3), "hit_rate": round(hit_rate, 4),
             "wall_mrays_per_s": round(n * args.steps / wall / 1e6, 2),
             "binned": binned, "binning_ms": round(bin_ms, 4), "traversal_kernel_ms": round(trace_ms, 4),
             # binning traffic per ray: the count kernel reads the 12-byte origin and writes a 2-byte key, the
             # place kernel reads key, ray and tMax and writes the 32-byte sorted record
             "binning_bytes_per_ray": (12 + 2 + 2 + 28 + 32) if binned else 0,
             "launch_config": describe({best: timings.get(best, 0.0)})[0],
             "autotune_ms": describe(timings)}

    def timed(fn, reps=5):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    if not args.no_extra:
        d_occ = torch.empty(n, dtype=torch.uint8, device=dev)
        ms_any = timed(lambda: bh.any_device(d_rays, d_tmax, d_occ))
        extra["incoherent_anyhit_mrays_per_s"] = round(n / (ms_any * 1e-3) / 1e6, 2)
        crays, ctmax = scenes.coherent_rays(4096)
        d_crays = torch.from_numpy(crays.view(np.uint8)).to(dev)
        d_ctmax = torch.from_numpy(ctmax).to(dev)
        nc = len(crays)
        d_chits = torch.empty(nc * 16, dtype=torch.uint8, device=dev)
        cbest, _ = autotune(bh, d_crays, d_ctmax, any_hit=False, sample=nc, reps=2)
        ms_c = timed(lambda: bh.closest_device(d_crays, d_ctmax, d_chits))
        ms_ca = timed(lambda: bh.any_device(d_crays, d_ctmax, d_occ))
        extra["coherent_closest_mrays_per_s"] = round(nc / (ms_c * 1e-3) / 1e6, 2)
        extra["coherent_anyhit_mrays_per_s"] = round(nc / (ms_ca * 1e-3) / 1e6, 2)
        extra["coherent_launch_config"] = describe({cbest: 0.0})[0]
        bh.set_trace_config(best.persistent, best.lds_stack, best.refill_lanes, best.grid_scale, best.chunk,
                            best.bin_min_rays, record_times=1)
        extra["device_build"] = device_build_leg(args, bh, d_rays, d_tmax, n, stream, mrays)
        # End to end through host buffers (H2D + trace + Intersection_C expansion + D2H), 1M rays.
        # (steady state: the first call grows the staging buffers and is not timed)
        m = 1 << 20
        bh.intersect_batch_gpu(rays[:m], tmax[:m])
        t1 = time.perf_counter()
        for _ in range(3):
            bh.intersect_batch_gpu(rays[:m], tmax[:m])
        t2 = time.perf_counter()
        extra["host_buffer_path_mrays_per_s"] = round(3 * m / (t2 - t1) / 1e6, 2)
        bh.check_traversal()
        log(f"any-hit {extra['incoherent_anyhit_mrays_per_s']} Mrays/s, coherent "
            f"{extra['coherent_closest_mrays_per_s']} / {extra['coherent_anyhit_mrays_per_s']}")

    cpu = None
    host = cpu_host()
    if not args.no_cpu:
        import ctypes

        import oracle

        cpu_build()
        desc, meshes_c = bh.descriptor()
        n1, nn = host["threads"]
        # calibrate at all share threads, then size the sample so that both timed runs together
        # take about --cpu-seconds
        calib = 1 << 15
        t1 = time.perf_counter()
        oracle.traverse(ctypes.addressof(desc), len(bh.nodes), rays[:calib], tmax[:calib], abi.INTERSECTION, nn)
        rate = calib / (time.perf_counter() - t1)
        m = int(min(n, max(calib, rate * args.cpu_seconds / 2)))
        m -= m % 4096 if m > 4096 else 0
        rates, outs = {}, {}
        for th in sorted({n1, nn}):
            t1 = time.perf_counter()
            outs[th] = oracle.traverse(ctypes.addressof(desc), len(bh.nodes), rays[:m], tmax[:m], abi.INTERSECTION,
                                       th, with_stats=True)
            rates[th] = round(m / (time.perf_counter() - t1) / 1e6, 4)
        rates.setdefault(n1, rates[nn])
        rates.setdefault(nn, rates[n1])
        ref, ref_st = outs[n1]
        # parity of the timed GPU kernel on the same sample: bit-exact t, u, v, the whole PrimId
        # of the hit primitive, the hit flag and the per-ray work counts
        hits = d_hits.view(torch.int32).view(n, 4)[:m].cpu().numpy().view(abi.GZ_HIT).reshape(-1)
        gst = d_stats.view(torch.int32).view(n, 2)[:m].cpu().numpy().view(abi.TRACE_STATS).reshape(-1)
        hit = hits["prim"] >= 0
        pid = bh.prim_ids[hits["prim"][hit]]
        ok = bool(np.array_equal(hits["t"].view(np.uint32), ref["tHit"].view(np.uint32))
                  and np.array_equal(hits["u"][hit].view(np.uint32), ref["u"][hit].view(np.uint32))
                  and np.array_equal(hits["v"][hit].view(np.uint32), ref["v"][hit].view(np.uint32))
                  and np.array_equal(hit, ref["hit"].astype(bool))
                  and all(np.array_equal(pid[f], ref["primId"][f][hit]) for f in ("id1", "id2", "materialIndex", "type"))
                  and gst.tobytes() == ref_st.tobytes()
                  and all(o[0].tobytes() == ref.tobytes() for o in outs.values()))
        extra["cpu_sample_parity_bit_exact"] = ok
        cpu = cpu_baseline_record("Mrays/s", rates, host, "port",
                                  f"first {m} of the {n} incoherent rays, same scene and BVH2, the oracle's "
                                  "mojo_cpu_traverse_batch loop over 4096-ray chunks")
        per_thread = rates[n1] / n1
        extra["gpu_over_cpu_baseline"] = round(mrays / rates[n1], 1)
        extra["cpu_extrapolated_all_host_threads"] = {
            "value": round(per_thread * host["nproc"], 3), "unit": "Mrays/s",
            "note": f"not measured: the per-thread rate at {n1} threads times the host's {host['nproc']} logical "
                    "CPUs (perfect scaling assumed); the box grants this process a share of "
                    f"{host['share']} threads", "gpu_over_it": round(mrays / (per_thread * host["nproc"]), 1)}
        log(f"cpu oracle {rates[n1]:.3f} / {rates[nn]:.3f} Mrays/s on {n1} / {nn} threads ({m} rays, "
            f"{host['model']}, nproc {host['nproc']}), parity {ok}")

    if not args.no_path:
        extra["path_trace"] = path_leg(args, dev, stream, host, not args.no_cpu)
    if not args.no_full_path:
        extra["path_trace_full"] = full_path_leg(args, dev, stream, host, not args.no_cpu)
    if not args.no_batch:
        extra["batch"] = batch_leg(args, dev, stream, host, not args.no_cpu)
    if not args.no_adjoint:
        extra["adjoint"] = adjoint_leg(args, dev, stream)

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"closest-hit traversal, {args.tris} triangle soup, one SAH BVH2, "
                               f"{n} incoherent rays (origins in the unit cube, directions uniform on the sphere, "
                               "tMax U[0.1,2])",
                   "tris": args.tris, "rays": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "algorithmic_bytes_per_launch": algo_bytes},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    bh.free()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
