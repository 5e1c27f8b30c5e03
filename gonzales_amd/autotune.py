"""Launch-configuration autotuners for the traversal kernels and the MIS path trace.

The results of a traversal never depend on its launch configuration (every configuration runs
the same per-ray algorithm); only the speed does, and the best choice depends on the scene and on
how coherent the rays are (refilling idle lanes often pays off for coherent batches, rarely for
incoherent ones). ``autotune`` times a small set of candidate configurations on a sample of the
caller's batch, keeps the fastest, sets it on the scene handle and can persist the choice as JSON
keyed by (scene signature, query kind, sample size).
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass
from pathlib import Path



@dataclass(frozen=True)
class TraceConfig:
    persistent: int = 3
    lds_stack: int = 16
    refill_lanes: int = 32
    grid_scale: int = 1
    chunk: int = 128
    bin_min_rays: int = 1 << 18


# binned by origin cell (the default bin_min_rays) and in the given order (bin_min_rays 0)
CANDIDATES = (
    TraceConfig(3, 16, 32, 1, 128),
    TraceConfig(3, 8, 32, 1, 128),
    TraceConfig(3, 16, 32, 1, 64),
    TraceConfig(3, 16, 16, 1, 64),
    TraceConfig(3, 16, 48, 1, 64),
    TraceConfig(3, 8, 32, 1, 64),
    TraceConfig(1, 16, 32, 1, 64),
    TraceConfig(0, 16, 16, 1, 64),
    TraceConfig(2, 16, 16, 2, 64),
    TraceConfig(1, 16, 56, 1, 64, 0),
    TraceConfig(3, 16, 56, 1, 64, 0),
    TraceConfig(1, 16, 16, 1, 64, 0),
)


def _apply(bh, cfg: TraceConfig) -> None:
    bh.set_trace_config(cfg.persistent, cfg.lds_stack, cfg.refill_lanes, cfg.grid_scale, cfg.chunk, cfg.bin_min_rays)


def scene_signature(bh) -> str:
    info = bh.info()
    return f"nodes{info.nodeCount}-prims{info.primCount}-depth{info.maxDepth}"


def _time(fn, reps: int) -> float:
    import torch

    fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune(bh, d_rays, d_tmax, any_hit: bool = False, sample: int = 1 << 21, reps: int = 3,
             candidates=CANDIDATES, cache: str | os.PathLike | None = None) -> tuple[TraceConfig, dict]:
    """Times every candidate on the first `sample` rays (device tensors) and keeps the fastest.

    Returns (best config, {config: ms}). With `cache`, a previous choice for the same scene, query
    and sample size is reused without timing.
    """
    import torch

    key = f"{scene_signature(bh)}|{'any' if any_hit else 'closest'}|{sample}"
    if cache is not None and Path(cache).exists():
        saved = json.loads(Path(cache).read_text())
        if key in saved:
            cfg = TraceConfig(**saved[key])
            _apply(bh, cfg)
            return cfg, {}
    n = int(d_tmax.numel())
    m = min(n, sample)
    rays = d_rays[: m * 24]
    tmax = d_tmax[:m]
    out = torch.empty(m * (1 if any_hit else 16), dtype=torch.uint8, device=d_tmax.device)
    timings = {}
    for cfg in candidates:
        _apply(bh, cfg)
        if any_hit:
            ms = _time(lambda: bh.any_device(rays, tmax, out), reps)
        else:
            ms = _time(lambda: bh.closest_device(rays, tmax, out), reps)
        timings[cfg] = ms
    bh.check_traversal()
    best = min(timings, key=timings.get)
    _apply(bh, best)
    if cache is not None:
        p = Path(cache)
        saved = json.loads(p.read_text()) if p.exists() else {}
        saved[key] = asdict(best)
        p.write_text(json.dumps(saved, indent=1, sort_keys=True))
    return best, timings


@dataclass(frozen=True)
class MisLaunch:
    queued: int = 1
    block: int = 256
    grid_scale: int = 1
    chunk: int = 256
    drain_level: int = 0


MIS_CANDIDATES = (
    MisLaunch(3, 256, 1, 64),
    MisLaunch(3, 256, 1, 64, 2),
    MisLaunch(3, 256, 1, 64, 3),
    MisLaunch(3, 256, 1, 64, 4),
    MisLaunch(3, 256, 1, 256, 3),
    MisLaunch(1, 256, 1, 64, 3),
    MisLaunch(1, 256, 1, 256, 1),
    MisLaunch(1, 256, 1, 256, 2),
    MisLaunch(1, 256, 1, 256, 3),
    MisLaunch(2, 256, 1, 64, 3),
    MisLaunch(0, 256, 1, 256),
    MisLaunch(1, 256, 1, 64),
    MisLaunch(1, 256, 1, 256),
    MisLaunch(1, 128, 1, 256, 3),
)


def autotune_mis(integrator, d_start, count: int, reps: int = 2, candidates=MIS_CANDIDATES) -> tuple[MisLaunch, dict]:
    """Times a whole MIS trace of the `count` path records in d_start (a device tensor, left
    unchanged: every run starts from a copy) under each candidate launch configuration, keeps the
    fastest on the integrator's scene handle and returns (best, {config: ms})."""
    import torch

    work = torch.empty_like(d_start)
    s = torch.cuda.current_stream()
    timings = {}
    for cfg in candidates:
        integrator.set_launch(cfg.queued, cfg.block, cfg.grid_scale, cfg.chunk, cfg.drain_level)
        total = 0.0
        for rep in range(reps + 1):
            work.copy_(d_start)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(s)
            integrator.trace(work, count)
            b.record(s)
            torch.cuda.synchronize()
            if rep > 0:  # the first run also grows the queue storage
                total += a.elapsed_time(b)
        timings[cfg] = total / reps
    best = min(timings, key=timings.get)
    integrator.set_launch(best.queued, best.block, best.grid_scale, best.chunk, best.drain_level)
    return best, timings


def describe(timings: dict) -> list[dict]:
    return [dict(asdict(c), ms=round(float(t), 3)) for c, t in sorted(timings.items(), key=lambda kv: kv[1])]


__all__ = ["TraceConfig", "CANDIDATES", "autotune", "describe", "scene_signature", "MisLaunch", "MIS_CANDIDATES",
           "autotune_mis"]
