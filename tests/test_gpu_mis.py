"""GPU parity of the wavefront MIS bounce (VolumePathIntegrator.swift:35-463) with the oracle.

Bar: the path records agree field by field. Paths whose sampled directions went through cos/sin
(ocml on the device, libm in the oracle) can differ in the last bits; a difference in a late
bit can, rarely, send a grazing ray to another triangle. So: at least 99% of paths must agree
exactly in bounce count, dimensions and flags and within 1e-4 relative in estimate and
throughput, and the mean image within 2e-3 relative.
"""
import ctypes

import numpy as np
import pytest

from gonzales_amd import abi, scenes
from gonzales_amd.integrator import MisOptions, VolumePathIntegrator, new_paths, pinhole_rays

from mis_helpers import (OracleMis, agree_mask, box_on_plane_scene, check_paths, light_over_plane, point_paths,
                         run_gpu_and_oracle, to_device, to_host)

pytestmark = pytest.mark.gpu


_to_device, _to_host, _agree, _check, _run_both = to_device, to_host, agree_mask, check_paths, run_gpu_and_oracle


def test_cornell_box_matches_oracle(gpu):
    bh, cam = scenes.cornell_hierarchy()
    w = h = 64
    spp = 4
    rays, px, py = pinhole_rays(cam, w, h)
    paths = abi.concat([new_paths(rays, px, py, s, start_dim=5) for s in range(spp)], abi.GZ_PATH)
    opt = MisOptions.for_image(w, h, spp, max_depth=5)
    got, ref = _run_both(bh, paths, opt)
    assert ref["bounce"].max() >= 4
    _check(got, ref)


@pytest.mark.parametrize("fix_sky", [False, True])
def test_sky_box_matches_oracle(gpu, fix_sky):
    bh = box_on_plane_scene()
    cam = {"eye": (0.0, 1.6, 4.0), "look": (0.0, 0.4, 0.0), "up": (0.0, 1.0, 0.0), "fov": 30.0}
    w = h = 32
    spp = 8
    rays, px, py = pinhole_rays(cam, w, h)
    paths = abi.concat([new_paths(rays, px, py, s) for s in range(spp)], abi.GZ_PATH)
    opt = MisOptions.for_image(w, h, spp, max_depth=4, fix_sky_direction=fix_sky)
    got, ref = _run_both(bh, paths, opt, sky=(1.0, 1.5, 2.0))
    _check(got, ref)


@pytest.mark.parametrize("facing_down", [True, False])
@pytest.mark.parametrize("fix", [False, True])
def test_triangle_light_matches_oracle(gpu, facing_down, fix):
    bh, _ = light_over_plane(facing_down)
    paths = point_paths((0.05, 2.0, -0.9), (0.0, -2.0, 1.0), 4096, log2spp=12)
    opt = MisOptions(max_depth=3, fix_light_pdf=fix, log2_spp=12, base4_digits=6)
    got, ref = _run_both(bh, paths, opt)
    _check(got, ref)


def test_bounce_steps_equal_trace(gpu):
    bh, cam = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    rays, px, py = pinhole_rays(cam, 32, 32)
    paths = new_paths(rays, px, py, 1, start_dim=5)
    opt = MisOptions.for_image(32, 32, 1, max_depth=4)
    integ = VolumePathIntegrator(bh, opt)
    a, b = _to_device(paths), _to_device(paths)
    integ.trace(a)
    for _ in range(opt.max_depth + 1):
        integ.bounce(b)
    ga, gb = _to_host(a), _to_host(b)
    assert ga.tobytes() == gb.tobytes()


def test_trace_stats_counts_work(gpu):
    import torch
    bh, cam = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    rays, px, py = pinhole_rays(cam, 32, 32)
    paths = new_paths(rays, px, py, 2, start_dim=4)
    opt = MisOptions.for_image(32, 32, 4, max_depth=5)
    integ = VolumePathIntegrator(bh, opt)
    a, b = _to_device(paths), _to_device(paths)
    stats = torch.zeros(len(paths) * 3, dtype=torch.int32, device="cuda")
    integ.trace(a)
    integ.trace_stats(b, stats)
    ga, gb = _to_host(a), _to_host(b)
    assert ga.tobytes() == gb.tobytes()
    st = stats.cpu().numpy().view(abi.PATH_STATS)
    assert np.array_equal(st["bounces"], gb["bounce"] + 1)
    # every bounce traces at least its closest-hit ray from the root of a 34-triangle BVH
    assert np.all(st["interior"] >= st["bounces"])
    assert st["tris"].sum() > 0


def test_inactive_paths_are_untouched(gpu):
    bh, cam = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    rays, px, py = pinhole_rays(cam, 16, 16)
    paths = new_paths(rays, px, py, 0, start_dim=5)
    paths["flags"][::3] = 0
    paths["estimate"][::3] = 7.0
    integ = VolumePathIntegrator(bh, MisOptions.for_image(16, 16, 1, max_depth=3))
    d = _to_device(paths)
    integ.trace(d)
    got = _to_host(d)
    assert got[::3].tobytes() == paths[::3].tobytes()
    assert not np.array_equal(got[1::3]["dim"], paths[1::3]["dim"])


def test_errors_are_reported(gpu):
    bh, cam = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    lib = abi.load()
    rays, px, py = pinhole_rays(cam, 4, 4)
    d = _to_device(new_paths(rays, px, py, 0))
    prm = MisOptions(max_depth=2).params()
    # no MIS scene on the handle yet
    rc = lib.gz_gpu_mis_trace_device(bh.handle, abi.ptr(d), 16, ctypes.byref(prm), None)
    assert rc == abi.GZ_E_INVALID and "gz_gpu_set_mis_scene" in abi.last_error()
    integ = VolumePathIntegrator(bh, MisOptions(max_depth=2))
    # unsupported material type
    bad = integ.materials.copy()
    bad["type"][0] = 7
    rc = lib.gz_gpu_set_mis_scene(bh.handle, abi.ptr(bad), len(bad), abi.ptr(integ.lights), len(integ.lights), None)
    assert rc == abi.GZ_E_INVALID
    # a light naming a mesh or a triangle that does not exist
    for field, value in (("mesh", 99), ("triangle", 1 << 40)):
        lights = integ.lights.copy()
        lights[field][0] = value
        rc = lib.gz_gpu_set_mis_scene(bh.handle, abi.ptr(integ.materials), len(integ.materials), abi.ptr(lights),
                                      len(lights), None)
        assert rc == abi.GZ_E_INVALID, field
    # no light at all
    rc = lib.gz_gpu_set_mis_scene(bh.handle, abi.ptr(integ.materials), len(integ.materials), None, 0, None)
    assert rc == abi.GZ_E_INVALID
    # interface materials (materialIndex < 0 on a surface) are refused: the reference's integrator
    # asserts a material at every surface hit (VolumePathIntegrator.swift:341-342)
    from gonzales_amd.accelerator import BoundingHierarchy
    ids = bh.prim_ids.copy()
    surf = np.nonzero(ids["type"] != 3)[0]
    ids["materialIndex"][surf[0]] = -1
    ib = BoundingHierarchy(bh.scene, ids, bh.nodes)
    ib.upload_to_gpu()
    rc = lib.gz_gpu_set_mis_scene(ib.handle, abi.ptr(integ.materials), len(integ.materials), abi.ptr(integ.lights),
                                  len(integ.lights), None)
    assert rc == abi.GZ_E_INVALID and "interface" in abi.last_error()
    ib.free()
    # bad parameters; the scene set before the failed calls stays usable
    prm = MisOptions(max_depth=-1).params()
    assert lib.gz_gpu_mis_trace_device(bh.handle, abi.ptr(d), 16, ctypes.byref(prm), None) == abi.GZ_E_INVALID
    integ.trace(d)
    got = _to_host(d)
    assert not np.any(got["flags"] & abi.GZ_PATH_ACTIVE)


@pytest.mark.parametrize("sky", [None, (0.6, 0.7, 0.8)])
def test_mixed_materials_match_oracle(gpu, sky):
    """Conductor, smooth and rough dielectrics, clear and scattering coated conductors (the
    layered BSDF's random walk draws from the same deterministic streams on both sides)."""
    from mis_helpers import mixed_material_scene
    bh, gz = mixed_material_scene()
    cam = {"eye": (0.0, 1.4, 3.5), "look": (0.1, 0.5, -0.2), "up": (0.0, 1.0, 0.0), "fov": 35.0}
    w = h = 32
    spp = 8
    rays, px, py = pinhole_rays(cam, w, h)
    paths = abi.concat([new_paths(rays, px, py, s) for s in range(spp)], abi.GZ_PATH)
    opt = MisOptions.for_image(w, h, spp, max_depth=5)
    got, ref = _run_both(bh, paths, opt, sky=sky, materials=gz)
    assert np.any(ref["flags"] == 0)
    _check(got, ref, min_frac=0.98)


@pytest.mark.parametrize("queued", [1, 3])
def test_materials_without_the_layered_bsdf_match_oracle(gpu, queued):
    """The mixed scene with its coated conductors replaced by conductors: the bounce kernels are
    then compiled without the layered BSDF (the set a scene without coated conductors gets), in
    the queued megakernel and the full wavefront."""
    from gonzales_amd.integrator import conductor_material
    from mis_helpers import mixed_material_scene
    bh, gz = mixed_material_scene()
    coated = gz["type"] == abi.GZ_MAT_COATED_CONDUCTOR
    assert coated.any()
    gz = gz.copy()
    gz[coated] = conductor_material(roughness=0.35)
    cam = {"eye": (0.0, 1.4, 3.5), "look": (0.1, 0.5, -0.2), "up": (0.0, 1.0, 0.0), "fov": 35.0}
    w = h = 32
    spp = 8
    rays, px, py = pinhole_rays(cam, w, h)
    paths = abi.concat([new_paths(rays, px, py, s) for s in range(spp)], abi.GZ_PATH)
    opt = MisOptions.for_image(w, h, spp, max_depth=5)
    got, ref = _run_both(bh, paths, opt, sky=(0.6, 0.7, 0.8), materials=gz, launch=(queued, 256, 1, 64, 3))
    assert np.any(ref["flags"] == 0)
    _check(got, ref, min_frac=0.98)


LAUNCHES = [(0, 256, 1, 256, 0), (1, 256, 1, 256, 0), (1, 64, 1, 64, 0), (1, 128, 2, 1024, 0), (1, 256, 4, 4096, 0),
            (2, 256, 1, 64, 0), (2, 128, 2, 256, 0), (3, 256, 1, 64, 0), (3, 128, 1, 256, 0), (1, 256, 1, 64, 1),
            (1, 256, 1, 256, 3), (2, 256, 1, 64, 2), (3, 256, 1, 64, 3), (3, 64, 1, 128, 6)]


@pytest.mark.parametrize("mixed", [False, True])
def test_queued_launches_equal_the_full_relaunch(gpu, mixed):
    """Every launch configuration (device path queues compacted by ballot, one lane per path
    every bounce, the split wavefront: a trace kernel sorting hits by material class, then one
    shade kernel per class, or the full wavefront, whose shadow and BSDF rays are traced by their
    own kernels between a shade and a finish kernel) leaves bit-identical path records and work
    counts: the configuration only changes which lane runs a path and when."""
    import torch
    from mis_helpers import mixed_material_scene
    if mixed:
        bh, gz = mixed_material_scene()
        cam = {"eye": (0.0, 1.4, 3.5), "look": (0.1, 0.5, -0.2), "up": (0.0, 1.0, 0.0), "fov": 35.0}
    else:
        bh, cam = scenes.cornell_hierarchy()
        gz = None
    bh.upload_to_gpu()
    w = h = 48
    spp = 4
    rays, px, py = pinhole_rays(cam, w, h)
    paths = abi.concat([new_paths(rays, px, py, s) for s in range(spp)], abi.GZ_PATH)
    paths["flags"][::7] = 0
    opt = MisOptions.for_image(w, h, spp, max_depth=6)
    integ = VolumePathIntegrator(bh, opt, materials=gz, sky=(0.3, 0.4, 0.5) if mixed else None)
    results = []
    for cfg in LAUNCHES:
        integ.set_launch(*cfg)
        assert integ.launch() == cfg
        d = _to_device(paths)
        stats = torch.zeros(len(paths) * 3, dtype=torch.int32, device="cuda")
        integ.trace_stats(d, stats)
        results.append((_to_host(d).tobytes(), stats.cpu().numpy().tobytes()))
        d2 = _to_device(paths)
        integ.trace(d2)
        assert _to_host(d2).tobytes() == results[-1][0], cfg
    bh.check_traversal()
    for cfg, r in zip(LAUNCHES[1:], results[1:]):
        assert r[0] == results[0][0], cfg
        assert r[1] == results[0][1], cfg
    got = np.frombuffer(results[0][0], abi.GZ_PATH)
    assert np.any(got["bounce"] >= 3) and not np.any(got["flags"] & abi.GZ_PATH_ACTIVE)


def test_queue_grows_and_stays_valid_across_sizes(gpu):
    """The per-stream queue storage grows with the path count; a smaller trace after a larger one,
    a trace on another stream and a trace after the storage is released give the same records as
    fresh runs."""
    import torch
    bh, cam = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    rays, px, py = pinhole_rays(cam, 64, 64)
    big = abi.concat([new_paths(rays, px, py, s) for s in range(2)], abi.GZ_PATH)
    opt = MisOptions.for_image(64, 64, 2, max_depth=5)
    integ = VolumePathIntegrator(bh, opt)
    small = big[:1000].copy()
    d_small0 = _to_device(small)
    integ.trace(d_small0)
    ref_small = _to_host(d_small0).tobytes()
    d_big = _to_device(big)
    integ.trace(d_big)
    ref_big = _to_host(d_big).tobytes()
    d_small = _to_device(small)
    integ.trace(d_small)
    assert _to_host(d_small).tobytes() == ref_small
    s = torch.cuda.Stream()
    d_big2 = _to_device(big)
    torch.cuda.synchronize()
    integ.trace(d_big2, stream=s)
    s.synchronize()
    assert _to_host(d_big2).tobytes() == ref_big
    d_part = _to_device(big)
    integ.trace(d_part, count=777)
    got = _to_host(d_part)
    assert got[:777].tobytes() == np.frombuffer(ref_big, abi.GZ_PATH)[:777].tobytes()
    assert got[777:].tobytes() == big[777:].tobytes()
    # released storage is allocated again by the next trace, in either queued mode
    for launch in [(1, 256, 1, 256, 3), (3, 256, 1, 64, 3)]:
        integ.set_launch(*launch)
        integ.release_queues()
        integ.release_queues()  # nothing left to free
        d_again = _to_device(big)
        integ.trace(d_again)
        assert _to_host(d_again).tobytes() == ref_big, launch
    assert abi.load().gz_gpu_release_mis_queues(None) == abi.GZ_E_INVALID


def test_bad_launch_configurations_are_rejected(gpu):
    bh, _ = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    lib = abi.load()
    for cfg in [(4, 256, 1, 256), (-1, 256, 1, 256), (1, 96, 1, 256), (1, 512, 1, 256), (1, 256, 0, 256),
                (1, 256, 17, 256),
                (1, 256, 1, 32), (1, 256, 1, 100), (1, 256, 1, 8192), (1, 256, 1, 256, -1), (1, 256, 1, 256, 65),
                (0, 256, 1, 256, 2)]:
        c = abi.GzMisLaunch(*cfg)
        assert lib.gz_gpu_set_mis_launch(bh.handle, ctypes.byref(c)) == abi.GZ_E_INVALID, cfg
    c = abi.GzMisLaunch(1, 256, 1, 256)
    c.reserved[1] = 5
    assert lib.gz_gpu_set_mis_launch(bh.handle, ctypes.byref(c)) == abi.GZ_E_INVALID
    out = abi.GzMisLaunch()
    assert lib.gz_gpu_get_mis_launch(bh.handle, ctypes.byref(out)) == abi.GZ_OK
    # the default of a small scene: the queued megakernel, the last levels drained
    assert (out.queued, out.block, out.gridScale, out.chunk, out.drainLevel) == (1, 256, 1, 256, 3)


def test_ingested_regression_scene_matches_oracle(gpu):
    """The reference's regression-test scene read from its pbrt file (gonzales_amd.ingest): camera
    paths of a 32x32 crop at 4 spp through every bounce, against the oracle from the same
    starting records."""
    from pathlib import Path
    import torch
    from gonzales_amd.accelerator import build_bounding_hierarchy
    from gonzales_amd.ingest import ingest
    s = ingest(str(Path(__file__).resolve().parent / "golden" / "regression-test.pbrt"))
    s.spp = 4
    bh = build_bounding_hierarchy(s.scene, s.prims)
    bh.upload_to_gpu()
    cam = s.camera(crop=(48, 80, 40, 72))
    opt = MisOptions(max_depth=s.max_depth, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    d = torch.empty(cam.path_count * abi.GZ_PATH.itemsize, dtype=torch.uint8, device="cuda")
    cam.paths_device(d)
    start = _to_host(d).copy()
    got, ref = _run_both(bh, start, opt, materials=s.materials)
    _check(got, ref)
    assert np.mean(ref["estimate"].sum(1) > 0) > 0.5
