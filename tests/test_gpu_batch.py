"""Multi-scene batch (gz_batch_*): every scene's film equals, bit for bit, the film of the same
scene rendered alone with FrameRenderer (camera paths -> MIS trace -> film add), in eager and
graph mode, across rounds that split scenes, and on repeated graph replays with other device
work in between."""
import numpy as np
import pytest

from gonzales_amd import abi, scenes
from gonzales_amd.batch import BatchEntry, SceneBatch, autotune_batch
from gonzales_amd.camera import DeviceFilm, Film, PerspectiveCamera, cornell_camera, look_at
from gonzales_amd.integrator import MisOptions, VolumePathIntegrator
from gonzales_amd.render import FrameRenderer

pytestmark = pytest.mark.gpu


def _options(cam, max_depth):
    return MisOptions(max_depth=max_depth, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                      base4_digits=cam.c.zsBase4Digits)


def _scenes():
    from mis_helpers import mixed_material_scene
    out = []
    bh, _ = scenes.cornell_hierarchy()
    bh.upload_to_gpu()
    cam = cornell_camera(res=40, spp=4, seed=3)
    integ = VolumePathIntegrator(bh, _options(cam, 5))
    out.append((integ, cam))
    cam2 = cornell_camera(res=36, spp=2, seed=9, crop=(5, 30, 3, 20))
    out.append((VolumePathIntegrator(bh, _options(cam2, 5)), cam2))
    mb, gz = mixed_material_scene()
    mb.upload_to_gpu()
    cam3 = PerspectiveCamera(look_at((0.0, 1.4, 3.5), (0.1, 0.5, -0.2), (0, 1, 0)), 35.0, Film(32, 24), 8, 5)
    out.append((VolumePathIntegrator(mb, _options(cam3, 6), sky=(0.6, 0.7, 0.8), materials=gz), cam3))
    # two scenes above 65536 primitives: the wavefront kernels bounce their items in wavefront rounds
    for k, kind in enumerate(("sphere", "soup")):
        bh, gz, (eye, at, up, fov) = scenes.batch_scene({"index": 90 + k, "kind": kind, "tris": 70000, "res": 20,
                                                          "spp": 4, "max_depth": 6, "seed": 4242 + k})
        bh.upload_to_gpu()
        cam = PerspectiveCamera(look_at(eye, at, up), fov, Film(20, 18), 4, 11 + k)
        out.append((VolumePathIntegrator(bh, _options(cam, 6), sky=(0.4, 0.45, 0.5), materials=gz), cam))
    return out


def _films(entries):
    return [np.concatenate([t.cpu().numpy() for t in (e.film.light, e.film.weight, e.film.albedo, e.film.normal)])
            for e in entries]


def test_batch_films_equal_single_scene_renders(gpu):
    import torch
    sc = _scenes()
    # Scenes 0 and 1 share one handle and MIS tables; the reference renders use the same.
    refs = []
    for integ, cam in sc:
        f = DeviceFilm(cam)
        FrameRenderer(cam, integ, f, batch_paths=1 << 12).render()
        torch.cuda.synchronize()
        refs.append(np.concatenate([t.cpu().numpy() for t in (f.light, f.weight, f.albedo, f.normal)]))
    entries = [BatchEntry(integ, cam, DeviceFilm(cam)) for integ, cam in sc]
    batch = SceneBatch(entries, round_paths=3000)
    info = batch.info()
    assert info.totalPaths == sum(cam.path_count for _, cam in sc)
    assert info.rounds > 3 and info.levels == 7 and info.allBsdfs == 1 and info.soloScenes == 0
    for mode in (False, True, True):
        batch.clear_films()
        batch.render(use_graph=mode)
        junk = torch.randn(1 << 16, device="cuda").sum()  # other device work between replays
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for k, (got, ref) in enumerate(zip(_films(entries), refs)):
            assert got.tobytes() == ref.tobytes(), (mode, k)
    for shape in [(128, 2, 64, 3), (256, 1, 256, 0), (256, 1, 128, 1), (256, 1, 64, 3, 3), (128, 1, 256, 0, 3),
                  (64, 2, 128, 2, 3)]:
        batch.set_launch(*shape)
        batch.clear_films()
        batch.render(use_graph=True)
        torch.cuda.synchronize()
        for k, (got, ref) in enumerate(zip(_films(entries), refs)):
            assert got.tobytes() == ref.tobytes(), ("reshaped", shape, k)
    best, timings = autotune_batch(batch, reps=1)
    assert best in timings
    batch.render(use_graph=True)
    torch.cuda.synchronize()
    for k, (got, ref) in enumerate(zip(_films(entries), refs)):
        assert got.tobytes() == ref.tobytes(), ("tuned", k)
    for e in entries:
        img = e.film.resolve()
        assert torch.isfinite(img).all()
    batch.close()


def test_batch_scenes_that_render_alone(gpu):
    """Scenes above the solo threshold render with the single-scene launch sequence after the
    rounds, in pieces of at most round_paths paths, with one queue storage for all of them: films
    bit-identical to FrameRenderer's, with some scenes alone and the rest in rounds, and with every
    scene alone (no rounds at all)."""
    import torch
    sc = _scenes()
    refs = []
    for integ, cam in sc:
        f = DeviceFilm(cam)
        FrameRenderer(cam, integ, f, batch_paths=1 << 12).render()
        torch.cuda.synchronize()
        refs.append(np.concatenate([t.cpu().numpy() for t in (f.light, f.weight, f.albedo, f.normal)]))
    counts = [cam.path_count for _, cam in sc]
    for solo_paths, want_solo in ((2000, sum(c > 2000 for c in counts)), (1, len(sc))):
        entries = [BatchEntry(integ, cam, DeviceFilm(cam)) for integ, cam in sc]
        # round_paths 3000: the 6400-path scene renders alone in pieces of 3000, 3000 and 400
        batch = SceneBatch(entries, round_paths=3000, solo_paths=solo_paths)
        info = batch.info()
        assert info.soloScenes == want_solo and 0 < want_solo
        assert 0 < info.sideScenes <= want_solo  # side lanes in use
        assert info.totalPaths == sum(counts)
        assert (info.rounds == 0) == (want_solo == len(sc))
        for mode in (False, True, True):
            batch.clear_films()
            batch.render(use_graph=mode)
            junk = torch.randn(1 << 16, device="cuda").sum()
            torch.cuda.synchronize()
            assert float(junk) == float(junk)
            for k, (got, ref) in enumerate(zip(_films(entries), refs)):
                assert got.tobytes() == ref.tobytes(), (solo_paths, mode, k)
        batch.check()
        batch.close()
    # the Cornell frame as two crops adding to ONE film, both alone: the second would go to an
    # empty side lane by path count, but shares the first one's lane (no two lanes add to one
    # film), and the film equals the whole frame's
    integ, cam = sc[0]
    top = cornell_camera(res=40, spp=4, seed=3, crop=(0, 40, 0, 17))
    bottom = cornell_camera(res=40, spp=4, seed=3, crop=(0, 40, 17, 40))
    shared = DeviceFilm(cam)
    others = [BatchEntry(i, c, DeviceFilm(c)) for i, c in sc[2:]]
    entries = [BatchEntry(integ, top, shared), BatchEntry(integ, bottom, shared)] + others
    batch = SceneBatch(entries, round_paths=3000, solo_paths=1)
    info = batch.info()
    assert info.soloScenes == len(entries) and 0 < info.sideScenes < len(entries)
    for mode in (False, True):
        batch.clear_films()
        batch.render(use_graph=mode)
        torch.cuda.synchronize()
        got = _films([entries[0]] + others)
        want = [refs[0]] + refs[2:]
        for k, (g, r) in enumerate(zip(got, want)):
            assert g.tobytes() == r.tobytes(), ("crops", mode, k)
    batch.check()
    batch.close()
    with pytest.raises(RuntimeError, match="soloPaths"):
        SceneBatch([BatchEntry(integ, cam, DeviceFilm(cam)) for integ, cam in sc], solo_paths=-1)


def test_frame_renderer_lanes_bit_identical(gpu):
    """FrameRenderer's lanes (pixel ranges on their own streams) and sequence sizes change no bit
    of the film: 1, 2 and 3 lanes, whole-frame and small sequences, one handle shared."""
    import torch
    sc = _scenes()
    for k, (integ, cam) in enumerate(sc):
        films = []
        for lanes, batch in ((1, 1 << 20), (2, 1 << 20), (3, 1 << 11), (2, 192)):
            f = DeviceFilm(cam)
            rnd = FrameRenderer(cam, integ, f, batch_paths=batch, lanes=lanes)
            assert rnd.lanes == lanes and len(rnd.side) == lanes - 1
            b = rnd.bounds
            assert b[0] == 0 and b[-1] == cam.path_count and all(y - x >= cam.spp for x, y in zip(b, b[1:]))
            assert all(x % cam.spp == 0 for x in b)
            rnd.render()
            rnd.render()  # twice: the second frame adds onto the first
            torch.cuda.synchronize()
            rnd.check()
            films.append(np.concatenate([t.cpu().numpy() for t in (f.light, f.weight, f.albedo, f.normal)]))
        for other in films[1:]:
            assert other.tobytes() == films[0].tobytes(), k
    integ, cam = sc[0]
    assert FrameRenderer(cam, integ, DeviceFilm(cam)).lanes == 1  # automatic: one lane below 4M paths


def test_batch_rejects_bad_input(gpu):
    sc = _scenes()
    entries = [BatchEntry(integ, cam, DeviceFilm(cam)) for integ, cam in sc]
    with pytest.raises(RuntimeError, match="roundPaths"):
        SceneBatch(entries, round_paths=4)
    batch = SceneBatch(entries, round_paths=1 << 14)
    import ctypes
    for cfg in [(0, 256, 1, 256), (2, 256, 1, 256), (1, 256, 1, 100), (1, 512, 1, 256), (1, 256, 1, 256, -1),
                (1, 256, 1, 256, 65)]:
        c = abi.GzMisLaunch(*cfg)
        assert abi.load().gz_batch_set_launch(batch.handle, ctypes.byref(c)) == abi.GZ_E_INVALID
    info = batch.info()
    assert info.roundPaths <= 1 << 14
    batch.render(use_graph=True)
    batch.check()
    # replacing a scene's MIS tables after the batch was created: the batch's copies of its table
    # pointers are stale, so the batch refuses to render until it is created again
    integ0 = sc[0][0]
    abi.check(abi.load().gz_gpu_set_mis_scene(integ0.bh.handle, abi.ptr(integ0.materials), len(integ0.materials),
                                              abi.ptr(integ0.lights), len(integ0.lights), None))
    with pytest.raises(abi.GzError, match="replaced"):
        batch.render(use_graph=True)
    batch.close()
    again = SceneBatch(entries, round_paths=1 << 14)
    again.render(use_graph=True)
    again.check()
    again.close()
    # slots for the largest round only: one round holds every path of a small batch
    big = SceneBatch(entries, round_paths=1 << 30)
    info = big.info()
    assert info.rounds == 1 and info.roundPaths == info.totalPaths
    big.close()
