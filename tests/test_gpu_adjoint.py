"""Path-replay adjoint (gz_gpu_mis_adjoint_device) against central finite differences of the same
renderer with common random numbers (identical Z-Sobol indices, Russian-roulette and BSDF
streams), loss in float64.

Reflectance: a diffuse path's contribution is a polynomial in the reflectances along it with
directions that do not depend on them, so with Russian roulette out of play (max depth 2) the
replayed derivative and the finite difference agree per parameter to the finite-difference
truncation and float32 image rounding: bar 1% relative (parameters with a gradient above the
noise floor). Roughness: the sampled directions move with roughness in the finite difference
(attached) but not in the adjoint (detached), so the two agree in expectation only. Without
visibility edges (conductor strips under a sky) they agree within 2% per parameter at 4M paths;
with occluders, shadow edges that the moving directions cross add a term the adjoint leaves out,
so there the bar is the gradient's direction (cosine > 0.8) and the signs of its large entries.
"""
import numpy as np
import pytest

from gonzales_amd import abi, scenes
from gonzales_amd.camera import cornell_camera
from gonzales_amd.diff import DiffRenderer, render_function
from gonzales_amd.integrator import MisOptions

pytestmark = pytest.mark.gpu

N_PARAMS = 8


def _setup(res, spp, max_depth, subdiv=4):
    bh, gz, ta, tr = scenes.inverse_rendering_scene(subdivisions=subdiv, n_params=N_PARAMS)
    bh.upload_to_gpu()
    cam = cornell_camera(res=res, spp=spp, seed=0)
    opt = MisOptions(max_depth=max_depth, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    return DiffRenderer(bh, cam, gz, opt), ta, tr


def _loss(img, target):
    d = img.double() - target
    return float((d * d).sum())


def _params(ta, tr, g):
    m = 3 + 2 * N_PARAMS
    albedo = np.full((m, 3), 0.5, np.float32)
    rough = np.full(m, 0.3, np.float32)
    albedo[3:3 + N_PARAMS] = np.clip(ta + g.uniform(-0.2, 0.2, ta.shape), 0.1, 0.9)
    rough[3 + N_PARAMS:] = np.clip(tr + g.uniform(-0.1, 0.1, tr.shape), 0.05, 0.6)
    albedo[:3] = [(0.725, 0.71, 0.68), (0.63, 0.1, 0.1), (0.14, 0.45, 0.1)]
    return albedo, rough


def test_replay_reproduces_the_forward_paths(gpu):
    import torch
    r, ta, tr = _setup(32, 4, 5)
    albedo, rough = _params(ta, tr, np.random.default_rng(1))
    r.set_parameters(albedo, rough)
    r.render()
    fwd = r.d_paths.clone()
    target = torch.zeros(32, 32, 3, dtype=torch.float64, device="cuda")
    img = r.render()
    r.backward(2 * (img.double() - target))
    torch.cuda.synchronize()
    assert torch.equal(r.d_paths, fwd)


def test_reflectance_gradient_matches_finite_differences(gpu):
    import torch
    r, ta, tr = _setup(48, 16, 2)
    g = np.random.default_rng(2)
    albedo, rough = _params(ta, tr, g)
    r.set_parameters(albedo, rough)
    target = torch.from_numpy(g.uniform(0, 0.5, (48, 48, 3))).cuda()
    img = r.render()
    ga, _ = r.backward(2 * (img.double() - target))
    ga = ga.cpu().numpy()
    h = 1e-3
    checked = 0
    for m, c in [(0, 0), (0, 1), (1, 0), (2, 1), (3, 0), (4, 2), (5, 1), (3 + N_PARAMS - 1, 2)]:
        fd = []
        for s in (1, -1):
            a = albedo.copy()
            a[m, c] += s * h
            r.set_parameters(a, rough)
            fd.append(_loss(r.render(), target))
        fd = (fd[0] - fd[1]) / (2 * h)
        print(f"albedo[{m},{c}]: adjoint {ga[m, c]:.6g}  fd {fd:.6g}", flush=True)
        if abs(fd) > 1e-2:
            assert abs(ga[m, c] - fd) <= 1e-2 * abs(fd)
            checked += 1
    assert checked >= 5


def _fd_roughness(r, albedo, rough, ids, target, h=1e-3):
    fd = np.zeros(len(ids))
    for j, m in enumerate(ids):
        vals = []
        for s in (1, -1):
            rr = rough.copy()
            rr[m] += s * h
            r.set_parameters(albedo, rr)
            vals.append(_loss(r.render(), target))
        fd[j] = (vals[0] - vals[1]) / (2 * h)
    r.set_parameters(albedo, rough)
    return fd


def test_roughness_gradient_matches_finite_differences_without_edges(gpu):
    """Conductor strips under a constant sky: no visibility edge, so finite differences with
    common random numbers (whose sampled directions move with the roughness) and the detached
    adjoint estimate the same derivative; they differ per sample, so the image is rendered at 1024
    spp (4M paths) and the bar is 2% per parameter."""
    import torch
    from mis_helpers import conductor_strips_scene
    from gonzales_amd.camera import Film, PerspectiveCamera, look_at
    n = 8
    bh, gz = conductor_strips_scene(n)
    bh.upload_to_gpu()
    cam = PerspectiveCamera(look_at((0.0, 2.5, 2.5), (0.0, 0.0, 0.0), (0, 1, 0)), 55.0, Film(64, 64), 1024)
    opt = MisOptions(max_depth=1, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    r = DiffRenderer(bh, cam, gz, opt, sky=(1.0, 0.9, 0.8))
    albedo = np.full((n, 3), 0.5, np.float32)
    rough = gz["roughness"][:, 0].copy()
    g = np.random.default_rng(3)
    target = torch.from_numpy(g.uniform(0, 0.3, (64, 64, 3))).cuda()
    img = r.render()
    _, gr = r.backward(2 * (img.double() - target))
    gr = gr.cpu().numpy()
    fd = _fd_roughness(r, albedo, rough, range(n), target)
    print("roughness adjoint", np.array2string(gr, precision=5), flush=True)
    print("roughness fd     ", np.array2string(fd, precision=5), flush=True)
    rel = np.abs(gr - fd) / np.abs(fd)
    print(f"relative errors {np.array2string(rel, precision=4)}", flush=True)
    assert np.all(rel <= 0.02)


def test_roughness_gradient_with_occluders(gpu):
    """Cornell box with occluders: visibility edges (shadows) move with the sampled directions in
    the finite difference but not in the adjoint, so the two differ by the edges' share; the
    gradient's direction agrees (cosine > 0.8) and the large components agree in sign."""
    import torch
    r, ta, tr = _setup(64, 64, 2)
    g = np.random.default_rng(3)
    albedo, rough = _params(ta, tr, g)
    r.set_parameters(albedo, rough)
    target = torch.from_numpy(g.uniform(0, 0.5, (64, 64, 3))).cuda()
    img = r.render()
    _, gr = r.backward(2 * (img.double() - target))
    ids = list(range(3 + N_PARAMS, 3 + 2 * N_PARAMS))
    gr = gr.cpu().numpy()[ids]
    fd = _fd_roughness(r, albedo, rough, ids, target)
    print("roughness adjoint", np.array2string(gr, precision=4), flush=True)
    print("roughness fd     ", np.array2string(fd, precision=4), flush=True)
    cos = float(gr @ fd / (np.linalg.norm(gr) * np.linalg.norm(fd)))
    print(f"cosine {cos:.4f}, vector relative error {np.linalg.norm(gr - fd) / np.linalg.norm(fd):.4f}", flush=True)
    assert cos > 0.8
    big = np.abs(fd) > 0.3 * np.abs(fd).max()
    assert np.all(np.sign(gr[big]) == np.sign(fd[big]))


def test_autograd_descent_reduces_the_loss(gpu):
    """A few Adam steps on the reflectances through torch.autograd lower the image loss."""
    import torch
    r, ta, tr = _setup(32, 8, 3)
    g = np.random.default_rng(4)
    albedo0, rough0 = _params(ta, tr, g)
    a_t, r_t = albedo0.copy(), rough0.copy()
    a_t[3:3 + N_PARAMS] = ta
    r.set_parameters(a_t, r_t)
    target = r.render().double()
    f = render_function()
    a = torch.tensor(albedo0, device="cuda", requires_grad=True)
    rg = torch.tensor(rough0, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([a], lr=0.02)
    losses = []
    for _ in range(8):
        opt.zero_grad()
        img = f.apply(a, rg, r)
        loss = ((img.double() - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            a.clamp_(0.05, 0.95)
        losses.append(float(loss))
    print("losses", [round(x, 4) for x in losses], flush=True)
    assert losses[-1] < 0.7 * losses[0]
    assert rg.grad is not None and a.grad is not None


def test_queued_adjoint_equals_the_full_relaunch(gpu):
    """The replay over device path queues (with and without the drain launch) and the
    one-lane-per-path replay give the same gradient (to the order of the double atomics) and the
    same path records."""
    import torch
    r, ta, tr = _setup(32, 8, 4)
    albedo, rough = _params(ta, tr, np.random.default_rng(5))
    r.set_parameters(albedo, rough)
    target = torch.zeros(32, 32, 3, dtype=torch.float64, device="cuda")
    out = []
    # relaunch, queued without and with the drain from level 2, the full wavefront forward
    # (the adjoint then replays the queued form) drained from level 3
    launches = [(0, 256, 1, 256, 0), (1, 256, 1, 256, 0), (1, 256, 1, 64, 2), (3, 256, 1, 64, 3)]
    for cfg in launches:
        r.integrator.set_launch(*cfg)
        img = r.render()
        ga, gr = r.backward(2 * (img.double() - target))
        out.append((img.cpu().numpy(), ga.cpu().numpy(), gr.cpu().numpy(), r.d_paths.cpu().numpy().tobytes()))
    r.integrator.set_launch()
    for cfg, o in zip(launches[1:], out[1:]):
        assert np.array_equal(o[0], out[0][0]), cfg
        assert o[3] == out[0][3], cfg
        for k in (1, 2):
            a, b = o[k], out[0][k]
            assert np.all(np.abs(a - b) <= 1e-9 * np.abs(b).max()), (cfg, k)
    assert np.abs(out[0][1]).max() > 0 and np.abs(out[0][2]).max() > 0
