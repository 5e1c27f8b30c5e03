"""Path-replay adjoint (gz_gpu_mis_adjoint_device) against central finite differences (h = 1e-3)
with common random numbers (identical Z-Sobol indices, Russian-roulette and BSDF streams), loss in
float64, on the device renderer, whose finite differences equal the oracle's (the CPU restatement
of the reference renderer) on the same samples (test_oracle_finite_differences_equal_the_device's).

Where the per-sample estimator is smooth in the parameters the two agree sample set by sample set:
- reflectance without roulette: a diffuse path's contribution is a polynomial in the reflectances
  along it with directions that do not depend on them (bar 1%);
- roughness under a triangle light with no occluder: the light strategy's directions do not move
  and the derivative goes through the BSDF value and the power heuristic's BSDF pdf (bar 0.1%).
Where the sampled directions move with the roughness (attached in the finite difference, detached in
the adjoint) the two agree in expectation only: strips under a constant sky, 2% at 4M paths.
With visibility edges (the inverse-rendering scene's occluders) the finite difference also picks up
edge crossings, each a jump of about dL / 2h with probability O(h), and with Russian roulette the
same for survival decisions that flip between the +h and -h renders: its noise grows as 1 / (h N)
and is heavy-tailed, so there the comparison sums K independent Z-Sobol seeds and the bar carries
the measured noise floor (the spread of the per-seed differences).
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


def test_recorded_replay_equals_traced_replay(gpu):
    """The backward pass from the forward pass's replay record (gz_gpu_mis_adjoint_record_device: the
    recorded levels' rays are not traced again) gives the traced replay's gradients and leaves the
    same path records, with a record of the levels below the drain level, of every level, and with
    a launch that records nothing; the forward images are the same with and without the record."""
    import torch
    res, spp, depth = 32, 8, 5
    g = np.random.default_rng(7)
    r, ta, tr = _setup(res, spp, depth)
    albedo, rough = _params(ta, tr, g)
    target = torch.from_numpy(g.uniform(0, 0.5, (res, res, 3))).cuda()
    plain = DiffRenderer(r.bh, r.camera, r.materials, r.options, record=False)
    for launch, want in (((3, 256, 1, 64, 3), 3), ((3, 256, 1, 64, 0), depth + 1), ((1, 256, 1, 256, 3), 0)):
        for rr in (r, plain):
            rr.set_parameters(albedo, rough)
            rr.integrator.set_launch(*launch)
        img = r.render()
        img0 = plain.render()
        assert r.recorded_levels == want and plain.recorded_levels == 0
        assert torch.equal(img, img0), launch
        ga, gr = r.backward(2 * (img.double() - target))
        after = r.d_paths.clone()
        ga0, gr0 = plain.backward(2 * (img0.double() - target))
        torch.cuda.synchronize()
        assert torch.equal(after, plain.d_paths), launch
        for x, y in ((ga, ga0), (gr, gr0)):
            x, y = x.cpu().numpy(), y.cpu().numpy()
            assert np.abs(y).max() > 0
            # the same terms, added in another order by the float64 atomics
            np.testing.assert_allclose(x, y, rtol=1e-9, atol=1e-12 * np.abs(y).max(), err_msg=str(launch))
    r.integrator.set_launch(3, 256, 1, 64, 3)


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


def test_roughness_gradient_under_a_triangle_light(gpu):
    """Conductor strips lit only by a triangle area light they all see whole (no sky, no occluder):
    the light strategy's directions do not depend on the roughness and the BSDF strategy carries
    weight 0, so the image is smooth in the roughness and the finite difference and the adjoint
    estimate the same derivative sample by sample, through the BSDF value and the power
    heuristic's BSDF pdf (adjoint_vertex's MIS-weight derivative)."""
    import torch
    from mis_helpers import conductor_strips_under_light
    from gonzales_amd.camera import Film, PerspectiveCamera, look_at
    n = 8
    bh, gz = conductor_strips_under_light(n)
    bh.upload_to_gpu()
    cam = PerspectiveCamera(look_at((0.0, 2.5, 2.5), (0.0, 0.0, 0.0), (0, 1, 0)), 55.0, Film(64, 64), 16)
    opt = MisOptions(max_depth=1, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    r = DiffRenderer(bh, cam, gz, opt)
    albedo = np.full((n, 3), 0.5, np.float32)
    rough = gz["roughness"][:, 0].copy()
    target = torch.from_numpy(np.random.default_rng(3).uniform(0, 0.3, (64, 64, 3))).cuda()
    img = r.render()
    _, gr = r.backward(2 * (img.double() - target))
    gr = gr.cpu().numpy()
    fd = _fd_roughness(r, albedo, rough, range(n), target)
    print("roughness adjoint", np.array2string(gr, precision=6), flush=True)
    print("roughness fd     ", np.array2string(fd, precision=6), flush=True)
    assert np.abs(fd).max() > 100
    assert np.all(np.abs(gr - fd) <= 1e-3 * np.abs(fd) + 1e-3)


# The inverse-rendering workload: Cornell box + 200K occluder triangles, 32 diffuse and 32 conductor
# parameter materials; target rendered at the hidden parameters (256 spp, its own seed), parameters
# at their initial 0.5, L2 image loss.
C5_PARAMS = 32
C5_ROUGH = [3 + C5_PARAMS + k for k in range(0, C5_PARAMS, 4)]
C5_ALBEDO = [(3 + k, (k // 4) % 3) for k in range(0, C5_PARAMS, 4)]


class Config5:
    def __init__(self, res, max_depth):
        import torch
        self.bh, self.gz, ta, tr = scenes.inverse_rendering_scene(subdivisions=6, n_params=C5_PARAMS)
        self.bh.upload_to_gpu()
        self.res, self.max_depth = res, max_depth
        m = 3 + 2 * C5_PARAMS
        self.albedo = np.full((m, 3), 0.5, np.float32)
        self.albedo[:3] = [(0.725, 0.71, 0.68), (0.63, 0.1, 0.1), (0.14, 0.45, 0.1)]
        self.rough = np.full(m, 0.5, np.float32)
        a_t, r_t = self.albedo.copy(), self.rough.copy()
        a_t[3:3 + C5_PARAMS] = ta
        r_t[3 + C5_PARAMS:] = tr
        rt = self.renderer(256, 99)
        rt.set_parameters(a_t, r_t)
        self.target = rt.render().double()
        del rt
        torch.cuda.empty_cache()

    def renderer(self, spp, seed):
        cam = cornell_camera(res=self.res, spp=spp, seed=seed)
        opt = MisOptions(max_depth=self.max_depth, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                         base4_digits=cam.c.zsBase4Digits)
        return DiffRenderer(self.bh, cam, self.gz, opt)

    def loss(self, img):
        import torch
        d = torch.as_tensor(img, device=self.target.device).double() - self.target
        return float((d * d).sum())

    def gradients(self, r):
        """(adjoint, finite differences) over C5_ROUGH then C5_ALBEDO for renderer r."""
        r.set_parameters(self.albedo, self.rough)
        img = r.render()
        ga, gr = r.backward(2 * (img.double() - self.target))
        ga, gr = ga.cpu().numpy(), gr.cpu().numpy()
        adj = np.concatenate([gr[C5_ROUGH], [ga[k, c] for k, c in C5_ALBEDO]])
        return adj, self.finite_differences(lambda a, rr: self.loss(self._render(r, a, rr)))

    @staticmethod
    def _render(r, albedo, rough):
        r.set_parameters(albedo, rough)
        return r.render()

    def finite_differences(self, loss_of, h=1e-3):
        fd = []
        for k in C5_ROUGH:
            v = []
            for s in (1, -1):
                rr = self.rough.copy()
                rr[k] += s * h
                v.append(loss_of(self.albedo, rr))
            fd.append((v[0] - v[1]) / (2 * h))
        for k, c in C5_ALBEDO:
            v = []
            for s in (1, -1):
                aa = self.albedo.copy()
                aa[k, c] += s * h
                v.append(loss_of(aa, self.rough))
            fd.append((v[0] - v[1]) / (2 * h))
        return np.array(fd)

    def seeds(self, spp, k):
        """Per-seed (adjoint, finite differences) for Z-Sobol seeds 0..k-1."""
        import torch
        adj, fd = [], []
        for seed in range(k):
            r = self.renderer(spp, seed)
            a, f = self.gradients(r)
            adj.append(a)
            fd.append(f)
            del r
            torch.cuda.empty_cache()
        return np.array(adj), np.array(fd)


def _check_summed(adj, fd, vec_bar):
    """Sums over seeds; per parameter |adjoint - fd| <= 2% |fd| + 3 x the noise floor (the spread of
    the per-seed differences, scaled to the sum), and the vector error of each group <= vec_bar."""
    k = len(adj)
    a, f = adj.sum(0), fd.sum(0)
    noise = (adj - fd).std(0, ddof=1) * np.sqrt(k)
    n = len(C5_ROUGH)
    for name, sl in (("roughness", slice(0, n)), ("albedo", slice(n, None))):
        vec = np.linalg.norm(a[sl] - f[sl]) / np.linalg.norm(f[sl])
        print(f"{name}: adjoint {np.array2string(a[sl], precision=4)}", flush=True)
        print(f"{name}: fd      {np.array2string(f[sl], precision=4)}", flush=True)
        print(f"{name}: noise   {np.array2string(noise[sl], precision=4)}  vector error {vec:.4f}", flush=True)
        assert vec <= vec_bar, (name, vec)
        assert np.all(np.abs(a[sl] - f[sl]) <= 0.02 * np.abs(f[sl]) + 3 * noise[sl]), name


def test_oracle_finite_differences_equal_the_device(gpu):
    """The finite differences of the oracle's renderer (camera paths -> every MIS bounce -> film, on
    the CPU) equal the device renderer's on the same samples, and the reflectance adjoint matches
    them within 1% per parameter (the roughness comparison needs more samples than the oracle
    renders in seconds, so it runs on the device renderer below)."""
    from mis_helpers import oracle_render
    c5 = Config5(64, 2)
    r = c5.renderer(16, 0)
    adj, fd_dev = c5.gradients(r)

    def oracle_loss(albedo, rough):
        r.set_parameters(albedo, rough)
        img = oracle_render(c5.bh, r.camera, r.options, r.integrator.materials, r.integrator.lights, nthreads=16)
        return c5.loss(img)

    fd_orc = c5.finite_differences(oracle_loss)
    print("fd device", np.array2string(fd_dev, precision=5), flush=True)
    print("fd oracle", np.array2string(fd_orc, precision=5), flush=True)
    print("adjoint  ", np.array2string(adj, precision=5), flush=True)
    # paths on which device and oracle round differently (at most 1% of them, the path bar of the
    # forward tests) move the difference by their change over 2h where they cross an edge
    assert np.all(np.abs(fd_dev - fd_orc) <= 1e-2 * np.abs(fd_orc) + 1e-4)
    n = len(C5_ROUGH)
    a, f = adj[n:], fd_orc[n:]
    big = np.abs(f) >= 0.1
    assert big.sum() >= 2
    assert np.all(np.abs(a - f)[big] <= 2e-3 * np.abs(f)[big])
    # small reflectance gradients: plus 5e-4, the float32 image rounding that the difference of two
    # renders 2h apart magnifies (about 3e-5 of a pixel value per pixel, over 4096 pixels)
    assert np.all(np.abs(a - f) <= 1e-2 * np.abs(f) + 5e-4)


def test_config5_gradients_match_central_differences(gpu):
    """Inverse-rendering scene at max depth 2 (no roulette), 128^2 x 1024 spp over 8 seeds (134M
    paths): roughness and reflectance gradients within 2% of the central differences, per
    parameter above the measured noise floor and as vectors."""
    adj, fd = Config5(128, 2).seeds(1024, 8)
    _check_summed(adj, fd, 0.02)


def test_config5_gradients_with_russian_roulette(gpu):
    """The same at max depth 5, where roulette decides after bounce 1: survival decisions that flip
    between the +h and -h renders add heavy-tailed jumps to the finite difference (measured vector
    error 2-3% at 16 seeds, falling with more), so the vector bar is 5% and the per-parameter bar
    carries the noise floor."""
    adj, fd = Config5(128, 5).seeds(1024, 16)
    _check_summed(adj, fd, 0.05)


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
