"""Scenes with known answers for the MIS integrator, and an independent brute-force estimator.

The brute-force estimator traces paths by cosine-weighted BSDF sampling only (no light sampling, no
MIS, no roulette), with its own numpy geometry and random numbers; only the ray casts go through
the oracle's traversal, which is pinned by its own tests. The integrator's estimate must agree with
it in expectation wherever the reference's estimator is unbiased (the constant sky).
"""
from __future__ import annotations

import ctypes

import numpy as np

import oracle
from gonzales_amd import abi
from gonzales_amd.accelerator import Primitives, Scene, TriangleMesh, build_bounding_hierarchy
from gonzales_amd.integrator import area_lights_from_scene, materials_from_scene, new_paths


def _materials(albedos) -> np.ndarray:
    m = np.zeros(len(albedos), abi.MATERIAL)
    m["type"] = 1
    m["albedo"] = albedos
    return m


def quad(p0, p1, p2, p3):
    return np.array([p0, p1, p2, p3], np.float32), np.array([0, 1, 2, 0, 2, 3], np.int64)


def plane_scene(albedo=(0.5, 0.25, 0.75), half=1000.0):
    """A diffuse square at y = 0 with its normal up (two triangles)."""
    v, i = quad([-half, 0, -half], [-half, 0, half], [half, 0, half], [half, 0, -half])
    mesh = TriangleMesh.from_xyz(v, i)
    return build_bounding_hierarchy(Scene([mesh], _materials([albedo])), Primitives.geometric(0, 2, 0))


def box_on_plane_scene():
    """The plane plus a rotated box (10 triangles) of another albedo standing on it."""
    from gonzales_amd.scenes import _box
    v, i = quad([-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50])
    bv, bi = _box(0.1, -0.2, 0.5, 1.0, 30.0, with_floor=False)
    meshes = [TriangleMesh.from_xyz(v, i),
              TriangleMesh.from_xyz(np.asarray(bv, np.float32), np.asarray(bi, np.int64))]
    prims = Primitives.concat([Primitives.geometric(0, 2, 0), Primitives.geometric(1, len(bi) // 3, 1)])
    return build_bounding_hierarchy(Scene(meshes, _materials([(0.6, 0.5, 0.4), (0.3, 0.7, 0.5)])), prims)


def conductor_strips_scene(n: int = 8, roughness=None):
    """n conductor strips side by side on the plane y = 0 (x in [-2, 2], z in [-2, 2]), strip i
    with material i (Cu, roughness from `roughness`). No occluder and no area light: lit by a
    constant sky, the image has no visibility edge. Returns (hierarchy, GzMaterial table)."""
    from gonzales_amd.integrator import conductor_material
    rough = np.linspace(0.1, 0.5, n) if roughness is None else np.asarray(roughness)
    verts, idx = [], []
    for i in range(n):
        x0, x1 = -2 + 4 * i / n, -2 + 4 * (i + 1) / n
        v, q = quad([x0, 0, -2], [x0, 0, 2], [x1, 0, 2], [x1, 0, -2])
        idx.append(q + 4 * i)
        verts.append(v)
    mesh = TriangleMesh.from_xyz(np.concatenate(verts), np.concatenate(idx))
    prims = Primitives.geometric(0, 2 * n, np.repeat(np.arange(n), 2))
    gz = abi.concat([conductor_material(roughness=float(r)) for r in rough], abi.GZ_MATERIAL)
    scene = Scene([mesh], _materials([(0.5, 0.5, 0.5)] * n))
    return build_bounding_hierarchy(scene, prims), gz


def mixed_material_scene():
    """Diffuse plane, a rotated box whose 10 triangles cycle through a conductor, smooth and rough
    dielectrics and two coated conductors (clear and scattering coat), and one triangle area light
    above. Returns (hierarchy, GzMaterial table)."""
    from gonzales_amd.integrator import (coated_conductor_material, conductor_material, dielectric_material,
                                         diffuse_material)
    from gonzales_amd.scenes import _box
    gz = abi.concat([diffuse_material((0.6, 0.5, 0.4)), conductor_material(roughness=0.2),
                     dielectric_material(eta=1.5, roughness=0.0), dielectric_material(eta=1.33, roughness=0.1),
                     coated_conductor_material(interface_roughness=0.05, conductor_roughness=0.3),
                     coated_conductor_material(reflectance=(0.9, 0.6, 0.3), albedo=(0.5, 0.7, 0.9), g=0.3,
                                               thickness=0.2, interface_roughness=0.2, conductor_roughness=0.1)],
                    abi.GZ_MATERIAL)
    v, i = quad([-50, 0, -50], [-50, 0, 50], [50, 0, 50], [50, 0, -50])
    bv, bi = _box(0.1, -0.2, 0.5, 1.0, 30.0, with_floor=False)
    tri = np.array([[-0.6, 2.2, -0.4], [0.9, 2.2, -0.5], [0.1, 2.2, 0.8]], np.float32)  # normal down
    meshes = [TriangleMesh.from_xyz(v, i),
              TriangleMesh.from_xyz(np.asarray(bv, np.float32), np.asarray(bi, np.int64)),
              TriangleMesh.from_xyz(tri, np.arange(3, dtype=np.int64))]
    box_mats = 1 + (np.arange(len(bi) // 3) % 5)
    prims = Primitives.concat([Primitives.geometric(0, 2, 0), Primitives.geometric(1, len(bi) // 3, box_mats),
                               Primitives.area_lights(2, 1, 0)])
    scene = Scene(meshes, _materials([(0.5, 0.5, 0.5)] * len(gz)),
                  light_emission=np.array([[6.0, 5.0, 4.0]], np.float32))
    return build_bounding_hierarchy(scene, prims), gz


# The light triangle of light_over_plane: corners above the plane at height 1.
LIGHT_TRI = np.array([[-0.3, 1.0, -0.2], [0.4, 1.0, -0.1], [0.1, 1.0, 0.5]], np.float64)


def light_over_plane(facing_down: bool, radiance=(4.0, 2.0, 1.0), albedo=(0.5, 0.25, 0.75)):
    """The plane plus one triangle area light at y = 1 (a closed-form direct-lighting setup)."""
    v, i = quad([-100, 0, -100], [-100, 0, 100], [100, 0, 100], [100, 0, -100])
    tri = LIGHT_TRI.astype(np.float32)
    # normal = cross(p1 - p0, p2 - p0); order the corners so it points where asked
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    if (n[1] < 0) != facing_down:
        tri = tri[[0, 2, 1]]
    meshes = [TriangleMesh.from_xyz(v, i), TriangleMesh.from_xyz(tri, np.arange(3, dtype=np.int64))]
    prims = Primitives.concat([Primitives.geometric(0, 2, 0), Primitives.area_lights(1, 1, 0)])
    scene = Scene(meshes, _materials([albedo]), light_emission=np.array([radiance], np.float32))
    return build_bounding_hierarchy(scene, prims), tri.astype(np.float64)


class OracleMis:
    """Runs the oracle integrator on a hierarchy (keeps the ctypes views alive)."""

    def __init__(self, bh, sky=None, materials=None, lights=None):
        self.bh = bh
        self.desc, self._meshes = bh.descriptor()
        self.materials = materials_from_scene(bh) if materials is None else materials
        self.lights = area_lights_from_scene(bh) if lights is None else lights
        self.sky = sky
        self.table = oracle.sobol_matrices()

    def params(self, max_depth, fix=False, fix_sky=False, seed=0, log2spp=0, digits=0,
               rr_seed=0x5EED) -> oracle.MisParams:
        sky = (0.0, 0.0, 0.0) if self.sky is None else tuple(float(x) for x in self.sky)
        flags = (abi.GZ_MIS_FIX_LIGHT_PDF if fix else 0) | (abi.GZ_MIS_FIX_SKY_DIRECTION if fix_sky else 0)
        return oracle.MisParams(max_depth, flags, seed, log2spp, digits, rr_seed,
                                (ctypes.c_float * 3)(*sky), 0 if self.sky is None else 1)

    def trace(self, paths, params, nthreads=8):
        oracle.mis_trace(ctypes.addressof(self.desc), len(self.bh.nodes), self.materials, self.lights, params,
                         paths, nthreads, self.table)


def point_paths(org, direction, n: int, log2spp: int, start_dim: int = 0) -> np.ndarray:
    """n paths of one pixel (0, 0) with sample indices 0..n-1, all along one camera ray."""
    rays = np.zeros(n, abi.RAY)
    rays["orgX"], rays["orgY"], rays["orgZ"] = org
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    rays["dirX"], rays["dirY"], rays["dirZ"] = d
    assert n <= (1 << log2spp)
    return new_paths(rays, 0, 0, np.arange(n), start_dim=start_dim)


def direct_quadrature(point, albedo, radiance, tri, light_facing_down: bool, fix: bool, grid: int = 1200):
    """Expected NEE + MIS-BSDF contribution at a point on the plane (normal +y) from one triangle
    light, by midpoint quadrature over the triangle in float64:
      light strategy : f cos_s L cos_l / d^2 * w_l,  w_l = p_l^2 / (p_l^2 + p_b^2)
      BSDF strategy  : f cos_s L cos_l / d^2 * w_b,  w_b = p_b^2 / (p_b^2 + p_l'^2)
    with p_l = d^2 / (cos_l A), p_b = cos_s / pi. The light strategy sees the emitting side
    (light facing the point); the BSDF strategy sees radiance only when the light faces away
    from the point, and its weight uses p_l' = p_l with the pdf correction, else 0 (weight 0)."""
    p0, p1, p2 = (np.asarray(t, np.float64) for t in tri)
    area = 0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0))
    # midpoint rule over the triangle in barycentric (u, v), u + v < 1
    k = (np.arange(grid) + 0.5) / grid
    u, v = np.meshgrid(k, k, indexing="ij")
    keep = u + v < 1
    u, v = u[keep], v[keep]
    x = p0[None] + u[:, None] * (p1 - p0)[None] + v[:, None] * (p2 - p0)[None]
    dA = area / len(u)
    d = x - np.asarray(point, np.float64)[None]
    d2 = np.einsum("ij,ij->i", d, d)
    w = d / np.sqrt(d2)[:, None]
    cos_s = np.clip(w[:, 1], 0, None)
    cos_l = np.abs(w[:, 1])  # light plane is horizontal
    f = np.asarray(albedo, np.float64) / np.pi
    L = np.asarray(radiance, np.float64)
    geom = cos_s * cos_l / d2 * dA
    p_l = d2 / (cos_l * area)
    p_b = cos_s / np.pi
    if light_facing_down:
        weight = p_l ** 2 / (p_l ** 2 + p_b ** 2)
    else:
        weight = p_b ** 2 / (p_b ** 2 + p_l ** 2) if fix else np.zeros_like(p_b)
    return (f[None] * L[None] * (geom * weight)[:, None]).sum(0)


def sky_plane_quadrature(albedo, sky, unit: bool, grid: int = 2000) -> np.ndarray:
    """Expected radiance off a diffuse plane (normal +y) under a constant sky, in float64:
      light strategy : E_uv[ f (d . n)+ Le w_l / p_l ],  p_l = 1/(4 pi), d = the equal-area map of
                       (u, v) (without the sqrt(2 - r^2) factor on x, y unless `unit`), and
                       w_l = p_l^2 / (p_l^2 + p_b(d/|d|)^2), p_b = cos / pi;
      BSDF strategy  : int_H f cos Le w_b dw,  w_b = p_b^2 / (p_b^2 + p_l^2)  (1-D in theta).
    With `unit` the two weights sum to one per direction and the total is albedo * Le."""
    f = np.asarray(albedo, np.float64) / np.pi
    Le = np.asarray(sky, np.float64)
    pl = 1.0 / (4 * np.pi)
    k = (np.arange(grid) + 0.5) / grid
    tu, tv = np.meshgrid(k, k, indexing="ij")
    u, v = 2 * tu - 1, 2 * tv - 1
    sd = 1 - (np.abs(u) + np.abs(v))
    r = 1 - np.abs(sd)
    phi = np.where(r == 0, 1.0, (np.abs(v) - np.abs(u)) / np.where(r == 0, 1.0, r) + 1) * (np.pi / 4)
    rr = r * np.sqrt(np.maximum(0, 2 - r * r)) if unit else r
    y = np.copysign(rr * np.sin(phi), v)
    x = np.copysign(rr * np.cos(phi), u)
    z = np.copysign(np.maximum(0, 1 - r * r), sd)
    length = np.sqrt(x * x + y * y + z * z)
    cos_d = np.maximum(y, 0)
    pb = np.maximum(y / length, 0) / np.pi
    wl = pl * pl / (pl * pl + pb * pb)
    light = (cos_d * wl).mean() / pl
    th = (np.arange(grid * 4) + 0.5) / (grid * 4) * (np.pi / 2)
    c = np.cos(th)
    pbt = c / np.pi
    wb = pbt * pbt / (pbt * pbt + pl * pl)
    bsdf = (c * wb * 2 * np.pi * np.sin(th)).sum() * (np.pi / 2) / len(th)
    return f * Le * (light + bsdf)


def _tri_vertices(bh) -> dict:
    return {mi: m.triangle_vertices().astype(np.float64) for mi, m in enumerate(bh.scene.meshes)}


def brute_force_sky(bh, sky, rays: np.ndarray, max_depth: int, seed: int, nthreads: int = 8) -> np.ndarray:
    """Radiance along each ray by BSDF sampling only, constant sky `sky`, diffuse surfaces
    (albedo from the scene's materials). A path that leaves the scene after k <= max_depth
    surface vertices collects throughput * sky. Returns (n, 3) float64."""
    g = np.random.default_rng(seed)
    verts = _tri_vertices(bh)
    albedo = np.asarray(bh.scene.materials["albedo"], np.float64)
    desc, keep = bh.descriptor()
    sky = np.asarray(sky, np.float64)
    n = len(rays)
    L = np.zeros((n, 3))
    T = np.ones((n, 3))
    cur = np.asarray(rays, abi.RAY).copy()
    alive = np.arange(n)
    for k in range(max_depth + 1):
        isect, _ = oracle.traverse(ctypes.addressof(desc), len(bh.nodes), cur, np.full(len(cur), np.inf, np.float32),
                                   abi.INTERSECTION, nthreads)
        hit = isect["hit"] != 0
        L[alive[~hit]] += T[alive[~hit]] * sky[None]
        if k == max_depth:
            break
        sel = np.nonzero(hit)[0]
        alive = alive[sel]
        if len(alive) == 0:
            break
        cur = cur[sel]
        pid = isect["primId"][sel]
        mesh = (pid["id2"] >> 32).astype(np.int64)
        tri = (pid["id2"] & 0xFFFFFFFF).astype(np.int64)
        P = np.zeros((len(sel), 3, 3))
        for mi, vv in verts.items():
            m = mesh == mi
            P[m] = vv[tri[m]]
        o = np.stack([cur["orgX"], cur["orgY"], cur["orgZ"]], 1).astype(np.float64)
        d = np.stack([cur["dirX"], cur["dirY"], cur["dirZ"]], 1).astype(np.float64)
        t = isect["tHit"][sel].astype(np.float64)
        x = o + t[:, None] * d
        nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        nrm *= -np.sign(np.einsum("ij,ij->i", nrm, d))[:, None]  # face the incoming ray
        # cosine-weighted direction about nrm: f cos / pdf = albedo
        r1, r2 = g.random(len(sel)), g.random(len(sel))
        r = np.sqrt(r1)
        phi = 2 * np.pi * r2
        lx, ly, lz = r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0, 1 - r1))
        a = np.where(np.abs(nrm[:, :1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
        tx = np.cross(a, nrm)
        tx /= np.linalg.norm(tx, axis=1, keepdims=True)
        ty = np.cross(nrm, tx)
        wi = lx[:, None] * tx + ly[:, None] * ty + lz[:, None] * nrm
        mat = pid["materialIndex"].astype(np.int64)
        T[alive] *= albedo[mat]
        org = x + 1e-4 * nrm
        cur = np.zeros(len(sel), abi.RAY)
        cur["orgX"], cur["orgY"], cur["orgZ"] = org[:, 0], org[:, 1], org[:, 2]
        cur["dirX"], cur["dirY"], cur["dirZ"] = wi[:, 0], wi[:, 1], wi[:, 2]
    return L


# ---- GPU-against-oracle comparison of whole path records and films ---------------------------

def to_device(paths):
    import torch
    return torch.from_numpy(np.ascontiguousarray(paths).view(np.uint8).copy()).cuda()


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().view(abi.GZ_PATH)


def oracle_params(om: OracleMis, opt) -> oracle.MisParams:
    return om.params(opt.max_depth, fix=opt.fix_light_pdf, fix_sky=opt.fix_sky_direction, seed=opt.zsobol_seed,
                     log2spp=opt.log2_spp, digits=opt.base4_digits, rr_seed=opt.rr_seed)


def agree_mask(got: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """Per path: same bounce count, dimensions and flags, estimate and throughput within 1e-4
    relative (+1e-6 absolute)."""
    e_g, e_r = got["estimate"].astype(np.float64), ref["estimate"].astype(np.float64)
    t_g, t_r = got["throughput"].astype(np.float64), ref["throughput"].astype(np.float64)
    close = (np.abs(e_g - e_r) <= 1e-4 * np.abs(e_r) + 1e-6).all(1) & \
            (np.abs(t_g - t_r) <= 1e-4 * np.abs(t_r) + 1e-6).all(1)
    same = (got["bounce"] == ref["bounce"]) & (got["dim"] == ref["dim"]) & \
           (got["lightDim"] == ref["lightDim"]) & (got["flags"] == ref["flags"])
    return close & same


def check_paths(got: np.ndarray, ref: np.ndarray, min_frac: float = 0.99) -> float:
    """The bar of the MIS parity tests: at least `min_frac` of the paths agree (agree_mask), the
    mean estimate within 2e-3 relative, every estimate finite and no path left active."""
    ok = agree_mask(got, ref)
    print(f"paths agreeing: {ok.mean():.5f} of {len(ok)}", flush=True)
    assert ok.mean() >= min_frac
    mg = got["estimate"].astype(np.float64).mean(0)
    mr = ref["estimate"].astype(np.float64).mean(0)
    assert np.allclose(mg, mr, rtol=2e-3, atol=1e-7), (mg, mr)
    assert np.all(np.isfinite(got["estimate"]))
    assert not np.any(got["flags"] & abi.GZ_PATH_ACTIVE)
    return float(ok.mean())


def run_gpu_and_oracle(bh, paths, opt, sky=None, materials=None, lights=None, launch=None, integ=None):
    """Traces `paths` (GzPath records) on the device with the MIS integrator and on the CPU with
    the oracle, from the same starting records. Returns (device records, oracle records)."""
    from gonzales_amd.integrator import VolumePathIntegrator
    if integ is None:
        bh.upload_to_gpu()
        integ = VolumePathIntegrator(bh, opt, sky=sky, materials=materials, lights=lights)
    if launch is not None:
        integ.set_launch(*launch)
    d = to_device(paths)
    integ.trace(d)
    got = to_host(d)
    bh.check_traversal()
    ref = paths.copy()
    om = OracleMis(bh, sky=sky, materials=integ.materials, lights=integ.lights)
    om.trace(ref, oracle_params(om, opt))
    return got, ref


def oracle_render(bh, camera, opt, materials, lights, sky=None, start=None, nthreads: int = 8) -> np.ndarray:
    """The oracle chain of a whole frame or crop: camera paths (oracle.camera_paths, unless the
    starting records `start` are given) -> every MIS bounce (oracle.mis_trace) -> film sums
    (oracle.film_add) -> resolve. Returns (res_y, res_x, 3) float32 (pixels outside the crop 0)."""
    c = camera.c
    n = camera.path_count
    table = oracle.sobol_matrices()
    paths = oracle.camera_paths(c, 0, n, abi.GZ_PATH, table) if start is None else np.array(start, copy=True)
    om = OracleMis(bh, sky=sky, materials=materials, lights=lights)
    om.trace(paths, oracle_params(om, opt), nthreads)
    px = c.resX * c.resY
    light, albedo, normal = (np.zeros(3 * px, np.float32) for _ in range(3))
    weight = np.zeros(px, np.float32)
    f = camera.film
    oracle.film_add(c, paths, f.iso, f.max_component_value, light, weight, albedo, normal)
    img, bad = oracle.film_resolve(light, weight)
    assert bad == 0
    return img.reshape(c.resY, c.resX, 3)


def check_image(img: np.ndarray, ref: np.ndarray, crop=None, rel: float = 1e-3, min_frac: float = 0.97) -> float:
    """Image bar against the oracle chain: at least `min_frac` of the crop's pixels within `rel`
    relative (+1e-5 absolute) in every channel, the crop's mean within 2e-3 relative."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    if crop is not None:
        x0, x1, y0, y1 = crop
        img, ref = img[y0:y1, x0:x1], ref[y0:y1, x0:x1]
    assert np.all(np.isfinite(img))
    close = (np.abs(img - ref) <= rel * np.abs(ref) + 1e-5).all(-1)
    print(f"pixels agreeing: {close.mean():.4f} of {close.size}", flush=True)
    assert close.mean() >= min_frac
    mi, mr = img.reshape(-1, 3).mean(0), ref.reshape(-1, 3).mean(0)
    assert np.allclose(mi, mr, rtol=2e-3, atol=1e-7), (mi, mr)
    return float(close.mean())


def conductor_strips_under_light(n: int = 8, roughness=None, radiance=(6.0, 5.0, 4.0)):
    """n conductor strips (as conductor_strips_scene) lit by one triangle area light above them,
    facing down, and no sky. Every strip point sees the whole light and nothing else: the light
    strategy's directions do not depend on the roughness, and the BSDF strategy's rays toward the
    light carry MIS weight 0 (the reference's triangle light pdf of a BSDF direction is 0), so the
    image is a smooth function of the roughness with no visibility edge, and its derivative goes
    through the BSDF value and the power heuristic's BSDF pdf. Returns (hierarchy, GzMaterial)."""
    from gonzales_amd.integrator import conductor_material
    rough = np.linspace(0.1, 0.5, n) if roughness is None else np.asarray(roughness)
    verts, idx = [], []
    for i in range(n):
        x0, x1 = -2 + 4 * i / n, -2 + 4 * (i + 1) / n
        v, q = quad([x0, 0, -2], [x0, 0, 2], [x1, 0, 2], [x1, 0, -2])
        idx.append(q + 4 * i)
        verts.append(v)
    tri = np.array([[-1.5, 2.0, -1.0], [1.5, 2.0, -1.2], [0.0, 2.0, 1.5]], np.float32)
    nrm = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    if nrm[1] > 0:
        tri = tri[[0, 2, 1]]
    meshes = [TriangleMesh.from_xyz(np.concatenate(verts), np.concatenate(idx)),
              TriangleMesh.from_xyz(tri, np.arange(3, dtype=np.int64))]
    prims = Primitives.concat([Primitives.geometric(0, 2 * n, np.repeat(np.arange(n), 2)),
                               Primitives.area_lights(1, 1, 0)])
    gz = abi.concat([conductor_material(roughness=float(r)) for r in rough], abi.GZ_MATERIAL)
    scene = Scene(meshes, _materials([(0.5, 0.5, 0.5)] * n), light_emission=np.array([radiance], np.float32))
    return build_bounding_hierarchy(scene, prims), gz
