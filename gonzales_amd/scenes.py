"""Seeded synthetic workloads (triangle scenes and ray batches).

* :func:`triangle_soup` -- N triangles, centroids uniform in [0,1]^3, edge lengths log-normal
  (median 2e-3, sigma 0.5), uniformly random orientations (the 1M-triangle traversal workload).
* :func:`incoherent_rays` -- origins uniform in the unit cube, directions uniform on the sphere,
  tMax uniform in [0.1, 2].
* :func:`coherent_rays` -- one pinhole primary ray per pixel centre of a square image looking
  into the unit cube (eye (0.5, 0.5, -0.5), vertical fov 60 degrees), tMax uniform in [0.1, 2].
* :func:`cornell_box` -- the Cornell-box geometry of the reference's Scenes/cornell-box.pbrt
  (5 wall quads, two boxes rotated about y, one emissive ceiling quad, radiance (17, 12, 4)).

All generators use numpy's PCG64 bit generator with fixed seeds, so every run, on any machine,
sees the same inputs.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .accelerator import Primitives, Scene, TriangleMesh, build_bounding_hierarchy


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def triangle_soup(n: int = 1 << 20, seed: int = 1) -> TriangleMesh:
    """Triangle soup: centroid c ~ U[0,1]^3, edge scale e ~ LogNormal(ln 2e-3, 0.5), vertices
    c + e R t_k with t_k the unit-edge equilateral triangle centred at 0 and R a uniform random
    rotation (normalised Gaussian quaternion)."""
    g = _rng(seed)
    c = g.random((n, 3))
    e = np.exp(np.log(2e-3) + 0.5 * g.standard_normal(n))
    q = g.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    s3 = np.sqrt(3.0)
    tmpl = np.array([[1 / s3, 0.0, 0.0], [-1 / (2 * s3), 0.5, 0.0], [-1 / (2 * s3), -0.5, 0.0]])
    verts = c[:, None, :] + e[:, None, None] * np.einsum("nij,kj->nki", R, tmpl)
    pts = np.ones((3 * n, 4), np.float32)
    pts[:, :3] = verts.reshape(-1, 3).astype(np.float32)
    return TriangleMesh(pts, np.arange(3 * n, dtype=np.int64))


def _tmax(n: int, seed: int) -> np.ndarray:
    return _rng(seed).uniform(0.1, 2.0, n).astype(np.float32)


def incoherent_rays(n: int = 1 << 24, seed_org: int = 2, seed_dir: int = 3, seed_tmax: int = 4):
    """Origins U[0,1]^3, directions uniform on S^2 (z = 1 - 2u, phi = 2 pi v), tMax U[0.1, 2]."""
    rays = np.empty(n, abi.RAY)
    o = _rng(seed_org).random((n, 3)).astype(np.float32)
    uv = _rng(seed_dir).random((n, 2))
    z = 1.0 - 2.0 * uv[:, 0]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * uv[:, 1]
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)
    rays["orgX"], rays["orgY"], rays["orgZ"] = o[:, 0], o[:, 1], o[:, 2]
    rays["dirX"], rays["dirY"], rays["dirZ"] = d[:, 0], d[:, 1], d[:, 2]
    return rays, _tmax(n, seed_tmax)


def coherent_rays(res: int = 4096, seed_tmax: int = 4):
    """Pinhole primary rays through pixel centres: eye (0.5, 0.5, -0.5) looking at (0.5, 0.5, 0.5),
    up +y, vertical fov 60 degrees, square image; directions normalised."""
    n = res * res
    j, i = np.meshgrid(np.arange(res), np.arange(res), indexing="ij")
    tan = np.tan(np.radians(30.0))
    sx = (2.0 * (i + 0.5) / res - 1.0) * tan
    sy = (1.0 - 2.0 * (j + 0.5) / res) * tan
    d = np.stack([sx, sy, np.ones_like(sx)], axis=-1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    rays = np.empty(n, abi.RAY)
    rays["orgX"], rays["orgY"], rays["orgZ"] = 0.5, 0.5, -0.5
    rays["dirX"], rays["dirY"], rays["dirZ"] = d[:, 0], d[:, 1], d[:, 2]
    return rays, _tmax(n, seed_tmax)


def soup_scene(n: int = 1 << 20, seed: int = 1):
    """Single-mesh triangle soup as geometric primitives with one diffuse material."""
    mesh = triangle_soup(n, seed)
    mats = np.zeros(1, abi.MATERIAL)
    mats["type"] = 1
    mats["albedo"] = 0.5
    scene = Scene([mesh], mats)
    prims = Primitives.geometric(0, mesh.number_triangles, 0)
    return scene, prims


def soup_hierarchy(n: int = 1 << 20, seed: int = 1):
    scene, prims = soup_scene(n, seed)
    return build_bounding_hierarchy(scene, prims)


# ---------------------------------------------------------------------------------------------
# Cornell box (Scenes/cornell-box.pbrt:49-162)
# ---------------------------------------------------------------------------------------------

def _quad(p0, p1, p2, p3):
    return [p0, p1, p2, p3], [0, 1, 2, 0, 2, 3]


def _box(cx, cz, half_w, height, angle_deg, with_floor=True):
    """Axis box of half width `half_w`, rotated about y by angle, standing on y = 0."""
    a = np.radians(angle_deg)
    ca, sa = np.cos(a), np.sin(a)

    def P(x, y, z):
        return [cx + ca * x + sa * z, y, cz - sa * x + ca * z]

    h = half_w
    v = [P(-h, 0, -h), P(h, 0, -h), P(h, 0, h), P(-h, 0, h),
         P(-h, height, -h), P(h, height, -h), P(h, height, h), P(-h, height, h)]
    faces = [(4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    if with_floor:
        faces.append((0, 3, 2, 1))
    idx = []
    for f in faces:
        idx += [f[0], f[1], f[2], f[0], f[2], f[3]]
    return v, idx


def cornell_box():
    """34-triangle Cornell box: walls x, z in [-1, 1], y in [0, 2]; light quad at y = 1.98 with
    radiance (17, 12, 4); a tall box rotated +18 degrees and a short box rotated -18 degrees
    (the short box's floor face is omitted: it is never visible and 36 - 2 = 34 triangles).
    Materials: white (0.725, 0.71, 0.68), red (0.63, 0.065, 0.05) and green (0.14, 0.45, 0.091)
    walls clamped into [0.1, 0.8]. Returns (scene, primitives, camera dict)."""
    meshes = []
    parts = []
    mats = []

    def add(verts, idx, material):
        mi = len(meshes)
        meshes.append(TriangleMesh.from_xyz(np.asarray(verts, np.float32), np.asarray(idx, np.int64)))
        parts.append(Primitives.geometric(mi, len(idx) // 3, material))

    white = (0.725, 0.71, 0.68)
    red = tuple(np.clip((0.63, 0.065, 0.05), 0.1, 0.8))
    green = tuple(np.clip((0.14, 0.45, 0.091), 0.1, 0.8))
    for c in (white, red, green):
        m = np.zeros(1, abi.MATERIAL)
        m["type"] = 1
        m["albedo"] = c
        mats.append(m)
    add(*_quad([-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]), 0)   # floor
    add(*_quad([-1, 2, -1], [-1, 2, 1], [1, 2, 1], [1, 2, -1]), 0)   # ceiling
    add(*_quad([-1, 0, -1], [-1, 2, -1], [1, 2, -1], [1, 0, -1]), 0)  # back wall
    add(*_quad([-1, 0, -1], [-1, 0, 1], [-1, 2, 1], [-1, 2, -1]), 1)  # left (red)
    add(*_quad([1, 0, -1], [1, 2, -1], [1, 2, 1], [1, 0, 1]), 2)      # right (green)
    add(*_box(-0.335, -0.29, 0.3, 1.2, 18.0, with_floor=True), 0)   # tall box
    add(*_box(0.33, 0.37, 0.3, 0.6, -18.0, with_floor=False), 0)    # short box
    # Light quad (two triangles) at y = 1.98, facing down.
    li = len(meshes)
    lv, lidx = _quad([-0.24, 1.98, -0.22], [0.23, 1.98, -0.22], [0.23, 1.98, 0.16], [-0.24, 1.98, 0.16])
    meshes.append(TriangleMesh.from_xyz(np.asarray(lv, np.float32), np.asarray(lidx, np.int64)))
    parts.append(Primitives.area_lights(li, 2, 0))
    scene = Scene(meshes, abi.concat(mats, abi.MATERIAL),
                  light_emission=np.array([[17, 12, 4], [17, 12, 4]], np.float32))
    camera = {"eye": (0.0, 1.0, 6.8), "look": (0.0, 1.0, 0.0), "up": (0.0, 1.0, 0.0), "fov": 19.5}
    return scene, Primitives.concat(parts), camera


def cornell_hierarchy():
    scene, prims, camera = cornell_box()
    return build_bounding_hierarchy(scene, prims), camera


def total_triangles(scene: Scene) -> int:
    return int(sum(m.number_triangles for m in scene.meshes))


# ---------------------------------------------------------------------------------------------
# Displaced icosphere (the full path-trace workload)
# ---------------------------------------------------------------------------------------------

def icosphere(subdivisions: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: the icosahedron with every triangle split into 4, `subdivisions` times
    (20 * 4^s faces, 10 * 4^s + 2 vertices), midpoints pushed to the sphere. Float64 vertices,
    int64 (F, 3) faces, counter-clockwise seen from outside."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(subdivisions):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(e, axis=1)
        uniq, inv = np.unique(key[:, 0] * len(v) + key[:, 1], return_inverse=True)
        a, b = uniq // len(v), uniq % len(v)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = (len(v) + inv).reshape(3, -1)  # midpoint ids of edges (0,1), (1,2), (2,0) per face
        v = np.concatenate([v, mid])
        m01, m12, m20 = m
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([m01, f[:, 1], m12], 1),
                            np.stack([m20, m12, f[:, 2]], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _value_noise(p: np.ndarray, seed: int, table: int = 256) -> np.ndarray:
    """3-D value noise in [-1, 1]: lattice values from a seeded permutation table, smoothstep
    trilinear interpolation."""
    g = _rng(seed)
    perm = g.permutation(table)
    vals = g.uniform(-1, 1, table)
    i = np.floor(p).astype(np.int64)
    fr = p - i
    s = fr * fr * (3 - 2 * fr)

    def lat(dx, dy, dz):
        h = perm[(perm[(perm[(i[:, 0] + dx) % table] + i[:, 1] + dy) % table] + i[:, 2] + dz) % table]
        return vals[h]

    x0 = lat(0, 0, 0) * (1 - s[:, 0]) + lat(1, 0, 0) * s[:, 0]
    x1 = lat(0, 1, 0) * (1 - s[:, 0]) + lat(1, 1, 0) * s[:, 0]
    x2 = lat(0, 0, 1) * (1 - s[:, 0]) + lat(1, 0, 1) * s[:, 0]
    x3 = lat(0, 1, 1) * (1 - s[:, 0]) + lat(1, 1, 1) * s[:, 0]
    y0 = x0 * (1 - s[:, 1]) + x1 * s[:, 1]
    y1 = x2 * (1 - s[:, 1]) + x3 * s[:, 1]
    return y0 * (1 - s[:, 2]) + y1 * s[:, 2]


def fbm(p: np.ndarray, seed: int, octaves: int = 4, lacunarity: float = 2.0, gain: float = 0.5,
        frequency: float = 4.0) -> np.ndarray:
    out = np.zeros(len(p))
    amp, freq = 1.0, frequency
    for _ in range(octaves):
        out += amp * _value_noise(p * freq, seed)
        amp *= gain
        freq *= lacunarity
    return out


def inverse_rendering_scene(subdivisions: int = 6, n_params: int = 32, seed: int = 8, patch: int = 64):
    """The Cornell box plus three displaced occluders: two icospheres (2 x 20 x 4^s triangles) and a
    lat-long sphere of about 36,160 x 4^(s-6) triangles, 200,320 in all at s = 6, whose 64-face
    patches cycle through n_params diffuse materials (parameter: RGB
    reflectance) and n_params conductor materials (parameter: GGX roughness). Material table:
    the 3 walls (white, red, green), then the diffuse, then the conductor materials. Returns
    (hierarchy, GzMaterial table, target reflectances (n_params, 3), target roughness (n_params,)),
    targets drawn with `seed`: reflectance U[0.1, 0.9], roughness U[0.05, 0.5]."""
    from .integrator import conductor_material, diffuse_material
    scene, prims, _ = cornell_box()
    meshes = list(scene.meshes)
    parts = [prims]
    for k, (c, r) in enumerate((((-0.42, 0.45, -0.25), 0.36), ((0.40, 0.36, 0.30), 0.3), ((0.05, 1.3, -0.35), 0.2))):
        v, f = icosphere(subdivisions) if k < 2 else latlong_sphere(int(round(36160 * 4.0 ** (subdivisions - 6))))
        v = v * (1 + 0.03 * fbm(v, 5 + k))[:, None]
        v = v * r + np.asarray(c)
        mi = len(meshes)
        meshes.append(TriangleMesh.from_xyz(v.astype(np.float32), f.reshape(-1)))
        patch_id = np.arange(len(f)) // patch + k * 7
        parts.append(Primitives.geometric(mi, len(f), 3 + (patch_id % (2 * n_params))))
    g = _rng(seed)
    target_albedo = g.uniform(0.1, 0.9, (n_params, 3)).astype(np.float32)
    target_rough = g.uniform(0.05, 0.5, n_params).astype(np.float32)
    walls = [diffuse_material(tuple(scene.materials["albedo"][i])) for i in range(3)]
    mats = walls + [diffuse_material(a) for a in target_albedo] + [conductor_material(roughness=r)
                                                                    for r in target_rough]
    gz = abi.concat(mats, abi.GZ_MATERIAL)
    placeholder = np.zeros(len(gz), abi.MATERIAL)
    placeholder["type"] = 1
    placeholder["albedo"] = 0.5
    new_scene = Scene(meshes, placeholder, light_emission=scene.light_emission)
    return build_bounding_hierarchy(new_scene, Primitives.concat(parts)), gz, target_albedo, target_rough


def displaced_sphere_scene(subdivisions: int = 9, amplitude: float = 0.05, seed_noise: int = 5,
                           seed_mat: int = 6, patch: int = 64):
    """Icosphere (r = 1, `subdivisions` levels) displaced along its normal by amplitude x fBm
    (value noise, 4 octaves, lacunarity 2, gain 0.5), on a 2-triangle ground plane at y = -1.05,
    with one quad area light (radiance 8) above. One material per `patch` consecutive faces,
    drawn with seed_mat: diffuse (albedo U[0.1, 0.8]^3), dielectric (eta 1.5) or coated
    conductor (Cu), GGX roughness U[0.05, 0.6]. Returns (hierarchy, GzMaterial table)."""
    from .integrator import coated_conductor_material, dielectric_material, diffuse_material
    v, f = icosphere(subdivisions)
    v = v * (1 + amplitude * fbm(v, seed_noise))[:, None]
    sphere = TriangleMesh.from_xyz(v.astype(np.float32), f.reshape(-1))
    half = 20.0
    gv, gi = _quad([-half, -1.05, -half], [-half, -1.05, half], [half, -1.05, half], [half, -1.05, -half])
    ground = TriangleMesh.from_xyz(np.asarray(gv, np.float32), np.asarray(gi, np.int64))
    lv, li = _quad([-0.75, 2.5, -0.75], [0.75, 2.5, -0.75], [0.75, 2.5, 0.75], [-0.75, 2.5, 0.75])  # faces down
    light = TriangleMesh.from_xyz(np.asarray(lv, np.float32), np.asarray(li, np.int64))
    n_faces = len(f)
    n_patch = (n_faces + patch - 1) // patch
    g = _rng(seed_mat)
    kind = g.integers(0, 3, n_patch)
    albedo = g.uniform(0.1, 0.8, (n_patch, 3))
    rough = g.uniform(0.05, 0.6, n_patch)
    mats = [diffuse_material((0.5, 0.5, 0.5))]  # ground
    for k in range(n_patch):
        if kind[k] == 0:
            mats.append(diffuse_material(albedo[k]))
        elif kind[k] == 1:
            mats.append(dielectric_material(eta=1.5, roughness=rough[k]))
        else:
            mats.append(coated_conductor_material(interface_roughness=rough[k], conductor_roughness=rough[k]))
    gz = abi.concat(mats, abi.GZ_MATERIAL)
    face_mat = 1 + np.arange(n_faces) // patch
    prims = Primitives.concat([Primitives.geometric(0, n_faces, face_mat), Primitives.geometric(1, 2, 0),
                               Primitives.area_lights(2, 2, 0)])
    placeholder = np.zeros(len(gz), abi.MATERIAL)
    placeholder["type"] = 1
    placeholder["albedo"] = 0.5
    scene = Scene([sphere, ground, light], placeholder,
                  light_emission=np.array([[8.0, 8.0, 8.0], [8.0, 8.0, 8.0]], np.float32))
    return build_bounding_hierarchy(scene, prims), gz


# ---------------------------------------------------------------------------------------------
# Many-scene batch: 64 scenes of uneven size
# ---------------------------------------------------------------------------------------------

def latlong_sphere(target_tris: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit sphere of m - 1 rings of n = 2m vertices between the poles (two pole fans and m - 2
    bands of quads): 4 m (m - 1) triangles, m chosen to bring that count nearest `target_tris`.
    Float64 vertices, int64 (F, 3) faces, counter-clockwise seen from outside."""
    m = max(3, int(round(0.5 + np.sqrt(0.25 + target_tris / 4.0))))
    n = 2 * m
    theta = np.pi * np.arange(1, m) / m
    phi = 2 * np.pi * np.arange(n) / n
    st, ct = np.sin(theta), np.cos(theta)
    ring = np.stack([st[:, None] * np.cos(phi)[None, :], np.broadcast_to(ct[:, None], (m - 1, n)),
                     st[:, None] * np.sin(phi)[None, :]], -1).reshape(-1, 3)
    v = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]])
    south = len(v) - 1
    j = np.arange(n)
    jn = (j + 1) % n
    faces = [np.stack([np.zeros(n, np.int64), 1 + jn, 1 + j], 1)]
    for r in range(m - 2):
        a, b = 1 + r * n + j, 1 + r * n + jn
        c, d = 1 + (r + 1) * n + j, 1 + (r + 1) * n + jn
        faces += [np.stack([a, b, d], 1), np.stack([a, d, c], 1)]
    last = 1 + (m - 2) * n
    faces.append(np.stack([np.full(n, south), last + j, last + jn], 1))
    return v, np.concatenate(faces).astype(np.int64)


def batch_scene_specs(count: int = 64, seed: int = 7) -> list[dict]:
    """Parameters of `count` scenes: triangle count log-uniform in [1e4, 2e6], kind alternating
    displaced sphere / triangle soup, resolution in {256, 512, 1024} squared and spp in {4, 16, 64}
    uniformly, max depth 6; per-scene seeds for geometry and materials."""
    g = _rng(seed)
    tris = np.exp(g.uniform(np.log(1e4), np.log(2e6), count)).astype(np.int64)
    res = g.choice([256, 512, 1024], count)
    spp = g.choice([4, 16, 64], count)
    return [{"index": k, "kind": "sphere" if k % 2 == 0 else "soup", "tris": int(tris[k]), "res": int(res[k]),
             "spp": int(spp[k]), "max_depth": 6, "seed": int(seed * 1000 + k)} for k in range(count)]


def batch_scene(spec: dict, patch: int = 64):
    """One batch scene: a displaced lat-long sphere (r = 1, fBm amplitude 0.05) on a ground plane,
    or a triangle soup in the unit cube (the traversal benchmark's generator), with 64-face
    patches of diffuse (albedo U[0.1, 0.8]^3), dielectric (eta 1.5, GGX roughness U[0.05, 0.6])
    and conductor (Cu, GGX roughness U[0.05, 0.6]) materials, a quad area light above and a
    constant sky. Returns (hierarchy, GzMaterial table, camera look-at (eye, target, up, fov))."""
    from .integrator import conductor_material, dielectric_material, diffuse_material
    g = _rng(spec["seed"])
    meshes, parts = [], []
    if spec["kind"] == "sphere":
        v, f = latlong_sphere(spec["tris"])
        v = v * (1 + 0.05 * fbm(v, spec["seed"]))[:, None]
        meshes.append(TriangleMesh.from_xyz(v.astype(np.float32), f.reshape(-1)))
        n_faces = len(f)
        gv, gi = _quad([-20, -1.05, -20], [-20, -1.05, 20], [20, -1.05, 20], [20, -1.05, -20])
        meshes.append(TriangleMesh.from_xyz(np.asarray(gv, np.float32), np.asarray(gi, np.int64)))
        lv, li = _quad([-0.75, 2.5, -0.75], [0.75, 2.5, -0.75], [0.75, 2.5, 0.75], [-0.75, 2.5, 0.75])
        cam = ((0.0, 0.45, 3.4), (0.0, -0.15, 0.0), (0.0, 1.0, 0.0), 40.0)
    else:
        soup = triangle_soup(spec["tris"], spec["seed"])
        meshes.append(soup)
        n_faces = soup.number_triangles
        lv, li = _quad([0.1, 1.8, 0.1], [0.9, 1.8, 0.1], [0.9, 1.8, 0.9], [0.1, 1.8, 0.9])
        cam = ((0.5, 0.7, -1.4), (0.5, 0.45, 0.5), (0.0, 1.0, 0.0), 45.0)
    n_patch = (n_faces + patch - 1) // patch
    kind = g.integers(0, 3, n_patch)
    albedo = g.uniform(0.1, 0.8, (n_patch, 3))
    rough = g.uniform(0.05, 0.6, n_patch)
    mats = [diffuse_material((0.5, 0.5, 0.5))]
    for k in range(n_patch):
        if kind[k] == 0:
            mats.append(diffuse_material(albedo[k]))
        elif kind[k] == 1:
            mats.append(dielectric_material(eta=1.5, roughness=rough[k]))
        else:
            mats.append(conductor_material(roughness=rough[k]))
    gz = abi.concat(mats, abi.GZ_MATERIAL)
    parts.append(Primitives.geometric(0, n_faces, 1 + np.arange(n_faces) // patch))
    if spec["kind"] == "sphere":
        parts.append(Primitives.geometric(1, 2, 0))
    light_mesh = len(meshes)
    meshes.append(TriangleMesh.from_xyz(np.asarray(lv, np.float32), np.asarray(li, np.int64)))
    parts.append(Primitives.area_lights(light_mesh, 2, 0))
    placeholder = np.zeros(len(gz), abi.MATERIAL)
    placeholder["type"] = 1
    placeholder["albedo"] = 0.5
    scene = Scene(meshes, placeholder, light_emission=np.array([[8.0, 8.0, 8.0], [8.0, 8.0, 8.0]], np.float32))
    return build_bounding_hierarchy(scene, Primitives.concat(parts)), gz, cam
