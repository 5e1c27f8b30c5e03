"""ORACLE — test infrastructure only.

ctypes front end of ``oracle/liboracle.so``, the CPU restatement of the reference's hot path
(Sources/mojoKernel/kernel.mojo traversal + shade, Sampler/ZSobolSampler.swift,
Integrator/VolumePathIntegrator.swift MIS bounce). Imported only by
tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg, always as the checker; the
product (``gonzales_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
JOE_KUO = HERE.parent / "gonzales_amd" / "data" / "new-joe-kuo-6.1023"

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32

_SIGS = {
    "orc_intersect_triangle": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_float, _P, _P, _P]),
    "orc_intersect_aabb": (ctypes.c_int, [_P, _P, _P, ctypes.c_float, _P]),
    "orc_traverse_one": (ctypes.c_int, [_P, _I64, _P, ctypes.c_float, _P, _P]),
    "orc_traverse_batch": (ctypes.c_int, [_P, _I64, _P, _P, _I64, _P, _P, ctypes.c_int]),
    "orc_occluded_batch": (ctypes.c_int, [_P, _I64, _P, _P, _I64, _P, _P, ctypes.c_int]),
    "orc_brute_closest": (None, [_P, _I64, _P, _P, _I64, _P, _P]),
    "orc_reverse_bits32": (ctypes.c_uint32, [ctypes.c_uint32]),
    "orc_owen": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
    "orc_encode_morton2": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32]),
    "orc_mix_bits": (ctypes.c_uint32, [ctypes.c_uint64]),
    "orc_next_pow2": (_I64, [_I64]),
    "orc_sobol_matrices": (ctypes.c_int, [ctypes.c_char_p, _P]),
    "orc_sobol_sample": (ctypes.c_float, [_I64, ctypes.c_int, ctypes.c_uint32, _P]),
    "orc_zsobol_init": (None, [_P, _I64, _I64, _I64, _I64]),
    "orc_zsobol_start": (None, [_P, _I32, _I32, _I64, _I32]),
    "orc_zsobol_sample_index": (ctypes.c_uint64, [_P]),
    "orc_zsobol_get1d": (ctypes.c_float, [_P, _P]),
    "orc_zsobol_get2d": (None, [_P, _P, _P]),
    "orc_zsobol_get3d": (None, [_P, _P, _P]),
    "orc_zsobol_batch": (None, [_I64, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _I32, _P, _I32, _P]),
    "orc_pcg32": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _P]),
    "orc_shade_batch": (None, [_P, _I64, _P, _P, _P]),
    "orc_mis_trace_paths": (ctypes.c_int, [_P, _I64, _P, _I32, _P, _I32, _P, _P, _P, _I64, ctypes.c_int]),
    "orc_power_heuristic": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_concentric_disk": (None, [ctypes.c_float, ctypes.c_float, _P]),
    "orc_cosine_hemisphere": (None, [ctypes.c_float, ctypes.c_float, _P]),
    "orc_equal_area_square_to_sphere": (None, [ctypes.c_float, ctypes.c_float, ctypes.c_int, _P]),
    "orc_rr_draw": (ctypes.c_float, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32]),
    "orc_fresnel_dielectric": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_fresnel_conductor": (None, [ctypes.c_float, _P, _P, _P]),
    "orc_refract": (ctypes.c_int, [_P, _P, ctypes.c_float, _P]),
    "orc_tr_D": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, _P]),
    "orc_tr_G1": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, _P]),
    "orc_tr_pdf": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, _P, _P]),
    "orc_tr_sample_h": (None, [ctypes.c_float, ctypes.c_float, _P, ctypes.c_float, ctypes.c_float, _P]),
    "orc_bx_f": (None, [_P, _P, _P, ctypes.c_uint64, _P]),
    "orc_bx_pdf": (ctypes.c_float, [_P, _P, _P, ctypes.c_uint64]),
    "orc_bx_sample": (None, [_P, _P, _P, ctypes.c_uint64, _P]),
    "orc_erfinv": (ctypes.c_float, [ctypes.c_float]),
    "orc_gaussian_evaluate": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "orc_gaussian_sample": (None, [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, _P]),
    "orc_camera_paths": (ctypes.c_int, [_P, _P, _I64, _I64, _P]),
    "orc_film_add": (ctypes.c_int, [_P, _P, _I64, _I64, ctypes.c_float, ctypes.c_float, _P, _P, _P, _P]),
    "orc_film_resolve": (_I64, [_P, _P, _I64, _P]),
}

_lib = None
_path = LIB


def use_native() -> str:
    """Switches to a build of the same sources with -march=native, compiled now on this machine
    (make native; the CPU baseline's build). Returns the compiler flags in effect: the native
    build's, or the portable library's if gcc cannot build it here."""
    global _lib, _path
    import subprocess
    native = HERE / "_native" / "liboracle.so"
    proc = subprocess.run(["make", "-s", "-C", str(HERE), "native"], capture_output=True, text=True)
    if proc.returncode != 0 or not native.exists():
        return "-O3 -ffp-contract=off (portable; the -march=native build failed)"
    _path, _lib = native, None
    return "-O3 -march=native -ffp-contract=off"


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _path.exists():
            raise RuntimeError(f"{_path} is not built (make -C oracle)")
        l = ctypes.CDLL(str(_path))
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


STATS = np.dtype({"names": ["interior", "tris"], "formats": ["<u4", "<u4"], "offsets": [0, 4], "itemsize": 8})


class ZSobolState(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("log2spp", ctypes.c_int32), ("nBase4Digits", ctypes.c_int32),
                ("morton", ctypes.c_uint64), ("dim", ctypes.c_int32), ("pad", ctypes.c_int32)]


# ---- traversal -------------------------------------------------------------------------------

def intersect_triangle(org, d, p0, p1, p2, t_max):
    a = [np.ascontiguousarray(x, np.float32) for x in (org, d, p0, p1, p2)]
    t, u, v = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    hit = lib().orc_intersect_triangle(*[_p(x) for x in a], float(t_max), ctypes.byref(t), ctypes.byref(u),
                                       ctypes.byref(v))
    return bool(hit), t.value, u.value, v.value


def intersect_aabb(bounds, org, d, t_max):
    a = [np.ascontiguousarray(x, np.float32) for x in (bounds, org, d)]
    tn = ctypes.c_float()
    hit = lib().orc_intersect_aabb(*[_p(x) for x in a], float(t_max), ctypes.byref(tn))
    return bool(hit), tn.value


def traverse(desc_ptr: int, node_count: int, rays: np.ndarray, t_max: np.ndarray, isect_dtype, nthreads: int = 1,
             with_stats: bool = False):
    """Closest hit for every ray (mojo_cpu_traverse_batch). Returns (Intersection array, stats|None)."""
    n = len(rays)
    out = np.zeros(n, isect_dtype)
    st = np.zeros(n, STATS) if with_stats else None
    rays = np.ascontiguousarray(rays)
    tm = np.ascontiguousarray(t_max, np.float32)
    rc = lib().orc_traverse_batch(desc_ptr, node_count, _p(rays), _p(tm), n,
                                  _p(out), _p(st) if st is not None else None, nthreads)
    if rc != 0:
        raise RuntimeError("oracle traversal stack overflow")
    return out, st


def occluded(desc_ptr: int, node_count: int, rays: np.ndarray, t_max: np.ndarray, nthreads: int = 1,
             with_stats: bool = False):
    n = len(rays)
    out = np.zeros(n, np.uint8)
    st = np.zeros(n, STATS) if with_stats else None
    rays = np.ascontiguousarray(rays)
    tm = np.ascontiguousarray(t_max, np.float32)
    rc = lib().orc_occluded_batch(desc_ptr, node_count, _p(rays), _p(tm), n,
                                  _p(out), _p(st) if st is not None else None, nthreads)
    if rc != 0:
        raise RuntimeError("oracle traversal stack overflow")
    return out.astype(bool), st


def brute_closest(desc_ptr: int, prim_count: int, rays: np.ndarray, t_max: np.ndarray):
    n = len(rays)
    hi = np.zeros(n, np.int64)
    t = np.zeros(n, np.float32)
    rays = np.ascontiguousarray(rays)
    tm = np.ascontiguousarray(t_max, np.float32)
    lib().orc_brute_closest(desc_ptr, prim_count, _p(rays), _p(tm), n, _p(hi), _p(t))
    return hi, t


# ---- Z-Sobol ---------------------------------------------------------------------------------

def sobol_matrices() -> np.ndarray:
    out = np.zeros(1024 * 52, np.uint32)
    if lib().orc_sobol_matrices(str(JOE_KUO).encode(), _p(out)) != 0:
        raise RuntimeError(f"cannot read {JOE_KUO}")
    return out


def zsobol_batch(spp, res_x, res_y, seed, px, py, sample_index, pattern, start_dim=0, table=None):
    table = sobol_matrices() if table is None else table
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    si = np.ascontiguousarray(sample_index, np.int64)
    pat = np.ascontiguousarray(pattern, np.int8)
    stride = int(pat.sum())
    out = np.zeros((len(px), stride), np.float32)
    lib().orc_zsobol_batch(spp, res_x, res_y, seed, _p(table), _p(px), _p(py), _p(si), len(px), start_dim, _p(pat),
                           len(pat), _p(out))
    return out


# ---- shade (mode A) --------------------------------------------------------------------------

def pcg32(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32(initstate, initseq, n, _p(out))
    return out


def shade(paths: np.ndarray, isects: np.ndarray, meshes_c_addr: int, materials: np.ndarray) -> None:
    """mojo_cpu_shade_batch: updates `paths` (PathState_C records) in place."""
    lib().orc_shade_batch(_p(paths), len(paths), _p(isects), meshes_c_addr, _p(materials))


# ---- MIS path integrator (mode B) ------------------------------------------------------------

class MisParams(ctypes.Structure):
    """mb_params of oracle_mis.c: the device GzMisParams plus the constant sky radiance."""
    _fields_ = [("maxDepth", ctypes.c_int32), ("flags", ctypes.c_int32), ("zsSeed", ctypes.c_int64),
                ("zsLog2Spp", ctypes.c_int32), ("zsBase4Digits", ctypes.c_int32), ("rrSeed", ctypes.c_uint64),
                ("sky", ctypes.c_float * 3), ("hasSky", ctypes.c_int32)]


def mis_trace(desc_ptr: int, node_count: int, materials: np.ndarray, lights: np.ndarray, params: MisParams,
              paths: np.ndarray, nthreads: int = 1, table=None) -> None:
    """Runs every active GzPath record to completion in place (VolumePathIntegrator bounces)."""
    table = sobol_matrices() if table is None else table
    rc = lib().orc_mis_trace_paths(desc_ptr, node_count, _p(materials), len(materials),
                                   _p(lights) if len(lights) else None, len(lights), _p(table),
                                   ctypes.byref(params), _p(paths), len(paths), nthreads)
    if rc != 0:
        raise RuntimeError("oracle MIS: bad arguments")


def power_heuristic(f: float, g: float) -> float:
    return lib().orc_power_heuristic(f, g)


def concentric_disk(u0: float, u1: float) -> np.ndarray:
    out = np.zeros(2, np.float32)
    lib().orc_concentric_disk(u0, u1, _p(out))
    return out


def cosine_hemisphere(u0: float, u1: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_cosine_hemisphere(u0, u1, _p(out))
    return out


def equal_area_square_to_sphere(u: float, v: float, unit: bool = False) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_equal_area_square_to_sphere(u, v, int(unit), _p(out))
    return out


def rr_draw(seed: int, px: int, py: int, si: int, bounce: int) -> float:
    return lib().orc_rr_draw(seed, px, py, si, bounce)


# ---- BSDFs (oracle_bsdf.c) -------------------------------------------------------------------

def _f3(v) -> np.ndarray:
    return np.ascontiguousarray(v, np.float32).reshape(3)


def fresnel_dielectric(cos_i: float, eta: float) -> float:
    return lib().orc_fresnel_dielectric(cos_i, eta)


# (every array handed to C is bound to a local name first, so that it outlives the call)

def fresnel_conductor(cos_t: float, eta, k) -> np.ndarray:
    e, kk, out = _f3(eta), _f3(k), np.zeros(3, np.float32)
    lib().orc_fresnel_conductor(cos_t, _p(e), _p(kk), _p(out))
    return out


def refract(wi, n, eta: float):
    w, nn, out = _f3(wi), _f3(n), np.zeros(4, np.float32)
    ok = lib().orc_refract(_p(w), _p(nn), eta, _p(out))
    return bool(ok), out[:3], float(out[3])


def tr_D(ax: float, ay: float, h) -> float:
    hh = _f3(h)
    return lib().orc_tr_D(ax, ay, _p(hh))


def tr_G1(ax: float, ay: float, w) -> float:
    ww = _f3(w)
    return lib().orc_tr_G1(ax, ay, _p(ww))


def tr_pdf(ax: float, ay: float, wo, h) -> float:
    o, hh = _f3(wo), _f3(h)
    return lib().orc_tr_pdf(ax, ay, _p(o), _p(hh))


def tr_sample_h(ax: float, ay: float, wo, u0: float, u1: float) -> np.ndarray:
    o, out = _f3(wo), np.zeros(3, np.float32)
    lib().orc_tr_sample_h(ax, ay, _p(o), u0, u1, _p(out))
    return out


def bsdf_f(material: np.ndarray, wo, wi, seed: int = 0) -> np.ndarray:
    """Local-frame f of one GzMaterial record (the layered BSDF draws from stream `seed`)."""
    m = np.ascontiguousarray(material).reshape(1)
    o, i, out = _f3(wo), _f3(wi), np.zeros(3, np.float32)
    lib().orc_bx_f(_p(m), _p(o), _p(i), seed, _p(out))
    return out


def bsdf_pdf(material: np.ndarray, wo, wi, seed: int = 0) -> float:
    m = np.ascontiguousarray(material).reshape(1)
    o, i = _f3(wo), _f3(wi)
    return lib().orc_bx_pdf(_p(m), _p(o), _p(i), seed)


def bsdf_sample(material: np.ndarray, wo, u, seed: int = 0):
    """Returns (f, wi, pdf) of the local-frame sample for uniform numbers u[0..2]."""
    m = np.ascontiguousarray(material).reshape(1)
    o, uu, out = _f3(wo), _f3(u), np.zeros(7, np.float32)
    lib().orc_bx_sample(_p(m), _p(o), _p(uu), seed, _p(out))
    return out[:3], out[3:6], float(out[6])


# ---- camera, filter, film (oracle_film.c) ------------------------------------------------------

def erfinv(y: float) -> float:
    return lib().orc_erfinv(y)


def gaussian_evaluate(x: float, y: float, sigma: float) -> float:
    return lib().orc_gaussian_evaluate(x, y, sigma)


def gaussian_sample(u0: float, u1: float, rx: float, ry: float, sigma: float):
    """(dx, dy, pdf) of GaussianFilter.sample."""
    out = np.zeros(3, np.float32)
    lib().orc_gaussian_sample(u0, u1, rx, ry, sigma, _p(out))
    return float(out[0]), float(out[1]), float(out[2])


def camera_paths(camera, first: int, count: int, path_dtype, table=None) -> np.ndarray:
    """GzPath records of paths first..first+count-1 for a GzCamera (ctypes structure)."""
    table = sobol_matrices() if table is None else table
    out = np.zeros(count, path_dtype)
    if lib().orc_camera_paths(ctypes.byref(camera), _p(table), first, count, _p(out)) != 0:
        raise RuntimeError("oracle camera: unsupported camera")
    return out


def film_add(camera, paths: np.ndarray, iso: float, max_component: float, light, weight, albedo, normal) -> None:
    paths = np.ascontiguousarray(paths)
    lib().orc_film_add(ctypes.byref(camera), _p(paths), 0, len(paths), iso, max_component, _p(light), _p(weight),
                       _p(albedo), _p(normal))


def film_resolve(light: np.ndarray, weight: np.ndarray):
    out = np.zeros_like(light)
    bad = lib().orc_film_resolve(_p(light), _p(weight), len(weight), _p(out))
    return out, int(bad)
