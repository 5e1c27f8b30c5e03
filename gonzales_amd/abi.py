"""ctypes binding of ``include/gonzales_amd.h`` (the library ``libgonzales_amd.so``).

numpy structured dtypes mirror the C structs byte for byte so host arrays can be handed to the
C-ABI without copies; torch tensors are passed by ``data_ptr()`` for the ``*_device`` entry
points. Loading fails loudly if the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libgonzales_amd.so"

# Byte layouts of mojo_kernel.h:4-58 (the reference's C structs).
PRIM_ID = np.dtype({"names": ["id1", "id2", "materialIndex", "type"],
                    "formats": ["<i8", "<i8", "<i8", "i1"], "offsets": [0, 8, 16, 24], "itemsize": 32})
MATERIAL = np.dtype({"names": ["type", "albedo", "emission"],
                     "formats": ["i1", ("<f4", 3), ("<f4", 3)], "offsets": [0, 4, 16], "itemsize": 28})
RAY = np.dtype({"names": ["orgX", "orgY", "orgZ", "dirX", "dirY", "dirZ"],
                "formats": ["<f4"] * 6, "offsets": [0, 4, 8, 12, 16, 20], "itemsize": 24})
INTERSECTION = np.dtype({"names": ["primId", "tHit", "u", "v", "hit"],
                         "formats": [PRIM_ID, "<f4", "<f4", "<f4", "i1"],
                         "offsets": [0, 32, 36, 40, 44], "itemsize": 48})
PATH_STATE = np.dtype({"names": ["ray", "throughput", "estimate", "albedo", "pcgState", "pcgInc", "active"],
                       "formats": [RAY, ("<f4", 3), ("<f4", 3), ("<f4", 3), "<u8", "<u8", "i1"],
                       "offsets": [0, 24, 36, 48, 64, 72, 80], "itemsize": 88})
BVH2_NODE = np.dtype({"names": ["boundsMinX", "boundsMinY", "boundsMinZ", "boundsMaxX", "boundsMaxY",
                                "boundsMaxZ", "offset", "count"],
                      "formats": ["<f4"] * 6 + ["<i4", "<i4"],
                      "offsets": [0, 4, 8, 12, 16, 20, 24, 28], "itemsize": 32})
GZ_HIT = np.dtype({"names": ["prim", "t", "u", "v"], "formats": ["<i4", "<f4", "<f4", "<f4"],
                   "offsets": [0, 4, 8, 12], "itemsize": 16})
TRACE_STATS = np.dtype({"names": ["interior", "tris"], "formats": ["<u4", "<u4"], "offsets": [0, 4],
                        "itemsize": 8})
# MIS integrator records (include/gonzales_amd.h, GzPath / GzMaterial / GzAreaLight).
GZ_PATH = np.dtype({"names": ["ray", "throughput", "estimate", "albedo", "firstNormal", "filterWeight", "pixelX",
                              "pixelY", "sampleIndex", "dim", "lightDim", "bounce", "flags"],
                    "formats": [RAY, ("<f4", 3), ("<f4", 3), ("<f4", 3), ("<f4", 3), "<f4", "<i4", "<i4", "<i4",
                                "<i4", "<i4", "<i4", "<i4"],
                    "offsets": [0, 24, 36, 48, 60, 72, 76, 80, 84, 88, 92, 96, 100], "itemsize": 112})
GZ_MATERIAL = np.dtype({"names": ["type", "flags", "reflectance", "eta", "roughness", "conductorRoughness",
                                  "conductorEta", "conductorK", "thickness", "albedo", "g", "maxDepth",
                                  "nSamples"],
                        "formats": ["<i4", "<i4", ("<f4", 3), "<f4", ("<f4", 2), ("<f4", 2), ("<f4", 3),
                                    ("<f4", 3), "<f4", ("<f4", 3), "<f4", "<i4", "<i4"],
                        "offsets": [0, 4, 8, 20, 24, 32, 40, 52, 64, 68, 80, 84, 88], "itemsize": 96})
GZ_AREA_LIGHT = np.dtype({"names": ["brightness", "mesh", "triangle"], "formats": [("<f4", 3), "<i4", "<i8"],
                          "offsets": [0, 12, 16], "itemsize": 24})
PATH_STATS = np.dtype({"names": ["interior", "tris", "bounces"], "formats": ["<u4", "<u4", "<u4"],
                       "offsets": [0, 4, 8], "itemsize": 12})
GZ_PATH_ACTIVE = 1
GZ_PATH_SPECULAR = 2
GZ_MAT_DIFFUSE = 1
GZ_MAT_CONDUCTOR = 2
GZ_MAT_DIELECTRIC = 3
GZ_MAT_COATED_CONDUCTOR = 4
GZ_MAT_REMAP_ROUGHNESS = 1
GZ_MIS_FIX_LIGHT_PDF = 1
GZ_MIS_FIX_SKY_DIRECTION = 2

GZ_OK = 0
GZ_E_INVALID = -1
GZ_E_NO_DEVICE = -2
GZ_E_HIP = -3
GZ_E_OVERFLOW = -4
GZ_E_ALLOC = -5
GZ_STACK_CAPACITY = 64
GZ_NOALLOC_MAX_RAYS = 524288
GZ_BATCH_SOLO_PATHS = 4194304


class TriangleMeshC(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("faceIndices", ctypes.c_void_p), ("vertexIndices", ctypes.c_void_p)]


class SceneDescriptor2C(ctypes.Structure):
    _fields_ = [("bvh2Nodes", ctypes.c_void_p), ("primIds", ctypes.c_void_p), ("meshes", ctypes.c_void_p),
                ("meshCount", ctypes.c_int64)]


class GzSceneInfo(ctypes.Structure):
    _fields_ = [("nodeCount", ctypes.c_int64), ("interiorCount", ctypes.c_int64), ("primCount", ctypes.c_int64),
                ("triRecordCount", ctypes.c_int64), ("meshCount", ctypes.c_int64), ("materialCount", ctypes.c_int64),
                ("deviceBytes", ctypes.c_int64), ("maxDepth", ctypes.c_int32), ("rootIsLeaf", ctypes.c_int32)]


class GzTraceConfig(ctypes.Structure):
    _fields_ = [("persistent", ctypes.c_int32), ("ldsStack", ctypes.c_int32), ("refillLanes", ctypes.c_int32),
                ("gridScale", ctypes.c_int32), ("chunk", ctypes.c_int32), ("binMinRays", ctypes.c_int32),
                ("recordTimes", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class GzZSobolParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_int64), ("log2SamplesPerPixel", ctypes.c_int32), ("nBase4Digits", ctypes.c_int32)]


class GzMisLaunch(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int32), ("block", ctypes.c_int32), ("gridScale", ctypes.c_int32),
                ("chunk", ctypes.c_int32), ("drainLevel", ctypes.c_int32), ("reserved", ctypes.c_int32 * 3)]


class GzPlyInfo(ctypes.Structure):
    _fields_ = [("vertexCount", ctypes.c_int64), ("faceCount", ctypes.c_int64), ("indexCount", ctypes.c_int64),
                ("hasNormals", ctypes.c_int32), ("hasUvs", ctypes.c_int32), ("hasFaceIndices", ctypes.c_int32),
                ("bigEndian", ctypes.c_int32)]


class GzBuildOptions(ctypes.Structure):
    _fields_ = [("maxLeaf", ctypes.c_int32), ("sah", ctypes.c_int32), ("reserved", ctypes.c_int32 * 6)]


class GzMisParams(ctypes.Structure):
    _fields_ = [("maxDepth", ctypes.c_int32), ("flags", ctypes.c_int32), ("zsSeed", ctypes.c_int64),
                ("zsLog2Spp", ctypes.c_int32), ("zsBase4Digits", ctypes.c_int32), ("rrSeed", ctypes.c_uint64)]


class GzCamera(ctypes.Structure):
    _fields_ = [("rasterToCamera", ctypes.c_float * 16), ("cameraToWorld", ctypes.c_float * 16),
                ("resX", ctypes.c_int32), ("resY", ctypes.c_int32), ("x0", ctypes.c_int32), ("y0", ctypes.c_int32),
                ("x1", ctypes.c_int32), ("y1", ctypes.c_int32), ("spp", ctypes.c_int32),
                ("filterType", ctypes.c_int32), ("filterRadius", ctypes.c_float * 2), ("filterSigma", ctypes.c_float),
                ("lensRadius", ctypes.c_float), ("focalDistance", ctypes.c_float), ("pad", ctypes.c_int32),
                ("zsSeed", ctypes.c_int64), ("zsLog2Spp", ctypes.c_int32), ("zsBase4Digits", ctypes.c_int32)]


class GzBatchScene(ctypes.Structure):
    _fields_ = [("handle", ctypes.c_void_p), ("camera", GzCamera), ("params", GzMisParams), ("iso", ctypes.c_float),
                ("maxComponentValue", ctypes.c_float), ("d_light", ctypes.c_void_p), ("d_weight", ctypes.c_void_p),
                ("d_albedo", ctypes.c_void_p), ("d_normal", ctypes.c_void_p)]


class GzBatchInfo(ctypes.Structure):
    _fields_ = [("totalPaths", ctypes.c_int64), ("rounds", ctypes.c_int32), ("items", ctypes.c_int32),
                ("roundPaths", ctypes.c_int32), ("maxItemsPerRound", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("allBsdfs", ctypes.c_int32), ("soloScenes", ctypes.c_int32), ("sideScenes", ctypes.c_int32)]


GZ_FILTER_GAUSSIAN = 0
GZ_FILTER_BOX = 1


class GzError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32

# name -> (restype, argtypes)
SIGNATURES = {
    "gz_gpu_available": (ctypes.c_bool, []),
    "gz_gpu_upload_scene": (_P, [_P, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _I64]),
    "gz_gpu_traverse_batch": (None, [_P, _P, _P, _I64, _P]),
    "gz_gpu_shade_batch": (None, [_P, _P, _I64, _P]),
    "gz_gpu_free_scene": (None, [_P]),
    "gz_traverse_bvh2": (None, [_P, _P, ctypes.c_float, _P]),
    "gz_cpu_traverse_batch": (None, [_P, _P, _P, _I64, _P]),
    "gz_cpu_shade_batch": (None, [_P, _I64, _P, _P, _P]),
    "gz_last_error": (ctypes.c_char_p, []),
    "gz_gpu_scene_info": (ctypes.c_int, [_P, _P]),
    "gz_gpu_occluded_batch": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "gz_gpu_traverse_batch_noalloc": (ctypes.c_int, [_P, _P, _P, _I64, _P]),
    "gz_gpu_trace_times": (ctypes.c_int, [_P, _P, _P, _P]),
    "gz_gpu_traverse_closest_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "gz_gpu_traverse_any_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "gz_gpu_traverse_closest_stats_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P]),
    "gz_gpu_traverse_any_stats_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P]),
    "gz_gpu_expand_hits_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P]),
    "gz_gpu_check_traversal": (ctypes.c_int, [_P, _P]),
    "gz_gpu_set_trace_config": (ctypes.c_int, [_P, _P]),
    "gz_gpu_get_trace_config": (ctypes.c_int, [_P, _P]),
    "gz_gpu_shade_device": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "gz_zsobol_params": (ctypes.c_int, [_I64, _I64, _I64, _I64, _P]),
    "gz_zsobol_samples_device": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I32, _P, _I32, _P, _I32, _P]),
    "gz_zsobol_tile_device": (ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P, _I32, _P]),
    "gz_sobol_matrices": (ctypes.c_int, [_P]),
    "gz_bvh2_build": (_I64, [_P, _I64, _P, _P]),
    "gz_gpu_set_mis_scene": (ctypes.c_int, [_P, _P, _I32, _P, _I32, _P]),
    "gz_gpu_mis_bounce_device": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "gz_gpu_mis_trace_device": (ctypes.c_int, [_P, _P, _I64, _P, _P]),
    "gz_gpu_mis_trace_stats_device": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    "gz_gpu_mis_adjoint_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P]),
    "gz_mis_record_bytes": (ctypes.c_int64, [_I64, ctypes.c_int32]),
    "gz_gpu_mis_trace_record_device": (ctypes.c_int, [_P, _P, _I64, _P, _P, ctypes.c_int32, _P]),
    "gz_gpu_mis_adjoint_record_device": (ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P, ctypes.c_int32, _P]),
    "gz_gpu_set_mis_launch": (ctypes.c_int, [_P, _P]),
    "gz_gpu_get_mis_launch": (ctypes.c_int, [_P, _P]),
    "gz_gpu_release_mis_queues": (ctypes.c_int, [_P]),
    "gz_gpu_camera_paths_device": (ctypes.c_int, [_P, _I64, _I64, _P, _P]),
    "gz_gpu_film_add_device": (ctypes.c_int, [_P, _P, _I64, ctypes.c_float, ctypes.c_float, _P, _P, _P, _P, _P]),
    "gz_gpu_film_resolve_device": (ctypes.c_int, [_P, _P, _I64, _P, _P, _P]),
    "gz_ply_parse": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P]),
    "gz_gpu_build_scene": (ctypes.c_void_p, [_P, _I64, _P, _I64, _P, _P, _P, _P, _I64, _P]),
    "gz_gpu_export_bvh2": (_I64, [_P, _P, _I64, _P]),
    "gz_batch_create": (ctypes.c_void_p, [_P, ctypes.c_int32, _I64, _I64]),
    "gz_batch_info": (ctypes.c_int, [_P, _P]),
    "gz_batch_set_launch": (ctypes.c_int, [_P, _P]),
    "gz_batch_render_device": (ctypes.c_int, [_P, ctypes.c_int32, _P]),
    "gz_batch_check": (ctypes.c_int, [_P, _P]),
    "gz_batch_free": (None, [_P]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Loads libgonzales_amd.so (built in-tree). Raises if it is missing: no fallback exists."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m gonzales_amd.build` "
                           "(the HIP library is required, there is no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().gz_last_error().decode("utf-8", "replace")


def check(code: int, what: str = "") -> None:
    if code != GZ_OK:
        raise GzError(code, f"{what}: {last_error()}" if what else last_error())


def concat(parts, dtype: np.dtype) -> np.ndarray:
    """Concatenates record arrays into `dtype` with its C layout. (np.concatenate of records with
    explicit offsets returns a packed dtype without the padding the C structs have.)"""
    parts = [np.asarray(p, dtype) for p in parts]
    out = np.zeros(sum(len(p) for p in parts), dtype)
    at = 0
    for p in parts:
        out[at:at + len(p)] = p
        at += len(p)
    return out


def ptr(a) -> int:
    """Address of a numpy array (host) or a torch tensor (device or host)."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())


def stream_ptr(stream=None) -> int:
    """hipStream_t of a torch stream (default: torch's current stream on the current device)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)
