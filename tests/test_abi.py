"""The C-ABI library loads, exports every entry point include/gonzales_amd.h declares, and its
host-only functions behave (no GPU needed)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import oracle
from gonzales_amd import abi

HEADER = Path(__file__).resolve().parent.parent / "include" / "gonzales_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(gz_\w+)\s*\(", text)))


def test_header_declares_the_reference_replacements():
    names = declared_functions()
    for ref in ("available", "upload_scene", "traverse_batch", "shade_batch", "free_scene"):
        assert f"gz_gpu_{ref}" in names  # one per mojo_gpu_* of mojo_kernel.h:65-91


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(abi.SIGNATURES)


def test_struct_layouts_match_the_reference():
    assert abi.PRIM_ID.itemsize == 32 and abi.MATERIAL.itemsize == 28 and abi.RAY.itemsize == 24
    assert abi.INTERSECTION.itemsize == 48 and abi.INTERSECTION.fields["tHit"][1] == 32
    assert abi.INTERSECTION.fields["hit"][1] == 44
    assert abi.PATH_STATE.itemsize == 88 and abi.PATH_STATE.fields["pcgState"][1] == 64
    assert abi.PATH_STATE.fields["active"][1] == 80
    assert abi.BVH2_NODE.itemsize == 32 and abi.GZ_HIT.itemsize == 16
    assert ctypes.sizeof(abi.SceneDescriptor2C) == 32 and ctypes.sizeof(abi.TriangleMeshC) == 24


def test_sobol_matrices_equal_the_oracle():
    lib = abi.load()
    out = np.zeros(1024 * 52, np.uint32)
    assert lib.gz_sobol_matrices(out.ctypes.data) == abi.GZ_OK
    assert np.array_equal(out, oracle.sobol_matrices())


@pytest.mark.parametrize("spp,w,h", [(16, 512, 512), (64, 1920, 1080), (8, 100, 60), (12, 64, 64), (1, 1, 1)])
def test_zsobol_params_match_the_sampler_init(spp, w, h):
    lib = abi.load()
    p = abi.GzZSobolParams()
    assert lib.gz_zsobol_params(spp, w, h, 1234, ctypes.byref(p)) == abi.GZ_OK
    st = oracle.ZSobolState()
    oracle.lib().orc_zsobol_init(ctypes.byref(st), spp, w, h, 1234)
    assert (p.seed, p.log2SamplesPerPixel, p.nBase4Digits) == (st.seed, st.log2spp, st.nBase4Digits)


def test_device_entry_points_fail_cleanly_without_a_handle():
    lib = abi.load()
    assert lib.gz_gpu_traverse_closest_device(None, None, None, 0, None, None) == abi.GZ_E_INVALID
    assert "null scene handle" in abi.last_error()
    assert lib.gz_gpu_scene_info(None, None) == abi.GZ_E_INVALID
    lib.gz_gpu_traverse_batch(None, None, None, 10, None)  # NULL handle: no-op (kernel.mojo:616-618)
    assert lib.gz_gpu_traverse_batch_noalloc(None, None, None, 10, None) == abi.GZ_E_INVALID
    lib.gz_gpu_shade_batch(None, None, 10, None)
    lib.gz_gpu_free_scene(None)
    c = abi.GzMisLaunch(1, 256, 1, 256)
    assert lib.gz_gpu_set_mis_launch(None, ctypes.byref(c)) == abi.GZ_E_INVALID
    assert lib.gz_gpu_get_mis_launch(None, ctypes.byref(c)) == abi.GZ_E_INVALID
    assert ctypes.sizeof(abi.GzMisLaunch) == 32
    assert lib.gz_batch_create(None, 0, 1 << 20, abi.GZ_BATCH_SOLO_PATHS) is None
    assert "soloPaths" in abi.last_error()
    assert ctypes.sizeof(abi.GzBatchInfo) == 40
    assert lib.gz_mis_record_bytes(10, 3) == 10 * 3 * 33 and lib.gz_mis_record_bytes(-1, 3) == -1
    assert lib.gz_gpu_mis_trace_record_device(None, None, 0, None, None, -1, None) == abi.GZ_E_INVALID
    assert lib.gz_gpu_mis_adjoint_record_device(None, None, None, 0, None, None, None, 1, None) == abi.GZ_E_INVALID


def test_zsobol_rejects_bad_patterns():
    lib = abi.load()
    p = abi.GzZSobolParams(0, 4, 9)
    bad = np.array([1, 4], np.int8)
    rc = lib.gz_zsobol_samples_device(ctypes.byref(p), None, None, None, 1, 0, bad.ctypes.data, 2, None, 0, None)
    assert rc == abi.GZ_E_INVALID and "pattern" in abi.last_error()
    rc = lib.gz_zsobol_tile_device(ctypes.byref(p), 0, 0, 4, 4, 4, 1024, bad.ctypes.data, 1, None, 0, None)
    assert rc == abi.GZ_E_INVALID
