"""gz_gpu_traverse_batch_noalloc: the same Intersection_C records as gz_gpu_traverse_batch (bit for
bit) with every device buffer in the library's fixed per-GPU array, over two calls in a row (the
second smaller, so stale records of the first would show), across two scene handles sharing the
array, in every traversal launch mode; more rays than the array holds are refused before anything
runs; and the calls take no device memory.
"""
import ctypes

import numpy as np
import pytest

from gonzales_amd import abi, scenes
from helpers import assert_isect_equal, random_rays, random_soup_hierarchy

pytestmark = pytest.mark.gpu


def _noalloc(bh, rays, tmax):
    out = np.zeros(len(rays), abi.INTERSECTION)
    rc = abi.load().gz_gpu_traverse_batch_noalloc(bh.handle, abi.ptr(rays), abi.ptr(tmax), len(rays), abi.ptr(out))
    assert rc == abi.GZ_OK, abi.last_error()
    return out


def test_noalloc_equals_the_staged_batch(gpu):
    import torch

    soup = random_soup_hierarchy(200_000, seed=11)
    soup.upload_to_gpu()
    box, _ = scenes.cornell_hierarchy()
    box.upload_to_gpu()
    big_rays, big_t = random_rays(abi.GZ_NOALLOC_MAX_RAYS, seed=12)
    small_rays, small_t = random_rays(70_001, seed=13, lo=-1.0, hi=2.0)
    # references first, through the staged entry (which grows each handle's own staging)
    ref_big = soup.intersect_batch_gpu(big_rays, big_t)
    ref_small = soup.intersect_batch_gpu(small_rays, small_t)
    box_rays, box_t = random_rays(50_000, seed=14, lo=-1.0, hi=2.0, tmax=(1.0, 8.0))
    ref_box = box.intersect_batch_gpu(box_rays, box_t)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()

    print("noalloc: soup, max rays", flush=True)
    assert_isect_equal(_noalloc(soup, big_rays, big_t), ref_big)
    print("noalloc: soup, 70001 rays", flush=True)
    assert_isect_equal(_noalloc(soup, small_rays, small_t), ref_small)
    print("noalloc: Cornell box, 50000 rays", flush=True)
    assert_isect_equal(_noalloc(box, box_rays, box_t), ref_box)
    assert ref_big["hit"].mean() > 0.05 and ref_box["hit"].mean() > 0.05

    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    soup.free()
    box.free()


@pytest.mark.parametrize("persistent", [0, 1, 2, 3])
def test_noalloc_in_every_launch_mode(gpu, persistent):
    lib = abi.load()
    soup = random_soup_hierarchy(50_000, seed=21)
    soup.upload_to_gpu()
    c = abi.GzTraceConfig()
    abi.check(lib.gz_gpu_get_trace_config(soup.handle, ctypes.byref(c)), "get trace config")
    c.persistent = persistent
    abi.check(lib.gz_gpu_set_trace_config(soup.handle, ctypes.byref(c)), "set trace config")
    for n, seed in ((200_000, 22), (333, 23)):
        rays, t = random_rays(n, seed=seed)
        print(f"noalloc: persistent {persistent}, {n} rays", flush=True)
        assert_isect_equal(_noalloc(soup, rays, t), soup.intersect_batch_gpu(rays, t))
    soup.free()


def test_noalloc_refuses_more_rays_than_its_array_holds(gpu):
    box, _ = scenes.cornell_hierarchy()
    box.upload_to_gpu()
    n = abi.GZ_NOALLOC_MAX_RAYS + 1
    rays, t = random_rays(n, seed=31)
    out = np.zeros(n, abi.INTERSECTION)
    out["tHit"] = -7.0
    rc = abi.load().gz_gpu_traverse_batch_noalloc(box.handle, abi.ptr(rays), abi.ptr(t), n, abi.ptr(out))
    assert rc == abi.GZ_E_INVALID
    assert "exceed the fixed staging" in abi.last_error()
    assert np.all(out["tHit"] == -7.0)
    assert abi.load().gz_gpu_traverse_batch_noalloc(box.handle, None, None, 0, None) == abi.GZ_OK
    box.free()
