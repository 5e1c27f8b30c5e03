"""GPU parity of the Z-Sobol sampler against the oracle and the committed golden vectors
(bit-exact float32 values, ZSobolSampler.swift:6-280)."""
import ctypes
from pathlib import Path

import numpy as np
import pytest

import oracle
from gonzales_amd import abi

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "zsobol_vectors.npz"


def _params(spp, rx, ry, seed):
    p = abi.GzZSobolParams()
    abi.check(abi.load().gz_zsobol_params(spp, rx, ry, seed, ctypes.byref(p)))
    return p


def _device_samples(spp, rx, ry, seed, px, py, si, pattern, start=0, layout=0):
    import torch

    lib = abi.load()
    p = _params(spp, rx, ry, seed)
    pat = np.ascontiguousarray(pattern, np.int8)
    n = len(px)
    stride = int(pat.sum())
    dpx = torch.from_numpy(np.ascontiguousarray(px, np.int32)).cuda()
    dpy = torch.from_numpy(np.ascontiguousarray(py, np.int32)).cuda()
    dsi = torch.from_numpy(np.ascontiguousarray(si, np.int64)).cuda()
    out = torch.empty(n * stride, dtype=torch.float32, device="cuda")
    abi.check(lib.gz_zsobol_samples_device(ctypes.byref(p), abi.ptr(dpx), abi.ptr(dpy), abi.ptr(dsi), n, start,
                                           pat.ctypes.data, len(pat), abi.ptr(out), layout, abi.stream_ptr()))
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    return o.reshape(n, stride) if layout == 0 else o.reshape(stride, n).T


def test_golden_vectors_bit_exact(gpu):
    data = np.load(GOLDEN)
    names = sorted({k.rsplit("_", 1)[0] for k in data.files})
    assert len(names) >= 6
    for name in names:
        spp, rx, ry, seed, start = (int(x) for x in data[name + "_args"])
        got = _device_samples(spp, rx, ry, seed, data[name + "_px"], data[name + "_py"], data[name + "_si"],
                              data[name + "_pattern"], start)
        assert np.array_equal(got.view(np.uint32), data[name + "_out"].view(np.uint32)), name


@pytest.mark.parametrize("spp,rx,ry,seed", [(64, 1920, 1080, 1234), (16, 512, 512, 0), (2, 4096, 4096, 9)])
def test_random_samples_match_oracle(gpu, spp, rx, ry, seed):
    g = np.random.default_rng(spp)
    n = 100000
    px, py, si = g.integers(0, rx, n), g.integers(0, ry, n), g.integers(0, spp, n)
    pat = [2, 2] + [2, 3, 2, 3] * 4
    ref = oracle.zsobol_batch(spp, rx, ry, seed, px, py, si, pat)
    got = _device_samples(spp, rx, ry, seed, px, py, si, pat, layout=1)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_tile_order_matches_tile_render_loop(gpu):
    import torch

    lib = abi.load()
    spp, w, h, x0, y0 = 16, 24, 10, 100, 40
    p = _params(spp, 512, 512, 1234)
    pat = np.array([2, 2, 3], np.int8)
    out = torch.empty(w * h * spp * 7, dtype=torch.float32, device="cuda")
    abi.check(lib.gz_zsobol_tile_device(ctypes.byref(p), x0, y0, w, h, spp, 0, pat.ctypes.data, 3, abi.ptr(out), 0,
                                        abi.stream_ptr()))
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(-1, 7)
    # for pixel in bounds (row-major), for sampleIndex in 0..<spp (Core/Tile.swift:28-29)
    pix = np.arange(w * h)
    px = np.repeat(x0 + pix % w, spp)
    py = np.repeat(y0 + pix // w, spp)
    si = np.tile(np.arange(spp), w * h)
    ref = oracle.zsobol_batch(spp, 512, 512, 1234, px, py, si, pat)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_full_frame_stratification(gpu):
    # 1920x1080 at 64 spp (132.7M samples): every sampled pixel's 64 camera samples form a
    # (0, 6, 2)-net in dims (0, 1).
    import torch

    lib = abi.load()
    spp = 64
    p = _params(spp, 1920, 1080, 1234)
    pat = np.array([2], np.int8)
    out = torch.empty(1920 * 1080 * spp * 2, dtype=torch.float32, device="cuda")
    abi.check(lib.gz_zsobol_tile_device(ctypes.byref(p), 0, 0, 1920, 1080, spp, 0, pat.ctypes.data, 1,
                                        abi.ptr(out), 0, abi.stream_ptr()))
    torch.cuda.synchronize()
    v = out.view(1920 * 1080, spp, 2)
    g = np.random.default_rng(1)
    for pix in g.integers(0, 1920 * 1080, 200):
        s = v[int(pix)].cpu().numpy()
        for a in range(7):
            cells = np.floor(s[:, 0] * (1 << a)).astype(int) * (1 << (6 - a)) + np.floor(s[:, 1] * (1 << (6 - a))).astype(int)
            assert len(set(cells.tolist())) == spp
    assert float(out.min()) >= 0.0 and float(out.max()) < 1.0
