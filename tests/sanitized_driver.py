"""Child process of tests/test_sanitizers.py: runs the host C++ of the library and the oracle, both
built with AddressSanitizer and UndefinedBehaviorSanitizer, over fuzzed and random inputs. A
sanitizer report ends the process with a non-zero status; the parent checks the status and the
output.

  argv[1]  the sanitized host library (gonzales_amd.build.build_host_sanitized)
  argv[2]  the sanitized oracle (make -C oracle asan)
  argv[3]  optional: "alloc" = run the allocation-failure injection of the host build only

Not collected by pytest (its name does not start with test_).
"""
from __future__ import annotations

import ctypes
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

from gonzales_amd import abi  # noqa: E402

HOST = ["gz_last_error", "gz_bvh2_build", "gz_ply_parse", "gz_sobol_matrices", "gz_traverse_bvh2",
        "gz_cpu_traverse_batch", "gz_cpu_shade_batch"]


def bind_host(path: str) -> ctypes.CDLL:
    lib = ctypes.CDLL(path)
    for name in HOST:
        res, args = abi.SIGNATURES[name]
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    abi._lib = lib  # the package's host helpers (BVH build, PLY parse) now call the sanitized build
    return lib


def ply_fuzz(lib, rng) -> int:
    """Valid files of every layout, then truncations, byte flips, random garbage and headers with
    absurd counts. Every call must return (GZ_OK or an error) without a sanitizer report; an OK
    header pass must not claim more vertices or faces than the bytes could hold."""
    from test_ply import LAYOUTS, write_ply

    def parse(data: bytes) -> int:
        buf = np.frombuffer(data, np.uint8) if data else np.zeros(1, np.uint8)
        info = abi.GzPlyInfo()
        rc = lib.gz_ply_parse(buf.ctypes.data, len(data), ctypes.byref(info), None, None, None, None, None)
        if rc != abi.GZ_OK:
            return rc
        assert info.vertexCount * 12 <= len(data) and info.faceCount <= len(data), (info.vertexCount, len(data))
        pts = np.zeros((max(info.vertexCount, 1), 3), np.float32)
        nrm = np.zeros_like(pts)
        uvs = np.zeros((max(info.vertexCount, 1), 2), np.float32)
        idx = np.zeros(max(info.indexCount, 1), np.int64)
        fidx = np.zeros(max(info.faceCount, 1), np.int64)
        return lib.gz_ply_parse(buf.ctypes.data, len(data), ctypes.byref(info), pts.ctypes.data, nrm.ctypes.data,
                                uvs.ctypes.data, idx.ctypes.data, fidx.ctypes.data)

    calls = 0
    seeds = []
    for name, props in LAYOUTS.items():
        for big in (False, True):
            nv, nf = int(rng.integers(3, 40)), int(rng.integers(1, 30))
            verts = rng.standard_normal((nv, len(props))).astype(np.float32)
            faces = [list(rng.integers(0, nv, int(rng.choice([3, 4])))) for _ in range(nf)]
            data = write_ply(props, verts, faces, big=big, list_type=str(rng.choice(["uchar", "uint"])),
                             face_indices=list(range(nf)) if rng.random() < 0.5 else None)
            assert parse(data) == abi.GZ_OK, (name, big)
            seeds.append(data)
            calls += 1
    for data in seeds:
        head = data.index(b"end_header") + len(b"end_header\n")
        cuts = sorted(set(int(c) for c in rng.integers(0, len(data), 40)) | set(range(0, head + 1, 7)))
        for c in cuts:
            parse(data[:c])
            calls += 1
        for _ in range(120):
            b = bytearray(data)
            for _ in range(int(rng.integers(1, 5))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
            parse(bytes(b))
            calls += 1
    for _ in range(200):
        parse(bytes(rng.integers(0, 256, int(rng.integers(0, 300)), dtype=np.uint8)))
        calls += 1
    base = ["ply", "format binary_little_endian 1.0"]
    for lines in (["element vertex 4294967295", "property float x", "property float y", "property float z",
                   "element face 0", "property list uchar int vertex_indices"],
                  ["element vertex 3", "property float x", "property float y", "property float z",
                   "element face 9223372036854775807", "property list uchar int vertex_indices"],
                  ["element vertex -5", "property float x"],
                  ["element vertex 3", "property double x", "property float y", "property float z"],
                  ["element vertex 3", "property float x", "property float y", "property float z",
                   "element face 1", "property list uchar int vertex_indices"],
                  ["element vertex 1", "property float x", "property float y", "property float z",
                   "element face 1", "property list uint int vertex_indices"]):
        head = "\n".join(base + lines + ["end_header"]).encode() + b"\n"
        for tail in (b"", b"\x00" * 12, b"\x00" * 12 + b"\xff" + b"\x00" * 8,
                     b"\x00" * 12 + b"\xff\xff\xff\xff" + b"\x01\x00\x00\x00" * 3):
            parse(head + tail)
            calls += 1
    return calls


def bvh_and_cpu_traversal(lib, rng) -> int:
    """SAH builds over random, degenerate and extreme boxes, then CPU traversal of random rays over
    the resulting hierarchy (gz_cpu_traverse_batch), with the oracle's traversal beside it."""
    import oracle
    from gonzales_amd.accelerator import Primitives, Scene, TriangleMesh, build_bounding_hierarchy, build_bvh2
    from helpers import random_rays

    runs = 0
    for n in (1, 2, 3, 5, 17, 64, 1000, 4000):
        for kind in ("random", "same", "flat", "huge"):
            c = rng.random((n, 3)).astype(np.float32)
            if kind == "same":
                c[:] = 0.5
            elif kind == "flat":
                c[:, 1] = 0.25
            elif kind == "huge":
                c *= 1e30
            v = c[:, None, :] + 0.01 * (rng.random((n, 3, 3)) - 0.5).astype(np.float32)
            mesh = TriangleMesh.from_xyz(v.reshape(-1, 3).astype(np.float32), np.arange(3 * n, dtype=np.int64))
            mats = np.zeros(1, abi.MATERIAL)
            mats["type"] = 1
            bh = build_bounding_hierarchy(Scene([mesh], mats), Primitives.geometric(0, n, 0))
            rays, tmax = random_rays(512, runs, lo=-0.2, hi=1.2)
            rays["dirX"][::5] = 0.0
            desc, keep = bh.descriptor()
            out = np.zeros(len(rays), abi.INTERSECTION)
            lib.gz_cpu_traverse_batch(ctypes.addressof(desc), abi.ptr(rays), abi.ptr(tmax), len(rays), abi.ptr(out))
            ref, _ = oracle.traverse(ctypes.addressof(desc), len(bh.nodes), rays, tmax, abi.INTERSECTION, 2)
            assert np.array_equal(out["tHit"].view(np.uint32), ref["tHit"].view(np.uint32)), (n, kind)
            runs += 1
    for bad in (np.full((4, 6), np.nan, np.float32), np.zeros((0, 6), np.float32)):
        try:
            build_bvh2(bad)
        except abi.GzError:
            pass
        runs += 1
    table = np.zeros(1024 * 52, np.uint32)
    assert lib.gz_sobol_matrices(table.ctypes.data) == abi.GZ_OK
    return runs


def oracle_paths(rng) -> int:
    """The oracle's camera, MIS trace (every BSDF, sky and triangle lights) and film on small
    scenes, multi-threaded."""
    import oracle
    from gonzales_amd.camera import Film, PerspectiveCamera, look_at
    from mis_helpers import mixed_material_scene, oracle_render
    from gonzales_amd.integrator import MisOptions, area_lights_from_scene

    # the camera's sampler constants from the oracle's ZSobolSampler init (the library's
    # gz_zsobol_params lives in its device part, which this process does not load)
    def zsobol_params(spp, res_x, res_y, seed, out):
        st = oracle.ZSobolState()
        oracle.lib().orc_zsobol_init(ctypes.byref(st), spp, res_x, res_y, seed)
        z = out._obj
        z.seed, z.log2SamplesPerPixel, z.nBase4Digits = st.seed, st.log2spp, st.nBase4Digits
        return abi.GZ_OK

    abi._lib.gz_zsobol_params = zsobol_params
    bh, gz = mixed_material_scene()
    cam = PerspectiveCamera(look_at((0.0, 1.4, 3.5), (0.1, 0.5, -0.2), (0, 1, 0)), 35.0, Film(24, 16), 8, 3)
    opt = MisOptions(max_depth=6, zsobol_seed=cam.c.zsSeed, log2_spp=cam.c.zsLog2Spp,
                     base4_digits=cam.c.zsBase4Digits)
    img = oracle_render(bh, cam, opt, gz, area_lights_from_scene(bh), sky=(0.6, 0.7, 0.8), nthreads=4)
    assert np.isfinite(img).all()
    z = oracle.zsobol_batch(16, 64, 64, 1234, rng.integers(0, 64, 256), rng.integers(0, 64, 256),
                            rng.integers(0, 16, 256), np.array([1, 2, 3, 2], np.int8))
    assert np.isfinite(z).all()
    return 2


def alloc_injection(lib) -> int:
    """GZ_FAIL_ALLOC_AFTER=k makes the k-th allocation of the SAH builder fail: every k up to the
    build's own count must give GZ_E_ALLOC (no crash, no leak the sanitizer sees at exit)."""
    from gonzales_amd.accelerator import build_bvh2
    bounds = np.random.default_rng(1).random((300, 6)).astype(np.float32)
    bounds[:, 3:] = bounds[:, :3] + 0.01
    failed = 0
    for k in range(0, 64):
        os.environ["GZ_FAIL_ALLOC_AFTER"] = str(k)
        try:
            build_bvh2(bounds)
        except abi.GzError as e:
            assert e.code == abi.GZ_E_ALLOC and "injected" in str(e), str(e)
            failed += 1
    del os.environ["GZ_FAIL_ALLOC_AFTER"]
    build_bvh2(bounds)
    assert failed > 0
    return failed


def main() -> int:
    import oracle
    lib = bind_host(sys.argv[1])
    oracle._path, oracle._lib = Path(sys.argv[2]), None
    rng = np.random.default_rng(20261203)
    if len(sys.argv) > 3 and sys.argv[3] == "alloc":
        print(f"allocation failures injected: {alloc_injection(lib)}", flush=True)
        return 0
    print(f"ply parses: {ply_fuzz(lib, rng)}", flush=True)
    print(f"builds + traversals: {bvh_and_cpu_traversal(lib, rng)}", flush=True)
    print(f"oracle path runs: {oracle_paths(rng)}", flush=True)
    print("sanitized run clean", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
