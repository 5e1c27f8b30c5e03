"""Build recipe for the native parts: the gfx950 HIP library and the CPU oracle.

``python -m gonzales_amd.build`` (or ``__graft_entry__.build()``) compiles
``gonzales_amd/libgonzales_amd.so`` in-tree with hipcc for gfx950 and, for the tests,
``oracle/liboracle.so`` with gcc. Compilation always runs (it is the "does it build" check).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libgonzales_amd.so"
ORACLE = ROOT / "oracle"

ARCH = os.environ.get("GZ_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["gz_scene.hip", "gz_traverse.hip", "gz_shade.hip", "gz_zsobol.hip", "gz_mis.hip", "gz_film.hip",
               "gz_batch.hip", "gz_build.hip"]
CXX_SOURCES = ["gz_error.cpp", "bvh_build.cpp", "sobol_table.cpp", "gz_ply.cpp", "gz_cpu.cpp"]

# Device code: no fused multiply-add contraction (the traversal is bit-exact against the
# reference's separately rounded arithmetic), IEEE division/sqrt (hipcc's default), and the
# packed-FP32 instructions disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build gonzales_amd)")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def _gen_sobol_data() -> Path:
    """Turns the Joe-Kuo direction numbers (data file) into a C array include."""
    src = PKG / "data" / "new-joe-kuo-6.1023"
    out = BUILD / "joe_kuo_data.inc"
    rows = []
    with open(src, "r", encoding="ascii") as f:
        for line in f:
            parts = line.split()
            if not parts or not parts[0].isdigit():
                continue
            s, a = int(parts[1]), int(parts[2])
            ms = [int(x) for x in parts[3:3 + s]]
            rows.append((s, a, ms))
    if len(rows) != 1023:
        raise RuntimeError(f"{src}: expected 1023 direction-number rows, found {len(rows)}")
    maxs = max(r[0] for r in rows)
    text = [f"// generated from {src.name}: dims 1..1023 as (s, a, m_1..m_s)", f"#define GZ_JK_MAXS {maxs}"]
    text.append("static const struct { int s; int a; unsigned m[GZ_JK_MAXS]; } kJoeKuo[1023] = {")
    for s, a, ms in rows:
        text.append("  {%d, %d, {%s}}," % (s, a, ", ".join(str(m) for m in ms)))
    text.append("};")
    out.write_text("\n".join(text) + "\n")
    return out


def build_library(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    _gen_sobol_data()
    hipcc = _hipcc()
    cxx = shutil.which("g++") or "g++"
    inc = [f"-I{INCLUDE}", f"-I{CSRC}", f"-I{BUILD}"]
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append([hipcc, *HIP_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
    for src in CXX_SOURCES:
        obj = BUILD / (src + ".o")
        objs.append(obj)
        jobs.append([cxx, *CXX_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in jobs]:
            if verbose:
                print(" ".join(cmd))
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
HOST_ASAN_LIB = BUILD / "asan" / "libgz_host_asan.so"


def build_host_sanitized() -> Path:
    """The host-only sources (CXX_SOURCES: error state, SAH builder, Sobol table, PLY reader, CPU
    traversal and shade) with AddressSanitizer and UndefinedBehaviorSanitizer, as one shared
    library for the CPU robustness tests (loaded into a process that preloads libasan)."""
    out = HOST_ASAN_LIB
    out.parent.mkdir(parents=True, exist_ok=True)
    _gen_sobol_data()
    cxx = shutil.which("g++") or "g++"
    flags = [f for f in CXX_FLAGS if f != "-O2"] + ["-O1", *SANITIZE]
    _run([cxx, *flags, f"-I{INCLUDE}", f"-I{CSRC}", f"-I{BUILD}", "-shared", "-o", str(out),
          *[str(CSRC / s) for s in CXX_SOURCES], "-lpthread"])
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE)])
    return ORACLE / "liboracle.so"


def main() -> int:
    lib = build_library(verbose="-v" in sys.argv)
    print(f"built {lib}")
    orc = build_oracle()
    print(f"built {orc}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
