"""CPU robustness: the library's host C++ (SAH builder, PLY reader, CPU traversal and shade, Sobol
table, error state) and the oracle, built with AddressSanitizer and UndefinedBehaviorSanitizer,
run over fuzzed PLY files (truncations, byte flips, garbage, absurd header counts), random and
degenerate SAH builds with CPU traversal of each, and the oracle's whole MIS chain; then the
allocation-failure injection (GZ_FAIL_ALLOC_AFTER) of the builder. Each run is a child process
(tests/sanitized_driver.py) with the sanitizer runtimes preloaded; any report fails the test.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent


def _runtime(name: str) -> str:
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else ""


@pytest.fixture(scope="module")
def sanitized_libs():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    from gonzales_amd import build
    host = build.build_host_sanitized()
    proc = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan"], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    return host, ROOT / "oracle" / "_asan" / "liboracle.so", f"{asan}:{ubsan}"


def _run(libs, *extra):
    host, orc, preload = libs
    env = dict(os.environ, LD_PRELOAD=preload,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=86",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=87")
    env.pop("GZ_FAIL_ALLOC_AFTER", None)
    proc = subprocess.run([sys.executable, str(HERE / "sanitized_driver.py"), str(host), str(orc), *extra],
                          capture_output=True, text=True, env=env, timeout=600)
    out = proc.stdout + proc.stderr
    print(proc.stdout[-2000:])
    assert "AddressSanitizer" not in out and "runtime error:" not in out, out[-6000:]
    assert proc.returncode == 0, out[-6000:]
    return proc.stdout


def test_fuzzed_and_random_inputs_are_clean_under_the_sanitizers(sanitized_libs):
    out = _run(sanitized_libs)
    assert "sanitized run clean" in out


def test_injected_allocation_failures_are_reported(sanitized_libs):
    out = _run(sanitized_libs, "alloc")
    assert "allocation failures injected" in out
