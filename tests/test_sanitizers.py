"""Host C/C++ code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Each test builds a driver with `-fsanitize=address,undefined -fno-sanitize-recover=all`
(host compilers only: no device code is involved) and runs it; any out-of-bounds
access, leak or undefined operation aborts the driver with a report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300, **kw)


def _check(run):
    print(run.stdout)
    assert run.returncode == 0, run.stderr[-3000:]
    assert "runtime error" not in run.stderr and "Sanitizer" not in run.stderr, run.stderr[-3000:]


def test_oracle_c_code_is_clean_under_sanitizers():
    """oracle/sanitize_check.c drives every C entry point of the oracle (precompute,
    colouring, steps under each initialisation strategy with two threads, damping,
    Chebyshev, energies, element derivatives, contact sweeps, distances)."""
    odir = os.path.join(ROOT, "oracle")
    if shutil.which("make") is None or shutil.which(os.environ.get("CC", "cc")) is None:
        pytest.skip("no C toolchain")
    build = _run(["make", "-C", odir, "sanitize_check"])
    if build.returncode != 0:
        pytest.skip("sanitizer runtime unavailable: " + build.stderr[-300:])
    _check(_run([os.path.join(odir, "sanitize_check")]))


def test_host_precompute_is_clean_under_sanitizers(tmp_path):
    """tests/cpp/sanitize_data.cpp drives the library's host precompute
    (pbat_vbd_data_construct, csrc/vbd_data.cpp) through every colouring strategy, with
    and without per-element materials, and its error paths (inverted element, indices
    out of range, null output, invalid strategy)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "sanitize_data")
    build = _run([gxx, "-std=c++17", "-Wall", "-Wextra", *SAN, "-I", os.path.join(ROOT, "include"),
                  "-I", os.path.join(ROOT, "physicsbasedanimationtoolkit_amd", "csrc"),
                  os.path.join(ROOT, "tests", "cpp", "sanitize_data.cpp"),
                  os.path.join(ROOT, "physicsbasedanimationtoolkit_amd", "csrc", "vbd_data.cpp"), "-o", exe])
    if build.returncode != 0:
        pytest.skip("sanitizer runtime unavailable: " + build.stderr[-300:])
    _check(_run([exe]))
