"""The header-only C++ wrapper (include/pbat_vbd.hpp) driven from C++ (tests/cpp)."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_wrapper")


def run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built: run __graft_entry__.build()")
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_wrapper_host():
    run()


@pytest.mark.gpu
def test_cpp_wrapper_gpu(gpu_device):
    run("gpu")
