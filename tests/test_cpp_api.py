"""Runs the C++ mirror of the gpu package (algo-fft_amd/host/algofft_gpu.hpp)
through its test program: the error paths of gpu/plan.go:17-29 without a
device, and transforms on device 0 with a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "gpu_plan_test")


def _run(mode):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/gpu_plan_test is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_api_cpu_paths():
    _run("cpu")


@pytest.mark.gpu
def test_cpp_api_on_gpu():
    _run("gpu")
