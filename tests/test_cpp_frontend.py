"""The C++ front end (xtensor_amd/cpp/xdevice.hpp) through its test program
tests/cpp/test_xdevice.cpp, built by xtensor_amd.build next to the library.

cpu: descriptor flattening, C++ result types and broadcast shapes, view
strides, exception types (no device access).
gpu: evaluation on the device checked against host loops (bit-exact
integer and affine work, exp within 2 ulp, sums within 1e-14 relative).
"""

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "xtensor_amd", "_build", "test_xdevice")


def _exe():
    if not os.path.exists(EXE):
        from xtensor_amd.build import build
        build()
    return EXE


def test_cpp_frontend_cpu():
    r = subprocess.run([_exe(), "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_frontend_gpu():
    npy_dir = os.path.join(ROOT, "tests", "golden", "xnpy")
    r = subprocess.run([_exe(), "gpu", npy_dir], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
