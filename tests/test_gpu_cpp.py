"""Runs the C++ drop-in driver (tests/cpp/test_index_bf.cpp): curplsh::IndexBF,
searchL2Distance and computeL2Norm used from C++ exactly as the reference's
drivers use them, over libcurplsh.so -> libcurplsh_amd.so."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_driver():
    exe = os.path.join(ROOT, "tests", "cpp", "build", "test_index_bf")
    assert os.path.exists(exe), "build the driver first (__graft_entry__.build())"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=110)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0
    assert "all C++ tests passed" in p.stdout
