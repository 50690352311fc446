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


def test_reference_test_search_body():
    """tests/cpp/test_search_ref.cpp: the reference driver's body
    (test/test_search.cpp:17-64) compiled against include/curplsh (only the
    CUDA-only calls replaced), run on unified memory, printing the reference's
    output and checking the closed-form answer."""
    exe = os.path.join(ROOT, "tests", "cpp", "build", "test_search_ref")
    assert os.path.exists(exe), "build the driver first (__graft_entry__.build())"
    p = subprocess.run([exe, "0"], capture_output=True, text=True, timeout=110)
    print(p.stdout[-2000:])
    print(p.stderr[-2000:])
    assert p.returncode == 0
    assert "test_search: closed form ok" in p.stdout
    first = p.stdout.splitlines()
    row = next(ln for ln in first if ln.startswith("("))
    assert row.split("\t")[:3] == ["(1,-128)", "(0,0)", "(2,0)"]
