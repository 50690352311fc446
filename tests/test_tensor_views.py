"""Runs tests/cpp/test_tensor_views.cpp (host only): the reference Tensor
interface (include/tensor.h:16-449) as include/curplsh/tensor.h declares it."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tensor_views():
    exe = os.path.join(ROOT, "tests", "cpp", "build", "test_tensor_views")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp")], check=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(p.stdout, p.stderr)
    assert p.returncode == 0 and "tensor views ok" in p.stdout
