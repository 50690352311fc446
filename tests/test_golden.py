"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the seeded inputs still hash to the recorded digest, and the oracle still
produces the recorded canonical and FP64 results."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402

NAMES = sorted(make_golden.CASES)


def load_golden(name):
    return dict(np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False))


@pytest.mark.parametrize("name", NAMES)
def test_golden_reproduces(name, orc):
    g = load_golden(name)
    gen, k = make_golden.CASES[name]
    x, q = gen()
    assert make_golden.digest(x, q) == str(g["sha"]), "generator output drifted"
    I, D = orc.search_canonical(x, q, k)
    assert np.array_equal(I, g["I"]) and np.array_equal(D.view(np.uint32), g["D"].view(np.uint32))
    I64, D64 = orc.search_fp64(x, q, k)
    assert np.array_equal(I64, g["I64"]) and np.array_equal(D64, g["D64"])
