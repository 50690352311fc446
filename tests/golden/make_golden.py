"""Writes the golden fixtures of tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`).

The reference ships no vectors for this path (no .fvecs / ground truth; its
drivers only print), so the fixtures are small problems regenerated from the
in-repo seeded generator, with the expected results of the pinned CPU oracle
(canonical FP32 restatement and FP64 brute force) and a checksum of the inputs
so generator drift is caught. Each case: name -> generator call, k.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import curplsh_amd as ca  # noqa: E402
import oracle  # noqa: E402

CASES = {
    # C1 shape, scaled down: U[0,1), seeds 0/1
    "uniform_2000x128_k10": (lambda: (ca.gen_uniform(2000, 128, 0), ca.gen_uniform(32, 128, 1)), 10),
    "uniform_2000x128_k100": (lambda: (ca.gen_uniform(2000, 128, 0), ca.gen_uniform(32, 128, 1)), 100),
    # C2 shape, scaled down: 256-centre mixture, sigma 0.1
    "mixture_4000x128_k100": (lambda: (ca.gen_mixture(4000, 128, 256, 0.1, 2, 3),
                                       ca.gen_mixture(32, 128, 256, 0.1, 2, 30)), 100),
    # C3 shape, scaled down: unit Gaussian d=96
    "unitgauss_5000x96_k10": (lambda: (ca.gen_gaussian(5000, 96, 4, True), ca.gen_gaussian(32, 96, 5, True)), 10),
    # C4 shape, scaled down: 5% duplicates
    "mixdup_4000x128_k10": (lambda: (ca.gen_mixture(4000, 128, 1024, 0.05, 6, 7, 0.05),
                                     ca.gen_mixture(32, 128, 1024, 0.05, 6, 70)), 10),
    # C5 dims: U[-1,1), d=64 and 256, odd sizes
    "uniform_3001x64_k1": (lambda: (ca.gen_uniform(3001, 64, 11, -1, 1), ca.gen_uniform(17, 64, 12, -1, 1)), 1),
    "uniform_1500x256_k33": (lambda: (ca.gen_uniform(1500, 256, 13, -1, 1), ca.gen_uniform(9, 256, 14, -1, 1)), 33),
}


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def make(name):
    gen, k = CASES[name]
    x, q = gen()
    I, D = oracle.search_canonical(x, q, k)
    I64, D64 = oracle.search_fp64(x, q, k)
    return dict(I=I, D=D, I64=I64, D64=D64, k=np.int32(k), sha=np.array(digest(x, q)))


def main():
    out = os.path.dirname(os.path.abspath(__file__))
    for name in CASES:
        np.savez_compressed(os.path.join(out, f"{name}.npz"), **make(name))
        print("wrote", name)


if __name__ == "__main__":
    main()
