"""C-ABI checks that need no GPU: the libraries load, every entry point the
headers declare is exported, argument errors come back as status codes with a
message (the reference aborts instead: include/internal/assertions.h:7-14),
and the host-side generator is deterministic."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "rplsh.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rplsh_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(ca):
    lib = ca.load()
    syms = declared_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ca._lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (rplsh_\w+)", out))
    assert set(syms) <= exported


def test_cpp_library_exports_index_api():
    so = os.path.join(ROOT, "curplsh_amd", "libcurplsh.so")
    assert os.path.exists(so)
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    for name in ["curplsh::IndexBF::search(int, float const*, int, int*, float*)",
                 "curplsh::IndexBF::add(int, float const*)",
                 "curplsh::IndexBF::reset()",
                 "curplsh::Index::Index(curplsh::DeviceResources*, int, curplsh::IndexConfig)",
                 "curplsh::searchL2Distance(",
                 "curplsh::computeL2Norm(",
                 "curplsh::l2Select(",
                 "curplsh::blockSelect(",
                 "curplsh::broadcastAlongRowsSum(",
                 "curplsh::IndexRPLSH::search(int, float const*, int, int*, float*)",
                 "curplsh::DeviceResources::getDefaultStream(int)"]:
        assert name in out, name


def test_python_binding_matches_header(ca):
    from curplsh_amd import _lib

    assert set(_lib._SIGS) == set(declared_symbols())


def test_invalid_arguments_return_status(ca):
    lib = ca.load()
    h = C.c_void_p()
    st = lib.rplsh_bf_create(None, 8, None, C.byref(h))
    assert st == 1 and b"resources" in lib.rplsh_last_error()
    res = C.c_void_p()
    assert lib.rplsh_resources_create(C.byref(res)) == 0
    st = lib.rplsh_bf_create(res, 0, None, C.byref(h))
    assert st == 1 and b"dimension" in lib.rplsh_last_error()
    assert lib.rplsh_bf_search(None, 1, None, 1, None, None) == 1
    assert lib.rplsh_gen_uniform(None, 10, 4, 0, 0.0, 1.0) == 1
    assert lib.rplsh_gen_mixture(None, 0, 4, 3, 0.1, 1, 2, 1.5) == 1
    assert lib.rplsh_resources_destroy(res) == 0


def test_no_device_here_is_not_a_crash(ca):
    n = ca.device_count()
    assert n >= 0


def test_generator_counter_based(ca):
    a = ca.gen_uniform(300, 17, 9, -1.0, 1.0)
    b = ca.gen_uniform(1000, 17, 9, -1.0, 1.0)
    assert np.array_equal(a, b[:300])
    assert a.min() >= -1.0 and a.max() < 1.0
    g1 = ca.gen_gaussian(2000, 96, 4, True)
    g2 = ca.gen_gaussian(100, 96, 4, True)
    assert np.array_equal(g1[:100], g2)
    np.testing.assert_allclose((g1.astype(np.float64) ** 2).sum(1), 1.0, rtol=1e-6)
    assert abs(g1.mean()) < 0.01


def test_generator_mixture_duplicates(ca):
    x = ca.gen_mixture(20000, 32, 64, 0.05, 6, 7, 0.05)
    _, inv, cnt = np.unique(x, axis=0, return_inverse=True, return_counts=True)
    dup_rows = int((cnt[inv.ravel()] > 1).sum() - (cnt > 1).sum())
    assert 0.04 * 20000 < dup_rows < 0.06 * 20000
    y = ca.gen_mixture(20000, 32, 64, 0.05, 6, 7, 0.0)
    assert np.unique(y, axis=0).shape[0] == 20000
    # same centres for another point seed: cluster structure shared
    z = ca.gen_mixture(100, 32, 64, 0.05, 6, 8, 0.0)
    assert not np.array_equal(z, y[:100])
