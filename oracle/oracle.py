"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_vp, _i32, _i64 = C.c_void_p, C.c_int, C.c_int64


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.oracle_norms_canonical.argtypes = [_vp, _i64, _i32, _vp]
    lib.oracle_norms_canonical.restype = None
    lib.oracle_dot_canonical.argtypes = [_vp, _vp, _i32]
    lib.oracle_dot_canonical.restype = C.c_float
    lib.oracle_search_canonical.argtypes = [_vp, _i64, _i32, _vp, _vp, _i64, _i32, _i32, _vp, _vp, _i32]
    lib.oracle_search_canonical.restype = None
    lib.oracle_search_fp64.argtypes = [_vp, _i64, _i32, _vp, _i64, _i32, _vp, _vp, _i32]
    lib.oracle_search_fp64.restype = None
    lib.oracle_pair_dist_fp64.argtypes = [_vp, _i32, _vp, _i64, _i32, _vp, _vp]
    lib.oracle_pair_dist_fp64.restype = None
    lib.oracle_recall.argtypes = [_vp, _vp, _i64, _i32]
    lib.oracle_recall.restype = C.c_double
    lib.oracle_bf16_rne.argtypes = [C.c_float]
    lib.oracle_bf16_rne.restype = C.c_uint16
    lib.oracle_bf16_to_float.argtypes = [C.c_uint16]
    lib.oracle_bf16_to_float.restype = C.c_float
    lib.oracle_e4m3_rne_sat.argtypes = [C.c_float]
    lib.oracle_e4m3_rne_sat.restype = C.c_uint8
    lib.oracle_e4m3_to_float.argtypes = [C.c_uint8]
    lib.oracle_e4m3_to_float.restype = C.c_float
    lib.oracle_bf16_tree_dot.argtypes = [_vp, _vp, _i32]
    lib.oracle_bf16_tree_dot.restype = C.c_float
    lib.oracle_lsh_hash.argtypes = [_vp, _i64, _i32, _vp, _vp, _i32, _i32, _vp, _vp]
    lib.oracle_lsh_hash.restype = None
    lib.oracle_lsh_search.argtypes = [_vp, _i64, _i32, _vp, _vp, _i32, _i32, _i32, _i32, _i32,
                                      _vp, _i64, _i32, _vp, _vp, _vp, _vp]
    lib.oracle_lsh_search.restype = None
    _lib = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def norms_canonical(x: np.ndarray) -> np.ndarray:
    x = _f32(x)
    out = np.empty(x.shape[0], np.float32)
    load().oracle_norms_canonical(x.ctypes.data, x.shape[0], x.shape[1], out.ctypes.data)
    return out


def search_canonical(x: np.ndarray, q: np.ndarray, k: int, exact: bool = True,
                     xnorms: np.ndarray | None = None, threads: int = 0):
    """Canonical FP32 restatement of searchL2Distance: (indices, distances)."""
    q = _f32(q)
    d = q.shape[1]
    x = _f32(x) if x is not None and x.size else np.zeros((0, d), np.float32)
    nq = q.shape[0]
    I = np.empty((nq, k), np.int32)
    D = np.empty((nq, k), np.float32)
    xn = None if xnorms is None else _f32(xnorms)
    load().oracle_search_canonical(x.ctypes.data, x.shape[0], d,
                                   xn.ctypes.data if xn is not None else None,
                                   q.ctypes.data, nq, k, int(exact), I.ctypes.data, D.ctypes.data,
                                   threads)
    return I, D


def search_fp64(x: np.ndarray, q: np.ndarray, k: int, threads: int = 0):
    """FP64 brute force: (indices, float64 distances), ties by lower index."""
    x, q = _f32(x), _f32(q)
    nq = q.shape[0]
    I = np.empty((nq, k), np.int32)
    D = np.empty((nq, k), np.float64)
    load().oracle_search_fp64(x.ctypes.data, x.shape[0], x.shape[1], q.ctypes.data, nq, k,
                              I.ctypes.data, D.ctypes.data, threads)
    return I, D


def pair_dist_fp64(x: np.ndarray, q: np.ndarray, idx: np.ndarray) -> np.ndarray:
    x, q = _f32(x), _f32(q)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.empty(idx.shape, np.float64)
    load().oracle_pair_dist_fp64(x.ctypes.data, x.shape[1], q.ctypes.data, idx.shape[0],
                                 idx.shape[1], idx.ctypes.data, out.ctypes.data)
    return out


def recall(result: np.ndarray, truth: np.ndarray) -> float:
    r = np.ascontiguousarray(result, dtype=np.int32)
    t = np.ascontiguousarray(truth, dtype=np.int32)
    assert r.shape == t.shape
    return load().oracle_recall(r.ctypes.data, t.ctypes.data, r.shape[0], r.shape[1])


def bf16_rne(v: float) -> int:
    return int(load().oracle_bf16_rne(float(v)))


def bf16_to_float(h: int) -> float:
    return float(load().oracle_bf16_to_float(int(h)))


def bf16_tree_dot(a: np.ndarray, b: np.ndarray) -> np.float32:
    """The LSH pre-rank's BF16 dot product (piece/tree order) of two vectors."""
    a, b = _f32(a), _f32(b)
    assert a.shape == b.shape and a.ndim == 1
    return np.float32(load().oracle_bf16_tree_dot(a.ctypes.data, b.ctypes.data, a.shape[0]))


def e4m3_rne_sat(v: float) -> int:
    return int(load().oracle_e4m3_rne_sat(float(v)))


def e4m3_to_float(b: int) -> float:
    return float(load().oracle_e4m3_to_float(int(b)))


def _off(off):
    return None if off is None else _f32(off)


def lsh_hash(x: np.ndarray, w: np.ndarray, L: int, K: int, off: np.ndarray | None = None):
    """LSH codes [L][n] (uint32) and margins dot - off [n][L*K] of the rows of x."""
    x, w, o = _f32(x), _f32(w), _off(off)
    n, d = x.shape
    codes = np.empty((L, n), np.uint32)
    marg = np.empty((n, L * K), np.float32)
    load().oracle_lsh_hash(x.ctypes.data, n, d, w.ctypes.data, o.ctypes.data if o is not None else None,
                           L, K, codes.ctypes.data, marg.ctypes.data)
    return codes, marg


def lsh_search(x: np.ndarray, w: np.ndarray, L: int, K: int, probes: int, budget: int,
               q: np.ndarray, k: int, off: np.ndarray | None = None, prerank: int = 0):
    """LSH search restatement: (indices, distances, raw entries, distinct rows)."""
    q = _f32(q)
    d = q.shape[1]
    x = _f32(x) if x is not None and x.size else np.zeros((0, d), np.float32)
    w, o = _f32(w), _off(off)
    nq = q.shape[0]
    I = np.empty((nq, k), np.int32)
    D = np.empty((nq, k), np.float32)
    raw, dist = C.c_int64(0), C.c_int64(0)
    load().oracle_lsh_search(x.ctypes.data, x.shape[0], d, w.ctypes.data,
                             o.ctypes.data if o is not None else None, L, K, probes, budget, prerank,
                             q.ctypes.data, nq, k, I.ctypes.data, D.ctypes.data, C.byref(raw),
                             C.byref(dist))
    return I, D, raw.value, dist.value
