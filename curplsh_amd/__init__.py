"""curplsh_amd: MI355X-native brute-force L2 kNN behind the cuRPLSH IndexBF API.

Python mirror of the reference's C++ operator interface for this path
(include/index.h, include/index_bf.h, include/device_resources.h,
include/internal/{search,norm}.h), over the C ABI of libcurplsh_amd.so.
All compute runs in the HIP library; nothing here computes distances.

    res = DeviceResources()
    idx = IndexBF(res, dim, IndexBFConfig(precision="bf16"))
    idx.add(x)                      # float32 [n, dim] (host numpy or DeviceArray)
    I, D = idx.search(q, k)         # int32 / float32 [nq, k], ascending
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import PREC_BF16, PREC_FP8, PREC_FP32, RplshError, call, device_count, load

__all__ = [
    "DeviceResources", "DeviceArray", "IndexBF", "IndexBFConfig", "RplshError",
    "search_l2_distance", "compute_l2_norm", "l2_select", "block_select", "broadcast_rows_sum",
    "encode_bf16", "encode_e4m3",
    "gen_uniform", "gen_gaussian", "gen_mixture", "device_count", "Event", "load",
    "PREC_FP32", "PREC_BF16", "PREC_FP8", "Server", "IndexRPLSH", "IndexRPLSHConfig",
]

_PREC = {"fp32": PREC_FP32, "plain": PREC_FP32, "bf16": PREC_BF16, "fp8": PREC_FP8,
         PREC_FP32: PREC_FP32, PREC_BF16: PREC_BF16, PREC_FP8: PREC_FP8}


def _ptr(a) -> int:
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    raise TypeError(f"unsupported buffer type {type(a)}")


class DeviceResources:
    """DeviceResources (include/device_resources.h:60-263): per-device streams."""

    def __init__(self, device: int | None = None):
        load()
        self._h = C.c_void_p()
        call("rplsh_resources_create", C.byref(self._h))
        if device is not None:
            self.initialize_for_device(device)

    def initialize_for_device(self, device: int) -> None:
        call("rplsh_resources_init_device", self._h, device)

    def default_stream(self, device: int = 0) -> int:
        s = C.c_void_p()
        call("rplsh_resources_default_stream", self._h, device, C.byref(s))
        return s.value or 0

    def alternate_streams(self, device: int = 0) -> list[int]:
        n = C.c_int(8)
        arr = (C.c_void_p * 8)()
        call("rplsh_resources_alternate_streams", self._h, device, arr, C.byref(n))
        return [arr[i] or 0 for i in range(n.value)]

    def async_copy_stream(self, device: int = 0) -> int:
        s = C.c_void_p()
        call("rplsh_resources_async_copy_stream", self._h, device, C.byref(s))
        return s.value or 0

    def sync_default_stream(self, device: int = 0) -> None:
        call("rplsh_resources_sync_default_stream", self._h, device)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            call("rplsh_resources_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """hipMalloc-backed row-major buffer (the role of DeviceTensor,
    include/device_tensor.h:8-186), owned unless wrapping a foreign pointer."""

    def __init__(self, shape, dtype, device: int = 0, ptr: int | None = None):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._own = ptr is None
        if ptr is None:
            p = C.c_void_p()
            call("rplsh_malloc", device, self.nbytes, 0, C.byref(p))
            self.ptr = p.value or 0
        else:
            self.ptr = ptr

    @classmethod
    def from_host(cls, a: np.ndarray, device: int = 0, stream: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.copy_from(a, stream)
        return d

    def copy_from(self, a: np.ndarray, stream: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, "size mismatch"
        call("rplsh_memcpy_async", self.ptr, a.ctypes.data, self.nbytes, stream)
        call("rplsh_stream_sync", stream)

    def to_host(self, stream: int = 0) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        call("rplsh_memcpy_async", out.ctypes.data, self.ptr, self.nbytes, stream)
        call("rplsh_stream_sync", stream)
        return out

    def free(self) -> None:
        if self._own and self.ptr:
            call("rplsh_free", self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    """hipEvent wrapper for timing on a given stream."""

    def __init__(self):
        load()
        self._h = C.c_void_p()
        call("rplsh_event_create", C.byref(self._h))

    def record(self, stream: int = 0) -> None:
        call("rplsh_event_record", self._h, stream)

    def synchronize(self) -> None:
        call("rplsh_event_sync", self._h)

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        call("rplsh_event_elapsed_ms", self._h, end._h, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self._h.value:
                call("rplsh_event_destroy", self._h)
        except Exception:
            pass


@dataclass
class IndexBFConfig:
    """IndexConfig + IndexBFConfig (include/index.h:8-16, include/index_bf.h:8-19).

    precision: "fp32" (plain path), "bf16" or "fp8" (precision forms).
    kprime_factor: over-fetch k' = kprime_factor * k for the precision forms.
    residual_clusters: BF16/FP8 on clustered rows: > 0 adds the cluster-
    residual copy (rows minus the nearest of this many centroids) that the
    full filter level reads; 0 (default) off.
    """

    device: int = 0
    memory_space: int = 0
    store_transposed: bool = False
    compute_exact_distances: bool = False
    precision: str | int = "fp32"
    kprime_factor: int = 4
    fp64_boundary: bool = True
    persistent_max_queries: int = 32
    residual_clusters: int = 0

    def to_c(self) -> _lib.BfConfig:
        c = _lib.BfConfig()
        c.device = self.device
        c.memory_space = self.memory_space
        c.store_transposed = int(self.store_transposed)
        c.compute_exact_distances = int(self.compute_exact_distances)
        c.precision = _PREC[self.precision]
        c.kprime_factor = self.kprime_factor
        c.fp64_boundary = int(self.fp64_boundary)
        c.persistent_max_queries = int(self.persistent_max_queries)
        c.residual_clusters = int(self.residual_clusters)
        return c


@dataclass
class IndexRPLSHConfig:
    """rplsh_lsh_config: random-projection LSH in front of the exact re-rank.

    tables x bits: hash tables and code bits per table; probes: extra
    single-bit-flip buckets per table (query-directed); max_candidates: bucket
    entries taken per query before de-duplication (<= 8192; above 4096 only
    with prerank > 0 and prerank * k <= 4096); seed: projection
    seed (standard normal, the library's generator); center: hyperplane
    offsets at the projections' medians over the first add's first 65536 rows
    (L2 proximity) instead of through the origin (angle); prerank: f > 0 keeps
    the f*k candidates with the smallest BF16 distance for the exact re-rank."""

    device: int = 0
    tables: int = 16
    bits: int = 16
    probes: int = 0
    max_candidates: int = 1024
    seed: int = 1234
    center: bool = True
    prerank: int = 0

    def to_c(self) -> _lib.LshConfig:
        c = _lib.LshConfig()
        for f, _ in c._fields_:
            setattr(c, f, int(getattr(self, f)))
        return c


class IndexRPLSH:
    """Approximate kNN: random-projection LSH candidates, exact re-rank.

    The "RPLSH" step of the reference's name, which the reference does not
    implement (its only piece is the Gaussian projection initialiser,
    src/internal/matrix_ops.cu:71-76). search() returns exact squared L2
    distances of the candidates, ascending, ties by lower index, padded with
    -1 / FLT_MAX when a query has fewer than k distinct candidates."""

    def __init__(self, resources: DeviceResources, dim: int, config: IndexRPLSHConfig | None = None):
        self.config = config or IndexRPLSHConfig()
        self.resources = resources
        self._h = C.c_void_p()
        cfg = self.config.to_c()
        call("rplsh_lsh_create", resources.handle, int(dim), C.byref(cfg), C.byref(self._h))
        self.dim = int(dim)

    def train(self, x) -> None:
        """No training: the projections depend on the seed only."""

    def add(self, x) -> None:
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float32)
        n = int(x.shape[0])
        if n and x.shape[1] != self.dim:
            raise ValueError(f"expected dim {self.dim}, got {x.shape[1]}")
        call("rplsh_lsh_add", self._h, n, _ptr(x) if n else None)

    def search(self, q, k: int, indices=None, distances=None):
        if isinstance(q, np.ndarray):
            q = np.ascontiguousarray(q, dtype=np.float32)
        nq = int(q.shape[0])
        if indices is None:
            indices = np.empty((nq, k), np.int32)
        if distances is None:
            distances = np.empty((nq, k), np.float32)
        if nq:
            call("rplsh_lsh_search", self._h, nq, _ptr(q), int(k), _ptr(indices), _ptr(distances))
        return indices, distances

    def reset(self) -> None:
        call("rplsh_lsh_reset", self._h)

    @property
    def num_data(self) -> int:
        n = C.c_int64()
        call("rplsh_lsh_num_data", self._h, C.byref(n))
        return n.value

    def last_stats(self) -> dict:
        s = _lib.LshStats()
        call("rplsh_lsh_last_stats", self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def projections(self) -> np.ndarray:
        """W [tables*bits][dim] (float32)."""
        c = self.config
        w = np.empty((c.tables * c.bits, self.dim), np.float32)
        call("rplsh_lsh_projections", self._h, w.ctypes.data)
        return w

    def offsets(self) -> np.ndarray:
        """Hyperplane offsets [tables*bits] (float32)."""
        c = self.config
        o = np.empty(c.tables * c.bits, np.float32)
        call("rplsh_lsh_offsets", self._h, o.ctypes.data)
        return o

    def tables(self):
        """(codes in row order, sorted codes, rows in bucket order), each [tables][n]."""
        n, L = self.num_data, self.config.tables
        codes = np.empty((L, n), np.uint32)
        srt = np.empty((L, n), np.uint32)
        perm = np.empty((L, n), np.int32)
        call("rplsh_lsh_tables", self._h, codes.ctypes.data, srt.ctypes.data, perm.ctypes.data)
        return codes, srt, perm

    def hash_queries(self, q):
        """(codes [tables][nq] uint32, dot - offset [nq][tables*bits] float32)."""
        q = np.ascontiguousarray(q, dtype=np.float32)
        nq, c = q.shape[0], self.config
        codes = np.empty((c.tables, nq), np.uint32)
        marg = np.empty((nq, c.tables * c.bits), np.float32)
        call("rplsh_lsh_hash_queries", self._h, nq, q.ctypes.data, codes.ctypes.data, marg.ctypes.data)
        return codes, marg

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            call("rplsh_lsh_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IndexBF:
    """IndexBF (include/index_bf.h:21-46): exact brute-force L2 kNN.

    search() returns exact squared L2 distances, ascending, ties broken by the
    lower index; k > n pads with -1 / FLT_MAX (include/index.h:42-44).
    """

    def __init__(self, resources: DeviceResources, dim: int, config: IndexBFConfig | None = None):
        self.config = config or IndexBFConfig()
        self.resources = resources
        self._h = C.c_void_p()
        cfg = self.config.to_c()
        call("rplsh_bf_create", resources.handle, int(dim), C.byref(cfg), C.byref(self._h))
        self.dim = int(dim)

    def save(self, path: str) -> None:
        """Configuration, installed tuning entries and the FP32 rows."""
        call("rplsh_bf_save", self._h, os.fsencode(path))

    @classmethod
    def load(cls, resources: DeviceResources, path: str, device: int = -1) -> "IndexBF":
        """A new index from a file written by save(); device < 0: the saved one."""
        obj = cls.__new__(cls)
        obj.resources = resources
        obj._h = C.c_void_p()
        call("rplsh_bf_load", resources.handle, os.fsencode(path), int(device), C.byref(obj._h))
        d, dev = C.c_int(0), C.c_int(0)
        call("rplsh_bf_dimension", obj._h, C.byref(d))
        call("rplsh_bf_device", obj._h, C.byref(dev))
        obj.dim = int(d.value)
        obj.config = IndexBFConfig(device=int(dev.value))
        return obj

    # Index::train / add / search / reset (include/index.h:29-55)
    def train(self, x) -> None:
        n = 0 if x is None else int(x.shape[0])
        call("rplsh_bf_train", self._h, n, _ptr(x) if n else None)

    def add(self, x) -> None:
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float32)
        n = int(x.shape[0])
        if n and x.shape[1] != self.dim:
            raise ValueError(f"expected dim {self.dim}, got {x.shape[1]}")
        call("rplsh_bf_add", self._h, n, _ptr(x) if n else None)

    def search(self, q, k: int, indices=None, distances=None):
        """Returns (indices int32 [nq,k], distances float32 [nq,k]).

        q and the optional outputs may be numpy arrays (host) or DeviceArrays."""
        if isinstance(q, np.ndarray):
            q = np.ascontiguousarray(q, dtype=np.float32)
        nq = int(q.shape[0])
        if indices is None:
            indices = np.empty((nq, k), np.int32)
        if distances is None:
            distances = np.empty((nq, k), np.float32)
        if nq:
            call("rplsh_bf_search", self._h, nq, _ptr(q), int(k), _ptr(indices), _ptr(distances))
        return indices, distances

    def reset(self) -> None:
        call("rplsh_bf_reset", self._h)

    def set_precision(self, precision, kprime_factor: int = 0) -> None:
        call("rplsh_bf_set_precision", self._h, _PREC[precision], int(kprime_factor))

    @property
    def num_data(self) -> int:
        n = C.c_int64()
        call("rplsh_bf_num_data", self._h, C.byref(n))
        return n.value

    getNumData = num_data

    def get_dimension(self) -> int:
        return self.dim

    def get_device(self) -> int:
        return self.config.device

    def data_pointers(self) -> tuple[int, int]:
        x, n = C.c_void_p(), C.c_void_p()
        call("rplsh_bf_data_dev", self._h, C.byref(x), C.byref(n))
        return x.value or 0, n.value or 0

    def set_persistent(self, max_queries: int) -> None:
        """Largest batch search() runs in the persistent kernel (0: never)."""
        call("rplsh_bf_set_persistent", self._h, int(max_queries))

    def autotune(self, sample, k: int, batch_nq: int, target_recall: float = 0.999) -> dict:
        """rplsh_bf_autotune: tunes (k' factor, grid, survivor target) for k
        and this batch size, installs the result and returns it."""
        if isinstance(sample, np.ndarray):
            sample = np.ascontiguousarray(sample, dtype=np.float32)
        r = _lib.TuneResult()
        call("rplsh_bf_autotune", self._h, _ptr(sample), int(sample.shape[0]), int(k), int(batch_nq),
             float(target_recall), C.byref(r))
        return {f: getattr(r, f) for f, _ in r._fields_}

    def set_tuning(self, t: dict) -> None:
        r = _lib.TuneResult()
        for f, _ in r._fields_:
            setattr(r, f, t[f])
        call("rplsh_bf_set_tuning", self._h, C.byref(r))

    def clear_tuning(self) -> None:
        call("rplsh_bf_clear_tuning", self._h)

    def last_stats(self) -> dict:
        s = _lib.SearchStats()
        call("rplsh_bf_last_stats", self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            call("rplsh_bf_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HbmBudget:
    """rplsh_memory: a byte budget on one device with reusable blocks, the
    implementation behind curplsh::DeviceMemory (include/device_memory.h:12-89).
    reserve() returns a device pointer charged to the budget until release();
    a Server admits its batches against its own budget or one given here."""

    def __init__(self, device: int = 0, budget_bytes: int = 0, _handle: int | None = None):
        load()
        self._own = _handle is None
        self._h = C.c_void_p(_handle) if _handle is not None else C.c_void_p()
        if self._own:
            call("rplsh_memory_create", int(device), int(budget_bytes), C.byref(self._h))

    def reserve(self, nbytes: int, stream: int = 0, timeout_s: float = -1.0) -> int:
        p = C.c_void_p()
        call("rplsh_memory_reserve", self._h, stream, int(nbytes), float(timeout_s), C.byref(p))
        return p.value or 0

    def release(self, ptr: int) -> None:
        call("rplsh_memory_release", self._h, C.c_void_p(ptr))

    def stats(self) -> dict:
        st = _lib.MemoryStats()
        call("rplsh_memory_get_stats", self._h, C.byref(st))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self) -> None:
        if self._own and getattr(self, "_h", None) and self._h.value:
            call("rplsh_memory_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over resident indexes (rplsh_server_*).

    submit() enqueues a request (host arrays) and returns a handle; wait()
    returns (indices, distances, info). Requests for the same index and k that
    queue behind each other run as one batch; batches on different slots
    overlap on the GPU; admission is bounded by an HBM budget."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget_bytes: int = 0,
                 max_batch_queries: int = 4096, use_graphs: bool = True,
                 graph_max_queries: int = 1024, memory: "HbmBudget | None" = None,
                 persistent_max_queries: int = 256):
        load()
        cfg = _lib.ServerConfig()
        _lib.load().rplsh_server_config_default(C.byref(cfg))
        cfg.device, cfg.slots = device, slots
        cfg.hbm_budget_bytes, cfg.max_batch_queries = hbm_budget_bytes, max_batch_queries
        cfg.use_graphs, cfg.graph_max_queries = int(use_graphs), graph_max_queries
        cfg.memory = memory._h.value if memory is not None else None
        cfg.persistent_max_queries = int(persistent_max_queries)
        self._memory = memory
        self._h = C.c_void_p()
        call("rplsh_server_create", C.byref(cfg), C.byref(self._h))
        self._pending = {}
        self._indexes = []

    def add_index(self, index: "IndexBF") -> int:
        i = C.c_int()
        call("rplsh_server_add_index", self._h, index._h, C.byref(i))
        self._indexes.append(index)
        return i.value

    def submit(self, index_id: int, q: np.ndarray, k: int) -> int:
        q = np.ascontiguousarray(q, dtype=np.float32)
        I = np.empty((q.shape[0], k), np.int32)
        D = np.empty((q.shape[0], k), np.float32)
        rid = C.c_uint64()
        call("rplsh_server_submit", self._h, int(index_id), q.shape[0], q.ctypes.data, int(k),
             I.ctypes.data, D.ctypes.data, C.byref(rid))
        self._pending[rid.value] = (q, I, D)
        return rid.value

    def done(self, rid: int) -> bool:
        d = C.c_int()
        call("rplsh_server_poll", self._h, C.c_uint64(rid), C.byref(d))
        return bool(d.value)

    def wait(self, rid: int, timeout_s: float = -1.0):
        info = _lib.RequestInfo()
        try:
            call("rplsh_server_wait", self._h, C.c_uint64(rid), float(timeout_s), C.byref(info))
        except RplshError as e:
            if e.status != 4:  # a finished request (failed) releases its buffers
                self._pending.pop(rid, None)
            raise
        _, I, D = self._pending.pop(rid)
        return I, D, {f: getattr(info, f) for f, _ in info._fields_}

    def drain(self) -> None:
        call("rplsh_server_drain", self._h)

    def stats(self) -> dict:
        s = _lib.ServerStats()
        call("rplsh_server_get_stats", self._h, C.byref(s))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def reset_stats(self) -> None:
        call("rplsh_server_reset_stats", self._h)

    def memory(self) -> HbmBudget:
        """The budget the server admits against (not owned by the result)."""
        m = C.c_void_p()
        call("rplsh_server_memory", self._h, C.byref(m))
        return HbmBudget(_handle=m.value)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            call("rplsh_server_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_bytes(index: "IndexBF", nq: int, k: int) -> int:
    b = C.c_size_t()
    call("rplsh_bf_batch_bytes", index._h, int(nq), int(k), C.byref(b))
    return b.value


def compute_l2_norm(x: DeviceArray, out: DeviceArray, squared: bool = True, stream: int = 0) -> None:
    """computeL2Norm (include/internal/norm.h:7-8) on device arrays."""
    n, d = x.shape
    call("rplsh_l2norm_dev", x.ptr, n, d, out.ptr, int(squared), stream)
    call("rplsh_stream_sync", stream)


def search_l2_distance(bases: DeviceArray, bases_norm: DeviceArray | None, queries: DeviceArray,
                       k: int, indices: DeviceArray, distances: DeviceArray,
                       exact: bool = True, stream: int = 0) -> None:
    """searchL2Distance (include/internal/search.h:8-11), plain FP32 path."""
    n, d = bases.shape if bases is not None else (0, queries.shape[1])
    nq = queries.shape[0]
    call("rplsh_search_l2_dev", bases.ptr if bases is not None else None, n, d,
         bases_norm.ptr if bases_norm is not None else None, queries.ptr, nq, int(k),
         indices.ptr, distances.ptr, int(exact), stream)


def search_l2_distance_noalloc(bases: DeviceArray, bases_norm: DeviceArray | None, queries: DeviceArray,
                               k: int, indices: DeviceArray, distances: DeviceArray,
                               exact: bool = True, stream: int = 0) -> None:
    """searchL2Distance through rplsh_search_l2_dev_noalloc: same results,
    every temporary from the library's fixed per-device scratch array (no
    device allocation); sizes that do not fit raise INVALID_ARG."""
    n, d = bases.shape if bases is not None else (0, queries.shape[1])
    nq = queries.shape[0]
    call("rplsh_search_l2_dev_noalloc", bases.ptr if bases is not None else None, n, d,
         bases_norm.ptr if bases_norm is not None else None, queries.ptr, nq, int(k),
         indices.ptr, distances.ptr, int(exact), stream)


def noalloc_scratch_bytes() -> int:
    return int(_lib.load().rplsh_search_l2_noalloc_scratch_bytes())


def l2_select(products: DeviceArray, base_norms: DeviceArray, k: int, distances: DeviceArray,
              indices: DeviceArray, stream: int = 0) -> None:
    """l2Select (include/internal/select.h:29-34): per row the k smallest
    base_norms[i] + products[r][i], ascending, ties by the lower i."""
    rows, n = products.shape
    call("rplsh_l2_select_dev", products.ptr, rows, n, n, base_norms.ptr, int(k), distances.ptr,
         indices.ptr, stream)


def block_select(values: DeviceArray, k: int, out_k: DeviceArray, out_v: DeviceArray,
                 labels: DeviceArray | None = None, select_max: bool = False, stream: int = 0) -> None:
    """blockSelect (include/internal/select.h:8-24): per row the k smallest
    (or largest) values with column or given labels, ties by the lower label."""
    rows, n = values.shape
    call("rplsh_block_select_dev", values.ptr, labels.ptr if labels is not None else None, rows, n,
         int(k), int(select_max), out_k.ptr, out_v.ptr, stream)


def broadcast_rows_sum(v: DeviceArray, m: DeviceArray, stream: int = 0) -> None:
    """broadcastAlongRowsSum (include/internal/broadcast.h:7-8): m[r][c] += v[r]."""
    rows, cols = m.shape
    call("rplsh_broadcast_rows_sum_dev", v.ptr, m.ptr, rows, cols, stream)
    call("rplsh_stream_sync", stream)


def encode_bf16(x: DeviceArray, ld: int, out: DeviceArray, stream: int = 0) -> None:
    n, d = x.shape
    call("rplsh_encode_bf16_dev", x.ptr, n, d, ld, out.ptr, stream)
    call("rplsh_stream_sync", stream)


def encode_e4m3(x: DeviceArray, ld: int, out: DeviceArray, scales: DeviceArray, stream: int = 0) -> None:
    n, d = x.shape
    call("rplsh_encode_e4m3_dev", x.ptr, n, d, ld, out.ptr, scales.ptr, stream)
    call("rplsh_stream_sync", stream)


# ----------------------------------------------------------- synthetic data
def gen_uniform(n: int, d: int, seed: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    call("rplsh_gen_uniform", out.ctypes.data, n, d, seed, lo, hi)
    return out


def gen_gaussian(n: int, d: int, seed: int, unit_normalize: bool = False) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    call("rplsh_gen_gaussian", out.ctypes.data, n, d, seed, int(unit_normalize))
    return out


def gen_mixture(n: int, d: int, centers: int, sigma: float, center_seed: int, point_seed: int,
                dup_fraction: float = 0.0) -> np.ndarray:
    out = np.empty((n, d), np.float32)
    call("rplsh_gen_mixture", out.ctypes.data, n, d, centers, sigma, center_seed, point_seed,
         dup_fraction)
    return out


# ---- datasets (include/dataset.h) -------------------------------------------

_XVECS_ELEM = {np.dtype(np.float32): 4, np.dtype(np.int32): 4, np.dtype(np.uint8): 1}


def xvecs_info(path: str, dtype=np.float32) -> tuple[int, int]:
    """(records, dimension) of a .fvecs / .ivecs / .bvecs file."""
    num, dim = C.c_int64(0), C.c_int(0)
    call("rplsh_xvecs_info", os.fsencode(path), _XVECS_ELEM[np.dtype(dtype)], C.byref(num), C.byref(dim))
    return int(num.value), int(dim.value)


def read_xvecs(path: str, dtype=np.float32, first: int = 0, count: int | None = None,
               device: int | None = None):
    """Records [first, first + count) as [count, dim]: a numpy array, or a
    DeviceArray on `device` (streamed through pinned staging)."""
    dt = np.dtype(dtype)
    num, dim = xvecs_info(path, dt)
    if count is None:
        count = num - first
    if device is None:
        out = np.empty((count, dim), dt)
        call("rplsh_xvecs_read", os.fsencode(path), _XVECS_ELEM[dt], first, count, out.ctypes.data, 0, 0)
        return out
    out = DeviceArray((count, dim), dt, device)
    call("rplsh_xvecs_read", os.fsencode(path), _XVECS_ELEM[dt], first, count, out.ptr, 1, device)
    return out


def read_fvecs(path: str, **kw):
    return read_xvecs(path, np.float32, **kw)


def read_ivecs(path: str, **kw):
    return read_xvecs(path, np.int32, **kw)


def read_bvecs(path: str, **kw):
    return read_xvecs(path, np.uint8, **kw)


def write_xvecs(path: str, a: np.ndarray, append: bool = False) -> None:
    a = np.ascontiguousarray(a)
    assert a.ndim == 2 and a.dtype in _XVECS_ELEM, "2-D float32 / int32 / uint8 array"
    call("rplsh_xvecs_write", os.fsencode(path), _XVECS_ELEM[a.dtype], a.ctypes.data, a.shape[0], a.shape[1],
         1 if append else 0)


def recall_at(result, truth, k: int | None = None, return_missed: bool = False):
    """Mean |first k of result ∩ first k of truth| / k over the rows (the
    reference's Dataset::evaluate); inputs are not modified."""
    r = np.ascontiguousarray(result.to_host() if isinstance(result, DeviceArray) else result, np.int32)
    t = np.ascontiguousarray(truth.to_host() if isinstance(truth, DeviceArray) else truth, np.int32)
    k = k or min(r.shape[1], t.shape[1])
    assert r.shape[0] == t.shape[0]
    missed = np.empty(r.shape[0], np.int32)
    v = C.c_double(0.0)
    call("rplsh_recall_at", r.ctypes.data, r.shape[0], r.shape[1], t.ctypes.data, t.shape[1], k,
         missed.ctypes.data, C.byref(v))
    return (float(v.value), missed) if return_missed else float(v.value)


class Dataset:
    """Base, query and ground-truth arrays of a TEXMEX-style dataset
    (Dataset / DatasetIrisa, dataset.h:49-190): `<dir>/<name>_base.fvecs`,
    `_query.fvecs`, `_groundtruth.ivecs`. Arrays are numpy, or DeviceArrays
    when `device` is given."""

    def __init__(self, name: str, datadir: str, device: int | None = None):
        self.name = name
        p = lambda s: os.path.join(datadir, f"{name}_{s}")  # noqa: E731
        self.base = read_fvecs(p("base.fvecs"), device=device)
        self.query = read_fvecs(p("query.fvecs"), device=device)
        self.groundtruth = read_ivecs(p("groundtruth.ivecs"))
        self.dimension = self.base.shape[1]
        self.num_base, self.num_query = self.base.shape[0], self.query.shape[0]
        self.gt_k = self.groundtruth.shape[1]
        assert self.groundtruth.shape[0] == self.num_query and self.query.shape[1] == self.dimension

    def evaluate(self, indices, k: int | None = None, return_missed: bool = False):
        return recall_at(indices, self.groundtruth, k or self.gt_k, return_missed)
