"""ctypes binding of libcurplsh_amd.so (the C ABI declared in include/rplsh.h).

The product path is the HIP library; this module only loads it and declares
the signatures. There is no CPU fallback: if the library is missing the import
of any compute entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcurplsh_amd.so")

OK = 0
STATUS_NAMES = {
    0: "OK",
    1: "INVALID_ARG",
    2: "HIP",
    3: "OUT_OF_MEMORY",
    4: "STATE",
    5: "UNSUPPORTED",
    6: "BUDGET",
    7: "NO_DEVICE",
}

PREC_FP32, PREC_BF16, PREC_FP8 = 0, 1, 2


class RplshError(RuntimeError):
    """A C-ABI call returned a non-zero rplsh_status."""

    def __init__(self, status: int, func: str, msg: str):
        self.status = status
        self.func = func
        super().__init__(f"{func}: {STATUS_NAMES.get(status, status)}: {msg}")


class BfConfig(C.Structure):
    """rplsh_bf_config: IndexConfig + IndexBFConfig (include/index.h:8-16,
    include/index_bf.h:8-19) plus the precision-form knobs."""

    _fields_ = [
        ("device", C.c_int),
        ("memory_space", C.c_int),
        ("store_transposed", C.c_int),
        ("compute_exact_distances", C.c_int),
        ("precision", C.c_int),
        ("kprime_factor", C.c_int),
        ("fp64_boundary", C.c_int),
        ("persistent_max_queries", C.c_int),
        ("residual_clusters", C.c_int),
    ]


class SearchStats(C.Structure):
    _fields_ = [
        ("queries", C.c_int64),
        ("candidates", C.c_int64),
        ("survivors", C.c_int64),
        ("levels", C.c_int32),
        ("retried_queries", C.c_int32),
        ("boundary_fp64", C.c_int32),
        ("kprime", C.c_int32),
        ("ms_total", C.c_float),
        ("ms_filter_full", C.c_float),
        ("ms_filter_samples", C.c_float),
        ("ms_select", C.c_float),
        ("ms_rerank", C.c_float),
        ("rows_filter_full", C.c_int64),
        ("filter_variant", C.c_int32),
        ("target_wgs", C.c_int32),
        ("persistent", C.c_int32),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("k", C.c_int),
        ("batch_nq", C.c_int64),
        ("kprime_factor", C.c_int),
        ("target_wgs", C.c_int),
        ("c_factor", C.c_int),
        ("variant", C.c_int),
        ("rho", C.c_int),
        ("recall", C.c_float),
        ("ms_default", C.c_float),
        ("ms_tuned", C.c_float),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("slots", C.c_int),
        ("hbm_budget_bytes", C.c_size_t),
        ("max_batch_queries", C.c_int),
        ("use_graphs", C.c_int),
        ("graph_max_queries", C.c_int),
        ("graph_cache_bytes", C.c_size_t),
        ("memory", C.c_void_p),
        ("persistent_max_queries", C.c_int),
    ]


class MemoryStats(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("budget_bytes", C.c_size_t),
        ("reserved_bytes", C.c_size_t),
        ("peak_reserved_bytes", C.c_size_t),
        ("cached_bytes", C.c_size_t),
        ("available_without_malloc", C.c_size_t),
        ("high_water_malloc_bytes", C.c_size_t),
        ("reservations", C.c_int64),
        ("waits", C.c_int64),
        ("mallocs", C.c_int64),
    ]


class RequestInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("latency_ms", C.c_double),
        ("queue_ms", C.c_double),
        ("batch_requests", C.c_int),
        ("batch_queries", C.c_int64),
        ("device_ms", C.c_float),
        ("kernel_batches", C.c_int),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("completed", C.c_int64),
        ("rejected", C.c_int64),
        ("queries", C.c_int64),
        ("batches", C.c_int64),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("mean_ms", C.c_double),
        ("max_ms", C.c_double),
        ("wall_s", C.c_double),
        ("qps", C.c_double),
        ("peak_reserved_bytes", C.c_size_t),
        ("peak_inflight", C.c_int),
        ("graph_launches", C.c_int64),
        ("graph_captures", C.c_int64),
        ("graph_cache_bytes", C.c_size_t),
        ("persistent_launches", C.c_int64),
        ("persistent_batches", C.c_int64),
    ]


_vp = C.c_void_p
_i32 = C.c_int
_i64 = C.c_int64
_u64 = C.c_uint64
_f32 = C.c_float
_pf = C.POINTER(C.c_float)

class LshConfig(C.Structure):
    """rplsh_lsh_config: the random-projection LSH front end."""

    _fields_ = [
        ("device", C.c_int),
        ("tables", C.c_int),
        ("bits", C.c_int),
        ("probes", C.c_int),
        ("max_candidates", C.c_int),
        ("seed", C.c_uint64),
        ("center", C.c_int),
        ("prerank", C.c_int),
    ]


class LshStats(C.Structure):
    """rplsh_lsh_stats: counters of the last IndexRPLSH search."""

    _fields_ = [
        ("queries", C.c_int64),
        ("raw_candidates", C.c_int64),
        ("distinct_candidates", C.c_int64),
        ("ms_total", C.c_float),
        ("ms_candidates", C.c_float),
        ("ms_rerank", C.c_float),
    ]


# name -> argtypes (restype is always c_int status unless listed in _RESTYPES)
_SIGS = {
    "rplsh_last_error": [],
    "rplsh_version": [],
    "rplsh_device_count": [C.POINTER(_i32)],
    "rplsh_get_device": [C.POINTER(_i32)],
    "rplsh_set_device": [_i32],
    "rplsh_pointer_device": [_vp, C.POINTER(_i32)],
    "rplsh_resources_create": [C.POINTER(_vp)],
    "rplsh_resources_destroy": [_vp],
    "rplsh_resources_init_device": [_vp, _i32],
    "rplsh_resources_default_stream": [_vp, _i32, C.POINTER(_vp)],
    "rplsh_resources_set_default_stream": [_vp, _i32, _vp],
    "rplsh_resources_alternate_streams": [_vp, _i32, C.POINTER(_vp), C.POINTER(_i32)],
    "rplsh_resources_async_copy_stream": [_vp, _i32, C.POINTER(_vp)],
    "rplsh_resources_sync_default_stream": [_vp, _i32],
    "rplsh_malloc": [_i32, C.c_size_t, _i32, C.POINTER(_vp)],
    "rplsh_free": [_vp],
    "rplsh_memcpy_async": [_vp, _vp, C.c_size_t, _vp],
    "rplsh_memset_async": [_vp, _i32, C.c_size_t, _vp],
    "rplsh_stream_sync": [_vp],
    "rplsh_device_sync": [],
    "rplsh_host_alloc": [C.c_size_t, C.POINTER(_vp)],
    "rplsh_host_free": [_vp],
    "rplsh_mem_info": [_i32, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "rplsh_event_create": [C.POINTER(_vp)],
    "rplsh_event_destroy": [_vp],
    "rplsh_event_record": [_vp, _vp],
    "rplsh_event_elapsed_ms": [_vp, _vp, _pf],
    "rplsh_event_sync": [_vp],
    "rplsh_l2norm_dev": [_vp, _i64, _i32, _vp, _i32, _vp],
    "rplsh_encode_bf16_dev": [_vp, _i64, _i32, _i32, _vp, _vp],
    "rplsh_encode_e4m3_dev": [_vp, _i64, _i32, _i32, _vp, _vp, _vp],
    "rplsh_search_l2_dev": [_vp, _i64, _i32, _vp, _vp, _i64, _i32, _vp, _vp, _i32, _vp],
    "rplsh_search_l2_dev_noalloc": [_vp, _i64, _i32, _vp, _vp, _i64, _i32, _vp, _vp, _i32, _vp],
    "rplsh_search_l2_noalloc_scratch_bytes": [],
    "rplsh_l2_select_dev": [_vp, _i64, _i64, _i64, _vp, _i32, _vp, _vp, _vp],
    "rplsh_block_select_dev": [_vp, _vp, _i64, _i64, _i32, _i32, _vp, _vp, _vp],
    "rplsh_broadcast_rows_sum_dev": [_vp, _vp, _i64, _i64, _vp],
    "rplsh_bf_config_default": [C.POINTER(BfConfig)],
    "rplsh_bf_create": [_vp, _i32, C.POINTER(BfConfig), C.POINTER(_vp)],
    "rplsh_bf_destroy": [_vp],
    "rplsh_bf_train": [_vp, _i64, _vp],
    "rplsh_bf_add": [_vp, _i64, _vp],
    "rplsh_bf_search": [_vp, _i64, _vp, _i32, _vp, _vp],
    "rplsh_bf_reset": [_vp],
    "rplsh_bf_num_data": [_vp, C.POINTER(_i64)],
    "rplsh_bf_dimension": [_vp, C.POINTER(_i32)],
    "rplsh_bf_device": [_vp, C.POINTER(_i32)],
    "rplsh_bf_set_persistent": [_vp, _i32],
    "rplsh_bf_set_precision": [_vp, _i32, _i32],
    "rplsh_bf_data_dev": [_vp, C.POINTER(_vp), C.POINTER(_vp)],
    "rplsh_bf_last_stats": [_vp, C.POINTER(SearchStats)],
    "rplsh_bf_autotune": [_vp, _vp, _i64, _i32, _i64, _f32, C.POINTER(TuneResult)],
    "rplsh_bf_set_tuning": [_vp, C.POINTER(TuneResult)],
    "rplsh_bf_clear_tuning": [_vp],
    "rplsh_memory_create": [_i32, C.c_size_t, C.POINTER(_vp)],
    "rplsh_memory_destroy": [_vp],
    "rplsh_memory_reserve": [_vp, _vp, C.c_size_t, C.c_double, C.POINTER(_vp)],
    "rplsh_memory_release": [_vp, _vp],
    "rplsh_memory_get_stats": [_vp, C.POINTER(MemoryStats)],
    "rplsh_server_memory": [_vp, C.POINTER(_vp)],
    "rplsh_server_config_default": [C.POINTER(ServerConfig)],
    "rplsh_server_create": [C.POINTER(ServerConfig), C.POINTER(_vp)],
    "rplsh_server_destroy": [_vp],
    "rplsh_server_add_index": [_vp, _vp, C.POINTER(_i32)],
    "rplsh_server_submit": [_vp, _i32, _i64, _vp, _i32, _vp, _vp, C.POINTER(C.c_uint64)],
    "rplsh_server_poll": [_vp, C.c_uint64, C.POINTER(_i32)],
    "rplsh_server_wait": [_vp, C.c_uint64, C.c_double, C.POINTER(RequestInfo)],
    "rplsh_server_drain": [_vp],
    "rplsh_server_get_stats": [_vp, C.POINTER(ServerStats)],
    "rplsh_server_reset_stats": [_vp],
    "rplsh_bf_batch_bytes": [_vp, _i64, _i32, C.POINTER(C.c_size_t)],
    "rplsh_gen_uniform": [_vp, _i64, _i32, _u64, _f32, _f32],
    "rplsh_gen_gaussian": [_vp, _i64, _i32, _u64, _i32],
    "rplsh_gen_mixture": [_vp, _i64, _i32, _i32, _f32, _u64, _u64, C.c_double],
    "rplsh_bf_save": [_vp, C.c_char_p],
    "rplsh_bf_load": [_vp, C.c_char_p, _i32, C.POINTER(_vp)],
    "rplsh_xvecs_info": [C.c_char_p, _i32, C.POINTER(_i64), C.POINTER(_i32)],
    "rplsh_xvecs_read": [C.c_char_p, _i32, _i64, _i64, _vp, _i32, _i32],
    "rplsh_xvecs_write": [C.c_char_p, _i32, _vp, _i64, _i32, _i32],
    "rplsh_recall_at": [_vp, _i64, _i32, _vp, _i32, _i32, _vp, C.POINTER(C.c_double)],
    "rplsh_lsh_config_default": [C.POINTER(LshConfig)],
    "rplsh_lsh_create": [_vp, _i32, C.POINTER(LshConfig), C.POINTER(_vp)],
    "rplsh_lsh_destroy": [_vp],
    "rplsh_lsh_add": [_vp, _i64, _vp],
    "rplsh_lsh_reset": [_vp],
    "rplsh_lsh_search": [_vp, _i64, _vp, _i32, _vp, _vp],
    "rplsh_lsh_last_stats": [_vp, C.POINTER(LshStats)],
    "rplsh_lsh_num_data": [_vp, C.POINTER(_i64)],
    "rplsh_lsh_get_config": [_vp, C.POINTER(LshConfig)],
    "rplsh_lsh_projections": [_vp, _vp],
    "rplsh_lsh_tables": [_vp, _vp, _vp, _vp],
    "rplsh_lsh_offsets": [_vp, _vp],
    "rplsh_lsh_hash_queries": [_vp, _i64, _vp, _vp, _vp],
}

_RESTYPES = {
    "rplsh_last_error": C.c_char_p,
    "rplsh_version": C.c_char_p,
    "rplsh_bf_config_default": None,
    "rplsh_server_config_default": None,
    "rplsh_lsh_config_default": None,
    "rplsh_search_l2_noalloc_scratch_bytes": C.c_size_t,
}

_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load the HIP library (once). Raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C curplsh_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point; raise RplshError on failure."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != OK:
        msg = lib.rplsh_last_error()
        raise RplshError(st, name, msg.decode() if msg else "")


def device_count() -> int:
    n = C.c_int(0)
    call("rplsh_device_count", C.byref(n))
    return n.value
