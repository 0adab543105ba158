"""ctypes binding of the C ABI in include/bolt_hip.h and include/bolt_ops.h.

The shared library is loaded from this package directory; there is no
fallback: a missing or stale library raises at import of the operator layer.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libbolt_hip.so"

# ---- status codes (bolt_ops.h) -------------------------------------------
BOLT_OK = 0
BOLT_ERR_PLAN = 1
BOLT_ERR_TYPE = 2
BOLT_ERR_MEMORY = 3
BOLT_DECLINE = 4
BOLT_ERR_UNSUPPORTED = 5
BOLT_ERR_OTHER = 6
BOLT_ERR_DEVICE = 7
BOLT_ERR_NO_DEVICE = 100000
BOLT_HIP_ERROR_NOT_READY = 600

# ---- dtypes ----------------------------------------------------------------
BOOL, I32, I64, F32, F64, DEC128, DATE32, TIMESTAMP = 1, 2, 3, 4, 5, 6, 7, 8
DTYPE_SIZE = {BOOL: 1, I32: 4, I64: 8, F32: 4, F64: 8, DEC128: 16, DATE32: 4, TIMESTAMP: 8}
# Timestamp units (a Timestamp column's `scale`): digits of a second
TS_S, TS_MS, TS_US, TS_NS = 0, 3, 6, 9

# ---- aggregates ------------------------------------------------------------
SUM, COUNT, COUNT_STAR, MIN, MAX, AVG, VAR_POP, VAR_SAMP, STDDEV_POP, STDDEV_SAMP = range(1, 11)
JOIN_CARRY = -2

# ---- comparisons -----------------------------------------------------------
LT, LE, GT, GE, EQ, NE, IS_NULL, IS_NOT_NULL = range(1, 9)
MAX_PRED_TERMS = 8

GROUPBY_SORTED = 1
GROUPBY_FORCE_GLOBAL = 2
GROUPBY_FORCE_PARTITIONED = 4


class BoltColumn(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("validity", C.c_void_p),
        ("length", C.c_int64),
        ("dtype", C.c_int32),
        ("scale", C.c_int32),
    ]


class BoltAgg(C.Structure):
    _fields_ = [("op", C.c_int32), ("col", C.c_int32)]


class BoltScalar(C.Structure):
    _fields_ = [
        ("dtype", C.c_int32),
        ("is_null", C.c_int32),
        ("i64", C.c_int64),
        ("f64", C.c_double),
        ("hi", C.c_int64),
    ]


class BoltExprOp(C.Structure):
    """bolt_expr_op_t: one op of a per-row expression program."""
    _fields_ = [
        ("kind", C.c_int32), ("dst", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("c", C.c_int32),
        ("col", C.c_int32), ("op", C.c_int32), ("dtype", C.c_int32),
        ("lit_i64", C.c_int64), ("lit_f64", C.c_double), ("lit_hi", C.c_int64),
    ]


class BoltPredTerm(C.Structure):
    _fields_ = [
        ("col", C.c_int32),
        ("op", C.c_int32),
        ("lit_i64", C.c_int64),
        ("lit_f64", C.c_double),
        ("lit_hi", C.c_int64),
    ]


class BoltFilter(C.Structure):
    _fields_ = [("terms", C.POINTER(BoltPredTerm)), ("n_terms", C.c_int32)]


class BoltGbTune(C.Structure):
    _fields_ = [("block", C.c_int32), ("tile_rows", C.c_int32), ("lds_slots", C.c_int32), ("grid", C.c_int32)]


class BoltServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("hbm_budget", C.c_int64), ("chunk_rows", C.c_int64)]


class BoltServeDim(C.Structure):
    _fields_ = [("key", BoltColumn), ("attrs", C.POINTER(BoltColumn)), ("n_attrs", C.c_int32)]


SERVE_MAX_DIMS, SERVE_MAX_AGGS, SERVE_MAX_TERMS, SERVE_MAX_LIMIT = 3, 4, 4, 1024


class BoltServeQuery(C.Structure):
    _fields_ = [
        ("n_terms", C.c_int32),
        ("terms", BoltPredTerm * SERVE_MAX_TERMS),
        ("n_joins", C.c_int32),
        ("join_fk", C.c_int32 * SERVE_MAX_DIMS),
        ("join_dim", C.c_int32 * SERVE_MAX_DIMS),
        ("n_group", C.c_int32),
        ("group_dim", C.c_int32 * 2),
        ("group_attr", C.c_int32 * 2),
        ("n_aggs", C.c_int32),
        ("aggs", BoltAgg * SERVE_MAX_AGGS),
        ("order_agg", C.c_int32),
        ("limit", C.c_int32),
        ("est_groups", C.c_int64),
    ]


class BoltServeResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("n_rows", C.c_int32),
        ("n_groups", C.c_int64),
        ("hbm_bytes", C.c_int64),
        ("n_cols", C.c_int32),
        ("cells", C.POINTER(BoltScalar)),
        ("submit_ms", C.c_double),
        ("admit_ms", C.c_double),
        ("done_ms", C.c_double),
        ("error", C.c_char_p),
    ]


# ---- errors ------------------------------------------------------------------

class BoltError(RuntimeError):
    """Base error; `code` is the C status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class PlanError(BoltError):
    pass


class BoltTypeError(BoltError):
    pass


class BoltMemoryError(BoltError):
    pass


class GpuCapacity(BoltError):
    """The operator declined the input; the caller reruns it on its host path."""


class UnsupportedError(BoltError):
    pass


class DeviceError(BoltError):
    pass


class NoDeviceError(BoltError):
    """No GPU is usable (the stub contract of the reference)."""


_ERR = {
    BOLT_ERR_PLAN: PlanError,
    BOLT_ERR_TYPE: BoltTypeError,
    BOLT_ERR_MEMORY: BoltMemoryError,
    BOLT_DECLINE: GpuCapacity,
    BOLT_ERR_UNSUPPORTED: UnsupportedError,
    BOLT_ERR_DEVICE: DeviceError,
    BOLT_ERR_NO_DEVICE: NoDeviceError,
}

_lib: C.CDLL | None = None

# name -> (restype, argtypes)
_SIGS = {
    "bolt_last_error": (C.c_char_p, []),
    "bolt_build_info": (C.c_char_p, []),
    "bolt_jit_cache_stats": (C.c_int, [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "bolt_profile_events": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bolt_hip_get_error_name": (C.c_int, [C.c_int, C.POINTER(C.c_char_p)]),
    "bolt_hip_get_error_string": (C.c_int, [C.c_int, C.POINTER(C.c_char_p)]),
    "bolt_hip_init": (C.c_int, [C.c_uint]),
    "bolt_hip_device_get": (C.c_int, [C.POINTER(C.c_int), C.c_int]),
    "bolt_hip_ctx_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint, C.c_int]),
    "bolt_hip_ctx_destroy": (C.c_int, [C.c_void_p]),
    "bolt_hip_ctx_set_current": (C.c_int, [C.c_void_p]),
    "bolt_hip_ctx_get_current": (C.c_int, [C.POINTER(C.c_void_p)]),
    "bolt_hip_device_get_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bolt_hip_device_get_name": (C.c_int, [C.c_char_p, C.c_int, C.c_int]),
    "bolt_hip_device_get_attribute": (C.c_int, [C.POINTER(C.c_int), C.c_int, C.c_int]),
    "bolt_hip_device_total_mem": (C.c_int, [C.POINTER(C.c_size_t), C.c_int]),
    "bolt_hip_set_device": (C.c_int, [C.c_int]),
    "bolt_hip_get_device": (C.c_int, [C.POINTER(C.c_int)]),
    "bolt_hip_device_synchronize": (C.c_int, []),
    "bolt_hip_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "bolt_hip_free": (C.c_int, [C.c_void_p]),
    "bolt_hip_malloc_async": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.c_void_p]),
    "bolt_hip_free_async": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bolt_hip_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]),
    "bolt_hip_host_free": (C.c_int, [C.c_void_p]),
    "bolt_hip_mem_get_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "bolt_hip_memcpy_htod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bolt_hip_memcpy_dtoh": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bolt_hip_memcpy_dtod": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bolt_hip_memcpy_htod_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "bolt_hip_memcpy_dtoh_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "bolt_hip_memcpy_dtod_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "bolt_hip_memset_d8": (C.c_int, [C.c_void_p, C.c_ubyte, C.c_size_t]),
    "bolt_hip_memset_d8_async": (C.c_int, [C.c_void_p, C.c_ubyte, C.c_size_t, C.c_void_p]),
    "bolt_hip_module_load_data": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p]),
    "bolt_hip_module_load_data_ex": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, C.c_uint, C.POINTER(C.c_int),
                                               C.POINTER(C.c_void_p)]),
    "bolt_hip_module_unload": (C.c_int, [C.c_void_p]),
    "bolt_hip_module_get_function": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, C.c_char_p]),
    "bolt_hip_launch_kernel": (C.c_int, [C.c_void_p, C.c_uint, C.c_uint, C.c_uint, C.c_uint, C.c_uint,
                                         C.c_uint, C.c_uint, C.c_void_p, C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_void_p)]),
    "bolt_hip_stream_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint]),
    "bolt_hip_stream_destroy": (C.c_int, [C.c_void_p]),
    "bolt_hip_stream_synchronize": (C.c_int, [C.c_void_p]),
    "bolt_hip_stream_pool_acquire": (C.c_int, [C.POINTER(C.c_void_p)]),
    "bolt_hip_stream_pool_release": (C.c_int, [C.c_void_p]),
    "bolt_hip_event_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint]),
    "bolt_hip_event_record": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bolt_hip_event_query": (C.c_int, [C.c_void_p]),
    "bolt_hip_event_synchronize": (C.c_int, [C.c_void_p]),
    "bolt_hip_event_elapsed_ms": (C.c_int, [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]),
    "bolt_hip_event_destroy": (C.c_int, [C.c_void_p]),
    "bolt_hip_stream_begin_capture": (C.c_int, [C.c_void_p, C.c_uint]),
    "bolt_hip_stream_end_capture": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "bolt_hip_graph_instantiate": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, C.POINTER(C.c_void_p), C.c_char_p,
                                             C.c_size_t]),
    "bolt_hip_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "bolt_hip_graph_exec_destroy": (C.c_int, [C.c_void_p]),
    "bolt_hip_graph_destroy": (C.c_int, [C.c_void_p]),
    "bolt_hip_pool_stats": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    # operators
    "bolt_aggregate": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltAgg), C.c_int32,
                                 C.POINTER(BoltFilter), C.POINTER(BoltScalar)]),
    "bolt_aggregate_noalloc": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltAgg), C.c_int32,
                                         C.POINTER(BoltFilter), C.POINTER(BoltScalar)]),
    "bolt_filter_compact": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltFilter),
                                      C.c_void_p, C.POINTER(BoltColumn), C.POINTER(C.c_int64)]),
    "bolt_prefix_scan_mask": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_int64)]),
    "bolt_project": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltExprOp), C.c_int32,
                               C.POINTER(BoltColumn), C.c_int32]),
    "bolt_aggregate_expr": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltExprOp), C.c_int32,
                                      C.c_int32, C.c_int32, C.POINTER(BoltAgg), C.c_int32, C.POINTER(BoltScalar)]),
    "bolt_groupby_expr": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltExprOp), C.c_int32,
                                    C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                    C.POINTER(BoltAgg), C.c_int32, C.c_uint32, C.POINTER(C.c_void_p)]),
    "bolt_join_probe_aggregate_expr": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BoltColumn), C.c_int32,
                                                 C.POINTER(BoltExprOp), C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                                 C.c_int32, C.c_int32, C.POINTER(BoltAgg), C.c_int32,
                                                 C.POINTER(BoltScalar)]),
    "bolt_take": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_void_p, C.c_int64, C.POINTER(BoltColumn)]),
    "bolt_groupby": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltColumn), C.c_int32,
                               C.POINTER(BoltAgg), C.c_int32, C.POINTER(BoltFilter), C.c_uint32,
                               C.POINTER(C.c_void_p)]),
    "bolt_groupby_last_plan": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int32)]),
    "bolt_groupby_autotune_partitions": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32,
                                                   C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltAgg), C.c_int32,
                                                   C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float)]),
    # Arrow C Data Interface ingestion (bolt_arrow.h)
    "bolt_arrow_format_dtype": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "bolt_gpu_table_from_arrow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "bolt_gpu_table_append_arrow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bolt_gpu_table_reserve": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p]),
    "bolt_gpu_table_num_rows": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "bolt_gpu_table_num_columns": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "bolt_gpu_table_column": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(BoltColumn)]),
    "bolt_gpu_table_column_name": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]),
    "bolt_gpu_table_release": (C.c_int, [C.c_void_p]),
    "bolt_column_to_arrow": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_char_p, C.c_void_p, C.c_void_p]),
    "bolt_table_num_rows": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "bolt_table_num_columns": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "bolt_table_column": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(BoltColumn)]),
    "bolt_table_release": (None, [C.c_void_p]),
    "bolt_join_build": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.POINTER(BoltColumn), C.POINTER(C.c_void_p)]),
    "bolt_join_table_unique": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "bolt_join_table_release": (None, [C.c_void_p]),
    "bolt_join_probe": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.c_int32,
                                  C.POINTER(BoltFilter), C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "bolt_join_probe_aggregate": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.c_int32,
                                            C.POINTER(BoltFilter), C.POINTER(BoltAgg), C.c_int32,
                                            C.POINTER(BoltScalar)]),
    "bolt_inner_join": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.POINTER(BoltColumn), C.c_void_p, C.c_void_p,
                                  C.c_int64, C.POINTER(C.c_int64)]),
    "bolt_join": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.POINTER(BoltColumn), C.c_int32, C.c_int32,
                            C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "bolt_sort_pairs_i64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p,
                                      C.c_void_p]),
    "bolt_sort_limit": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.c_int64, C.c_int64, C.c_void_p, C.POINTER(C.c_int64)]),
    "bolt_sort_indices": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int32), C.c_void_p]),
    # batched GROUP BY (bolt_batch.h)
    "bolt_gb_batch_create": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltColumn),
                                       C.c_int32, C.c_int32, C.POINTER(BoltAgg), C.c_int32, C.POINTER(BoltGbTune),
                                       C.POINTER(C.c_void_p)]),
    "bolt_gb_batch_run": (C.c_int, [C.c_void_p]),
    "bolt_gb_batch_num_groups": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]),
    "bolt_gb_batch_column": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(BoltColumn)]),
    "bolt_gb_batch_get_tune": (C.c_int, [C.c_void_p, C.POINTER(BoltGbTune), C.POINTER(C.c_int64)]),
    "bolt_gb_batch_release": (None, [C.c_void_p]),
    "bolt_gb_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(BoltGbTune), C.POINTER(C.c_float)]),
    # served form (bolt_serve.h)
    "bolt_server_create": (C.c_int, [C.c_void_p, C.POINTER(BoltColumn), C.c_int32, C.POINTER(BoltServeDim), C.c_int32,
                                     C.POINTER(BoltServeConfig), C.POINTER(C.c_void_p)]),
    "bolt_server_submit": (C.c_int, [C.c_void_p, C.POINTER(BoltServeQuery), C.POINTER(C.c_int64)]),
    "bolt_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(BoltServeResult)]),
    "bolt_server_result_release": (None, [C.c_void_p, C.c_int64]),
    "bolt_server_now_ms": (C.c_double, [C.c_void_p]),
    "bolt_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bolt_server_destroy": (None, [C.c_void_p]),
}


def header_symbols() -> list[str]:
    """Every function name declared in include/*.h (parsed from the headers)."""
    import re

    names = []
    for h in sorted((PKG.parent / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(bolt_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.append(m.group(1))
    return names


def lib() -> C.CDLL:
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m craton_bolt_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().bolt_last_error().decode(errors="replace")


def check(status: int, what: str = "") -> None:
    """Raise the typed error for a non-zero operator status."""
    if status == BOLT_OK:
        return
    msg = last_error()
    cls = _ERR.get(status, BoltError)
    raise cls(status, f"{what}: {msg}" if what else msg)


def check_rt(status: int, what: str = "") -> None:
    """Raise for a non-zero runtime (bolt_hip_*) status."""
    if status == 0:
        return
    msg = last_error()
    if status == BOLT_ERR_NO_DEVICE:
        raise NoDeviceError(status, msg)
    raise DeviceError(status, f"{what}: {msg} (status {status})")
