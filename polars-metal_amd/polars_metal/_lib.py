"""ctypes binding of libpmh.so, the C-ABI MI355X backend (include/pmh.h).

This is the binding a maintainer adds in place of the reference's PyO3 module
(crates/polars-metal-core/src/lib.rs:57-83).  It raises ImportError when the
shared library has not been built: there is no CPU fallback behind it.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("PMH_LIBRARY", _HERE / "libpmh.so"))

PMH_OK, PMH_E_VALUE, PMH_E_KEY, PMH_E_NOTIMPL, PMH_E_RUNTIME, PMH_E_COMPUTE = range(6)
PMH_HOST, PMH_DEVICE = 0, 1
PMH_OUT_HOST, PMH_OUT_DEVICE = 0, 1
PMH_NOALLOC_MAX_ROWS = 262144  # include/pmh.h

TAGS = {
    "I64": 0, "F64": 1, "Bool": 2, "I32": 3, "F32": 4, "I8": 5, "I16": 6,
    "U8": 7, "U16": 8, "U32": 9, "U64": 10, "Utf8": 11,
}
TAG_NAMES = {v: k for k, v in TAGS.items()}
TAG_ALIASES = {"Date": "I32", "String": "Utf8"}
TAG_SIZE = {"I64": 8, "F64": 8, "U64": 8, "I32": 4, "F32": 4, "U32": 4, "Utf8": 4,
            "I16": 2, "U16": 2, "I8": 1, "U8": 1, "Bool": 0}
CMP_OPS = {"Eq": 0, "Ne": 1, "Lt": 2, "Le": 3, "Gt": 4, "Ge": 5}


class PmhColumn(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char_p),
        ("tag", ctypes.c_int32),
        ("location", ctypes.c_int32),
        ("data", ctypes.c_void_p),
        ("data_len", ctypes.c_uint64),
        ("valid", ctypes.c_void_p),
        ("valid_len", ctypes.c_uint64),
    ]


class PmhOutColumn(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char_p),
        ("tag", ctypes.c_int32),
        ("location", ctypes.c_int32),
        ("data", ctypes.c_void_p),
        ("data_len", ctypes.c_uint64),
        ("valid", ctypes.c_void_p),
        ("valid_len", ctypes.c_uint64),
    ]


class PmhExecStats(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("n_passes", ctypes.c_int32),
        ("initial_capacity", ctypes.c_uint64),
        ("final_capacity", ctypes.c_uint64),
        ("n_groups", ctypes.c_uint64),
        ("n_selected", ctypes.c_uint64),
        ("device_ms", ctypes.c_double),
        ("sample_rows", ctypes.c_uint32),
        ("sample_distinct", ctypes.c_uint32),
        ("sample_f1", ctypes.c_uint32),
        ("sample_f2", ctypes.c_uint32),
        ("hot_keys", ctypes.c_uint32),
        ("partitions", ctypes.c_uint32),
        ("tail_rows", ctypes.c_uint64),
        ("distinct_estimate", ctypes.c_double),
        ("scan_kernel", ctypes.c_int32),
        ("sizer_choice", ctypes.c_int32),
    ]


class PmhSketchInfo(ctypes.Structure):
    _fields_ = [
        ("sample_size", ctypes.c_uint32),
        ("sample_rows", ctypes.c_uint32),
        ("distinct", ctypes.c_uint32),
        ("f1", ctypes.c_uint32),
        ("f2", ctypes.c_uint32),
        ("dropped", ctypes.c_uint32),
        ("null_keys", ctypes.c_uint32),
        ("hot_keys", ctypes.c_uint32),
        ("lane_bits", ctypes.c_uint32),
        ("scale", ctypes.c_uint32),
        ("hist", ctypes.c_uint32 * 1024),
        ("features", ctypes.c_float * 16),
        ("prediction", ctypes.c_float * 2),
        ("distinct_estimate", ctypes.c_double),
    ]


class PmhChunk(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_uint64),
        ("data", ctypes.c_void_p),
        ("valid", ctypes.c_void_p),
        ("data_bit_offset", ctypes.c_uint32),
        ("valid_bit_offset", ctypes.c_uint32),
    ]


class PmhResidentColumn(ctypes.Structure):
    _fields_ = [
        ("entry", ctypes.c_uint64),
        ("n_rows", ctypes.c_uint64),
        ("data", ctypes.c_void_p),
        ("data_len", ctypes.c_uint64),
        ("valid", ctypes.c_void_p),
        ("valid_len", ctypes.c_uint64),
        ("hit", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class PmhColcacheInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("entries", "bytes_resident", "capacity", "hits", "misses", "bytes_uploaded", "h2d_bytes")]


class PmhUtf8Column(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_uint64),
        ("location", ctypes.c_int32),
        ("large", ctypes.c_int32),
        ("offsets", ctypes.c_void_p),
        ("bytes", ctypes.c_void_p),
        ("bytes_len", ctypes.c_uint64),
        ("valid", ctypes.c_void_p),
        ("valid_bit_offset", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
    ]


class PmhBatchQuery(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_uint64),
        ("columns", ctypes.POINTER(PmhColumn)),
        ("n_columns", ctypes.c_uint32),
        ("predicate_json", ctypes.c_char_p),
    ]


class PmhBatchStats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("n_queries", ctypes.c_uint32),
        ("n_tiles", ctypes.c_uint32),
        ("n_partitions", ctypes.c_uint64),
        ("n_rerun", ctypes.c_uint32),
        ("tile_rows", ctypes.c_uint32),
        ("part_slots", ctypes.c_uint32),
        ("graph_reused", ctypes.c_int32),
        ("learned", ctypes.c_int32),
        ("n_fallback_keys", ctypes.c_uint32),
    ]


class PmhResult(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_uint64),
        ("n_columns", ctypes.c_uint32),
        ("columns", ctypes.POINTER(PmhOutColumn)),
        ("stats", PmhExecStats),
    ]


# Every symbol include/pmh.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_U8P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "pmh_last_error": (ctypes.c_char_p, []),
    "pmh_version_string": (ctypes.c_char_p, []),
    "pmh_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "pmh_execute_groupby": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(PmhColumn),
                                           ctypes.c_uint32, ctypes.POINTER(ctypes.POINTER(PmhResult))]),
    "pmh_execute_filter_groupby": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(PmhColumn),
                                                  ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.POINTER(ctypes.POINTER(PmhResult))]),
    "pmh_execute_filter_groupby_noalloc": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(PmhColumn),
                                                          ctypes.c_uint32, ctypes.POINTER(PmhOutColumn), ctypes.c_uint32,
                                                          ctypes.POINTER(ctypes.c_uint64)]),
    "pmh_execute_filter_compact": (ctypes.c_int, [_U8P, _U8P, ctypes.c_uint64, ctypes.POINTER(PmhColumn),
                                                  ctypes.c_uint32, ctypes.POINTER(ctypes.POINTER(PmhResult))]),
    "pmh_cmp_i64_col_scalar": (ctypes.c_int, [_P, _U8P, ctypes.c_int64, ctypes.c_int, ctypes.c_uint64, _U8P, _U8P]),
    "pmh_cmp_f64_col_scalar": (ctypes.c_int, [_P, _U8P, ctypes.c_double, ctypes.c_int, ctypes.c_uint64, _U8P, _U8P]),
    "pmh_cmp_i64_col_col": (ctypes.c_int, [_P, _U8P, _P, _U8P, ctypes.c_int, ctypes.c_uint64, _U8P, _U8P]),
    "pmh_cmp_f64_col_col": (ctypes.c_int, [_P, _U8P, _P, _U8P, ctypes.c_int, ctypes.c_uint64, _U8P, _U8P]),
    "pmh_bool_and": (ctypes.c_int, [_U8P, _U8P, _U8P, _U8P, ctypes.c_uint64, _U8P, _U8P]),
    "pmh_bool_or": (ctypes.c_int, [_U8P, _U8P, _U8P, _U8P, ctypes.c_uint64, _U8P, _U8P]),
    "pmh_compute_lifting_plan": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "pmh_warmup_common_fused_signatures": (ctypes.c_int, []),
    "pmh_buffer_alloc": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "pmh_buffer_zeroed": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "pmh_buffer_from_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "pmh_buffer_to_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    "pmh_buffer_free": (ctypes.c_int, [ctypes.c_void_p]),
    "pmh_buffer_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    "pmh_result_free": (None, [ctypes.POINTER(PmhResult)]),
    "pmh_string_free": (None, [ctypes.c_void_p]),
    "pmh_set_profiling": (ctypes.c_int, [ctypes.c_int]),
    "pmh_kernel_times": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "pmh_sketch_features": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(PmhColumn),
                                           ctypes.c_uint32, _FP, ctypes.POINTER(PmhExecStats)]),
    "pmh_utf8_encode": (ctypes.c_int, [ctypes.POINTER(PmhUtf8Column), ctypes.POINTER(ctypes.c_void_p)]),
    "pmh_utf8_dict_codes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64)]),
    "pmh_utf8_decode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_uint64)]),
    "pmh_utf8_dict_free": (None, [ctypes.c_void_p]),
    "pmh_column_resident": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(PmhChunk), ctypes.c_uint32,
                                           ctypes.c_uint64, ctypes.POINTER(PmhResidentColumn)]),
    "pmh_column_release": (ctypes.c_int, [ctypes.c_uint64]),
    "pmh_column_evict": (ctypes.c_int, [ctypes.c_uint64]),
    "pmh_colcache_clear": (ctypes.c_int, []),
    "pmh_colcache_set_capacity": (ctypes.c_int, [ctypes.c_uint64]),
    "pmh_colcache_stats": (ctypes.c_int, [ctypes.POINTER(PmhColcacheInfo)]),
    "pmh_sketch_inspect": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(PmhColumn),
                                          ctypes.c_uint32, ctypes.POINTER(PmhSketchInfo)]),
    "pmh_mlp_new": (ctypes.c_int, [_FP, ctypes.POINTER(_P)]),
    "pmh_mlp_free": (None, [_P]),
    "pmh_mlp_params": (ctypes.c_int, [_P, _FP]),
    "pmh_mlp_forward": (ctypes.c_int, [_P, _FP, ctypes.c_uint32, _FP]),
    "pmh_mlp_grad": (ctypes.c_int, [_P, _FP, _FP, ctypes.c_uint32, ctypes.c_float, _FP, _FP]),
    "pmh_mlp_set_data": (ctypes.c_int, [_P, _FP, _FP, ctypes.c_uint32]),
    "pmh_mlp_train": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_uint32, _FP]),
    "pmh_sizer_install": (ctypes.c_int, [_FP]),
    "pmh_execute_batch": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(PmhBatchQuery), ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.POINTER(PmhResult)),
                                         ctypes.POINTER(PmhBatchStats)]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load libpmh.so once; raise ImportError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"polars_metal: {LIB_PATH} is not built; run __graft_entry__.build() "
            "(the MI355X backend has no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class ComputeError(RuntimeError):
    """Stand-in for polars.exceptions.ComputeError when Polars is absent."""


def _compute_error_cls():
    try:
        from polars.exceptions import ComputeError as PolarsComputeError  # type: ignore
        return PolarsComputeError
    except Exception:
        return ComputeError


def check(status: int) -> None:
    """Raise the reference's exception class for a non-zero status."""
    if status == PMH_OK:
        return
    msg = load().pmh_last_error().decode("utf-8", "replace")
    if status == PMH_E_VALUE:
        raise ValueError(msg)
    if status == PMH_E_KEY:
        raise KeyError(msg)
    if status == PMH_E_NOTIMPL:
        raise NotImplementedError(msg)
    if status == PMH_E_RUNTIME:
        raise RuntimeError(msg)
    raise _compute_error_cls()(msg)


def take_string(ptr: ctypes.c_void_p) -> str:
    s = ctypes.cast(ptr, ctypes.c_char_p).value.decode("utf-8")
    load().pmh_string_free(ptr)
    return s
