"""ctypes binding of libfpdb_amd.so (the C ABI declared in include/fpdb_amd.h).

The library is loaded from flexpushdowndb_amd/lib/.  There is no fallback: if
the HIP library is missing or cannot be loaded, importing the operators raises.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfpdb_amd.so")

FPDB_OK = 0
FPDB_ERR_INVALID = 1
FPDB_ERR_HIP = 2
FPDB_ERR_BUDGET = 3
FPDB_ERR_CAPACITY = 4
FPDB_ERR_INTERNAL = 5

TYPE_INT8, TYPE_INT16, TYPE_INT32, TYPE_INT64 = 1, 2, 3, 4
TYPE_UINT8, TYPE_UINT16, TYPE_UINT32, TYPE_UINT64 = 5, 6, 7, 8
TYPE_FLOAT32, TYPE_FLOAT64, TYPE_DATE32, TYPE_DATE64 = 9, 10, 11, 12
TYPE_UTF8 = 13

NP_TO_TYPE = {
    np.dtype(np.int8): TYPE_INT8, np.dtype(np.int16): TYPE_INT16,
    np.dtype(np.int32): TYPE_INT32, np.dtype(np.int64): TYPE_INT64,
    np.dtype(np.uint8): TYPE_UINT8, np.dtype(np.uint16): TYPE_UINT16,
    np.dtype(np.uint32): TYPE_UINT32, np.dtype(np.uint64): TYPE_UINT64,
    np.dtype(np.float32): TYPE_FLOAT32, np.dtype(np.float64): TYPE_FLOAT64,
}
TYPE_TO_NP = {v: k for k, v in NP_TO_TYPE.items()}
TYPE_TO_NP[TYPE_DATE32] = np.dtype(np.int32)
TYPE_TO_NP[TYPE_DATE64] = np.dtype(np.int64)

JOIN_INNER, JOIN_LEFT, JOIN_RIGHT, JOIN_FULL, JOIN_LEFT_SEMI, JOIN_RIGHT_SEMI = range(6)
AGG_COUNT_ALL, AGG_COUNT_VALID, AGG_SUM, AGG_MIN, AGG_MAX = range(5)


class Column(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("length", ctypes.c_int64), ("offset", ctypes.c_int64),
                ("values", ctypes.c_void_p), ("validity", ctypes.c_void_p),
                ("data", ctypes.c_void_p), ("data_bytes", ctypes.c_int64)]


class UniversalHash(ctypes.Structure):
    _fields_ = [("a", ctypes.c_int64), ("b", ctypes.c_int64),
                ("m", ctypes.c_int64), ("p", ctypes.c_int64)]


class AggSpec(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("input_type", ctypes.c_int32),
                ("input_nullable", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class StarDimension(ctypes.Structure):
    _fields_ = [("key", Column), ("filter_attr", Column), ("filter_op", ctypes.c_int32),
                ("has_group_attr", ctypes.c_int32), ("filter_constant", ctypes.c_int64),
                ("group_attr", Column), ("fact_fk", Column)]


class StarQuery(ctypes.Structure):
    _fields_ = [("num_dims", ctypes.c_int32), ("group_dims", ctypes.c_int32 * 2),
                ("filter_order", ctypes.c_int32 * 8), ("dims", StarDimension * 8),
                ("fact_measure", Column), ("fact_rows", ctypes.c_int64), ("num_partitions", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("partition_offsets", ctypes.POINTER(ctypes.c_int64)),
                ("max_groups", ctypes.c_int64), ("partition_fact_fk", ctypes.POINTER(Column)),
                ("partition_measure", ctypes.POINTER(Column))]


class StarTuning(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int32), ("barrier_timeout_us", ctypes.c_int32),
                ("ht_load_shift", ctypes.c_int32), ("items_per_thread", ctypes.c_int32)]


class ServeDimension(ctypes.Structure):
    _fields_ = [("key", Column), ("filter_attr", Column), ("group_attr", Column), ("group_domain", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("fact_fk", Column)]


class ServeSchema(ctypes.Structure):
    _fields_ = [("num_dims", ctypes.c_int32), ("reserved", ctypes.c_int32), ("dims", ServeDimension * 8),
                ("fact_measure", Column), ("fact_rows", ctypes.c_int64)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("max_active", ctypes.c_int32), ("chunks", ctypes.c_int32), ("workspace_bytes", ctypes.c_uint64)]


class ServeQuery(ctypes.Structure):
    _fields_ = [("num_filters", ctypes.c_int32), ("num_group", ctypes.c_int32), ("filter_dim", ctypes.c_int32 * 4),
                ("filter_op", ctypes.c_int32 * 4), ("filter_constant", ctypes.c_int64 * 4),
                ("group_dim", ctypes.c_int32 * 2)]


CMP_LT, CMP_LE, CMP_GT, CMP_GE, CMP_EQ, CMP_NE = range(6)
CMP_OR = 0x10


class FpdbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fpdb error {code}: {msg}")
        self.code = code


# (name, restype, argtypes) for every exported entry point
_P = ctypes.c_void_p
_I = ctypes.c_int
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PCOL = ctypes.POINTER(Column)
SIGNATURES = [
    ("fpdb_context_create", _I32, [_I, _U64, _P, ctypes.POINTER(_P)]),
    ("fpdb_context_destroy", _I32, [_P]),
    ("fpdb_context_stream", _P, [_P]),
    ("fpdb_context_synchronize", _I32, [_P]),
    ("fpdb_context_memory", _I32, [_P, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    ("fpdb_device_info", _I32, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_U64)]),
    ("fpdb_last_error", ctypes.c_char_p, []),
    ("fpdb_device_alloc", _I32, [_P, _U64, ctypes.POINTER(_P)]),
    ("fpdb_device_free", _I32, [_P, _P]),
    ("fpdb_copy_to_device", _I32, [_P, _P, _P, _U64]),
    ("fpdb_copy_to_host", _I32, [_P, _P, _P, _U64]),
    ("fpdb_copy_device", _I32, [_P, _P, _P, _U64]),
    ("fpdb_fill", _I32, [_P, _P, ctypes.c_uint8, _U64]),
    ("fpdb_timer_start", _I32, [_P]),
    ("fpdb_timer_stop", _I32, [_P, ctypes.POINTER(ctypes.c_float)]),
    ("fpdb_context_set_profiling", _I32, [_P, _I]),
    ("fpdb_profile_kernel", _I32, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), _PI64]),
    ("fpdb_profile_reset", _I32, [_P]),
    ("fpdb_hash_keys", _I32, [_P, _PCOL, _I, _I64, _P]),
    ("fpdb_bloom_masks", _I32, [_P]),
    ("fpdb_bloom_blocked_log_num_blocks", _I, [_I64]),
    ("fpdb_bloom_blocked_build", _I32, [_P, _PCOL, _I, _I64, _P, _I, _P]),
    ("fpdb_bloom_blocked_insert", _I32, [_P, _PCOL, _I, _I64, _P, _I, _P]),
    ("fpdb_bloom_blocked_find", _I32, [_P, _P, _I, _PCOL, _I, _I64, _P, _P]),
    ("fpdb_bloom_legacy_params", _I32, [_I64, ctypes.c_double, _PI64, _PI64]),
    ("fpdb_bloom_legacy_make_hash_functions", _I32, [ctypes.c_uint32, _I64, _I, ctypes.POINTER(UniversalHash)]),
    ("fpdb_bloom_legacy_build", _I32, [_P, _PCOL, _I, _I64, _P, ctypes.POINTER(UniversalHash), _I, _I64, _P]),
    ("fpdb_bloom_legacy_find", _I32, [_P, ctypes.POINTER(UniversalHash), _I, _I64, _I64, _P, _PCOL, _I, _I64, _P, _P]),
    ("fpdb_compare_bitmap", _I32, [_P, _PCOL, _I, _I64, ctypes.c_double, _P, _I64, _P, _P]),
    ("fpdb_filter_compact", _I32, [_P, _PCOL, _I, _I64, _P, _PCOL, _PI64]),
    ("fpdb_filter_compact_noalloc", _I32, [_P, _PCOL, _I, _I64, _P, _PCOL, _P]),
    ("fpdb_divide_to_float64", _I32, [_P, _PCOL, _PCOL, _I64, _PCOL]),
    ("fpdb_bloom_blocked_filter", _I32, [_P, _P, _I, _PCOL, _I, ctypes.POINTER(_I), _I, _I64, _P, _PCOL, _PI64]),
    ("fpdb_hj_build", _I32, [_P, _PCOL, _I, _I64, _PCOL, _I, ctypes.POINTER(_P)]),
    ("fpdb_hj_table_destroy", _I32, [_P]),
    ("fpdb_hj_table_info", _I32, [_P, _PI64, _PI64, _PI64, ctypes.POINTER(_I)]),
    ("fpdb_hj_table_form", _I32, [_P, ctypes.POINTER(_I)]),
    ("fpdb_hj_probe", _I32, [_P, _P, _I, _PCOL, _I, _I64, _PCOL, _I, _PCOL, _I, _PCOL, _I64, _PI64]),
    ("fpdb_hj_probe_async", _I32, [_P, _P, _I, _PCOL, _I, _I64, _PCOL, _I, _PCOL, _I, _PCOL, _I64, _P]),
    ("fpdb_hj_probe_count", _I32, [_P, _P, _I, _PCOL, _I, _I64, _PI64]),
    ("fpdb_hj_build_side_rows", _I32, [_P, _P, _I, _P, _I64, _PI64]),
    ("fpdb_gather", _I32, [_P, _PCOL, _I, _P, _I64, _PCOL]),
    ("fpdb_copy_rows", _I32, [_P, _PCOL, _I, _I64, _PCOL, _I64]),
    ("fpdb_shuffle", _I32, [_P, _PCOL, _I, _PCOL, _I, _I64, _I, _PCOL, _P, _PI64]),
    ("fpdb_bitmap_to_words", _I32, [_P, _P, _I64, _I64, _P]),
    ("fpdb_bitmap_from_words", _I32, [_P, _P, _I64, _I64, _I64, _P]),
    ("fpdb_bloom_blocked_export", _I32, [_P, _P, _I, _P, _P]),
    ("fpdb_bloom_blocked_import", _I32, [_P, _P, _P, _U64, ctypes.POINTER(_I), _P]),
    ("fpdb_cache_create", _I32, [_P, _U64, _I, ctypes.POINTER(_P)]),
    ("fpdb_cache_destroy", _I32, [_P]),
    ("fpdb_cache_store", _I32, [_P, ctypes.c_char_p, ctypes.c_char_p, _I64, _I64, _PCOL, ctypes.POINTER(_I)]),
    ("fpdb_cache_load", _I32, [_P, ctypes.c_char_p, ctypes.c_char_p, _I64, _I64, _PCOL, ctypes.POINTER(_I)]),
    ("fpdb_cache_release", _I32, [_P, _P]),
    ("fpdb_cache_remove", _I32, [_P, ctypes.c_char_p, ctypes.c_char_p, _PI64]),
    ("fpdb_cache_clear", _I32, [_P]),
    ("fpdb_cache_info", _I32, [_P, _PI64, ctypes.POINTER(_U64), ctypes.POINTER(_U64), _PI64, _PI64]),
    ("fpdb_cache_layout", _I32, [_P, ctypes.c_char_p, _U64, ctypes.POINTER(_U64)]),
    ("fpdb_groupby_create", _I32, [_P, ctypes.POINTER(_I32), _I, ctypes.POINTER(AggSpec), _I, _I64, ctypes.POINTER(_P)]),
    ("fpdb_groupby_create_packed", _I32, [_P, ctypes.POINTER(_I32), _I, _PI64, ctypes.POINTER(_I32),
                                          ctypes.POINTER(AggSpec), _I, _I64, ctypes.POINTER(_P)]),
    ("fpdb_groupby_destroy", _I32, [_P]),
    ("fpdb_column_min_max", _I32, [_P, _PCOL, _I64, _PI64, _PI64, _PI64]),
    ("fpdb_groupby_consume", _I32, [_P, _P, _PCOL, _PCOL, _I64]),
    ("fpdb_groupby_num_groups", _I32, [_P, _P, _PI64]),
    ("fpdb_groupby_output_type", _I32, [AggSpec]),
    ("fpdb_groupby_finalize", _I32, [_P, _P, _PCOL, _PCOL, _PI64]),
    ("fpdb_star_prepare", _I32, [_P, ctypes.POINTER(StarQuery), ctypes.POINTER(StarTuning), ctypes.POINTER(_P)]),
    ("fpdb_star_destroy", _I32, [_P]),
    ("fpdb_star_run_plain", _I32, [_P, _P]),
    ("fpdb_star_run_batched", _I32, [_P, _P]),
    ("fpdb_star_graph_capture", _I32, [_P, _P]),
    ("fpdb_star_graph_launch", _I32, [_P, _P]),
    ("fpdb_star_result", _I32, [_P, _P, _P, _P, _P, _P, _I64, _PI64]),
    ("fpdb_star_filter", _I32, [_P, _P, _I, _I, _P, _I64, ctypes.POINTER(_I), _PI64]),
    ("fpdb_serve_create", _I32, [_P, ctypes.POINTER(ServeSchema), ctypes.POINTER(ServeConfig), ctypes.POINTER(_P)]),
    ("fpdb_serve_destroy", _I32, [_P]),
    ("fpdb_serve_submit", _I32, [_P, ctypes.POINTER(ServeQuery), _PI64]),
    ("fpdb_serve_step", _I32, [_P, ctypes.POINTER(_I)]),
    ("fpdb_serve_status", _I32, [_P, _I64, ctypes.POINTER(_I)]),
    ("fpdb_serve_result", _I32, [_P, _I64, _P, _P, _P, _P, _I64, _PI64]),
    ("fpdb_serve_run_trace", _I32, [_P, ctypes.POINTER(ServeQuery), _PI64, _I64, _PI64, _PI64, _PI64]),
    ("fpdb_utf8_take_bytes", _I32, [_P, _PCOL, _PCOL, _I64, _PI64]),
    ("fpdb_utf8_take", _I32, [_P, _PCOL, _PCOL, _I64, _PCOL]),
    ("fpdb_iota_int64", _I32, [_P, _P, _I64, _I64]),
    ("fpdb_compare_bitmap_utf8", _I32, [_P, _PCOL, _I, ctypes.c_char_p, _I64, _P, _I64, _P, _P]),
    ("fpdb_keydict_create", _I32, [_P, ctypes.POINTER(_I32), _I, _I, ctypes.POINTER(_P)]),
    ("fpdb_keydict_destroy", _I32, [_P]),
    ("fpdb_keydict_size", _I32, [_P, _PI64]),
    ("fpdb_keydict_encode", _I32, [_P, _P, _PCOL, _I, _I64, _I, _I, _PCOL, _PI64]),
    ("fpdb_keydict_decode_bytes", _I32, [_P, _P, _I, _PCOL, _I64, _PI64]),
    ("fpdb_keydict_decode", _I32, [_P, _P, _I, _PCOL, _I64, _PCOL]),
]

_lib = None


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FpdbError(FPDB_ERR_INVALID,
                            f"{LIB_PATH} not built: run python -m flexpushdowndb_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != FPDB_OK:
        msg = lib().fpdb_last_error()
        raise FpdbError(status, msg.decode() if msg else "")
    return status


def columns_array(cols):
    arr = (Column * max(1, len(cols)))()
    for i, c in enumerate(cols):
        arr[i] = c
    return arr
