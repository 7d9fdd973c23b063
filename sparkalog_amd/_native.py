"""ctypes bindings of the native libraries (include/sparkalog_hip.h and
include/sparkalog_workload.h).

The HIP library is required: importing an operator without it raises, so no
operator ever falls back to a host implementation.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "_lib"
HIP_LIB = LIB_DIR / "libsparkalog_hip.so"
WORKLOAD_LIB = LIB_DIR / "libsparkalog_workload.so"

u32p = POINTER(c_uint32)
u64p = POINTER(c_uint64)


class Rule(ctypes.Structure):
    """sparkalog_hip_rule (RecursiveRulePlan)."""

    _fields_ = [
        ("target", c_uint32), ("delta_input", c_uint32), ("right_input", c_uint32),
        ("left_key", c_uint32), ("right_key", c_uint32),
        ("out_input", c_uint32 * 2), ("out_column", c_uint32 * 2),
        ("distinct_columns", c_uint32 * 2),
        ("anti_left", c_uint32 * 2), ("anti_right", c_uint32 * 2),
        ("union_left", c_uint32 * 2), ("union_right", c_uint32 * 2),
    ]


class Summary(ctypes.Structure):
    """sparkalog_hip_summary (SccSummary + device timing)."""

    _fields_ = [
        ("iterations", c_uint64), ("total_new_rows", c_uint64), ("delta_rows_left", c_uint64),
        ("device_ms", c_double), ("algorithmic_bytes", c_double),
    ]


# name -> (restype, argtypes)
_HIP_SIGNATURES = {
    "sparkalog_hip_add_one_i32": (c_int, [c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_stream_synchronize": (c_int, [c_void_p]),
    "sparkalog_hip_filter_u32_temporary_bytes": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, POINTER(c_size_t), c_void_p]),
    "sparkalog_hip_filter_u32": (c_int, [c_void_p, c_size_t, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_fill_mod_u32": (c_int, [c_void_p, c_size_t, c_uint32, c_void_p]),
    "sparkalog_hip_copy_u64": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_join_u32_temporary_bytes": (c_int, [c_void_p, c_void_p, c_size_t, POINTER(c_size_t), c_void_p]),
    "sparkalog_hip_join_u32_count": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_join_u32_count_no_alloc": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sparkalog_hip_join_u32_emit": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_uint32, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "sparkalog_hip_distinct_u32_temporary_bytes": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, POINTER(c_size_t), c_void_p]),
    "sparkalog_hip_distinct_u32": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_anti_join_u32_temporary_bytes": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t, POINTER(c_size_t), c_void_p]),
    "sparkalog_hip_anti_join_u32": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_union_u32_temporary_bytes": (c_int, [c_void_p, c_size_t, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, POINTER(c_size_t), c_void_p]),
    "sparkalog_hip_union_u32": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_range_index_u32_temporary_bytes": (c_int, [c_size_t, POINTER(c_size_t)]),
    "sparkalog_hip_range_index_u32": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_text_lines_temporary_bytes": (c_int, [c_size_t, POINTER(c_size_t)]),
    "sparkalog_hip_text_lines": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_load_binary_u32_temporary_bytes": (c_int, [c_size_t, c_uint64, POINTER(c_size_t)]),
    "sparkalog_hip_load_binary_u32": (c_int, [c_void_p, c_size_t, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "sparkalog_hip_store_create": (c_int, [POINTER(c_void_p), c_void_p]),
    "sparkalog_hip_store_destroy": (None, [c_void_p]),
    "sparkalog_hip_store_insert_static": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p, c_uint64, c_int]),
    "sparkalog_hip_store_insert_recursive": (c_int, [c_void_p, c_uint32, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_int]),
    "sparkalog_hip_store_clear": (c_int, [c_void_p]),
    "sparkalog_hip_store_rows": (c_int, [c_void_p, c_uint32, c_int, u64p]),
    "sparkalog_hip_store_read": (c_int, [c_void_p, c_uint32, c_int, c_void_p, c_void_p, c_uint64, c_int]),
    "sparkalog_hip_store_digest": (c_int, [c_void_p, c_uint32, u64p, POINTER(c_int), u64p]),
    "sparkalog_hip_store_set_profiling": (c_int, [c_void_p, c_int]),
    "sparkalog_hip_store_set_tuning": (c_int, [c_void_p, c_uint32, c_int, c_int, c_uint64]),
    "sparkalog_hip_store_set_option": (c_int, [c_void_p, ctypes.c_char_p, c_uint64]),
    "sparkalog_hip_store_graph_stats": (c_int, [c_void_p, u64p, u64p]),
    "sparkalog_hip_store_graph_capture_stats": (c_int, [c_void_p, u64p, ctypes.POINTER(ctypes.c_double)]),
    "sparkalog_hip_store_op_stats": (c_int, [c_void_p, c_int, u64p, POINTER(c_double), POINTER(c_double)]),
    "sparkalog_hip_store_op_kernel_stats": (c_int, [c_void_p, c_int, POINTER(ctypes.c_char_p), POINTER(ctypes.c_char_p),
                                                     u64p, POINTER(c_double)]),
    "sparkalog_hip_executor_compile": (c_int, [c_void_p, u32p, c_uint32, POINTER(Rule), c_uint32]),
    "sparkalog_hip_executor_run": (c_int, [c_void_p, u32p, c_uint32, POINTER(Rule), c_uint32, c_uint64, POINTER(Summary)]),
    "sparkalog_hip_session_begin": (c_int, [c_void_p, u32p, c_uint32, POINTER(Rule), c_uint32]),
    "sparkalog_hip_session_step": (c_int, [c_void_p, u64p]),
    "sparkalog_hip_session_end": (c_int, [c_void_p, POINTER(Summary)]),
    "sparkalog_hip_set_instance_tracking": (c_int, [c_void_p, c_uint32, c_uint32]),
    "sparkalog_hip_instance_iterations": (c_int, [c_void_p, c_void_p, c_uint32]),
    "sparkalog_hip_instance_delta_rows": (c_int, [c_void_p, c_void_p, c_uint32]),
    "sparkalog_hip_store_instance_stats": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p]),
    "sparkalog_hip_store_range_stats": (c_int, [c_void_p, c_uint32, c_int, c_uint32, c_uint32, c_uint32, u64p, u64p]),
    "sparkalog_hip_store_range_rows": (c_int, [c_void_p, c_uint32, c_int, c_void_p, c_uint32, c_void_p]),
    "sparkalog_hip_store_merge_rows": (c_int, [c_void_p, c_uint32, c_int, c_void_p, c_void_p, c_uint64, c_int]),
    "sparkalog_hip_store_remove_range": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32]),
    "sparkalog_hip_store_difference": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, u64p]),
    "sparkalog_hip_store_remove_ranges": (c_int, [c_void_p, c_uint32, u32p, c_uint32]),
    "sparkalog_hip_last_hip_error": (c_int, []),
    "sparkalog_hip_status_string": (ctypes.c_char_p, [c_int]),
}

_WORKLOAD_SIGNATURES = {
    "sparkalog_wl_splitmix64": (c_uint64, [u64p]),
    "sparkalog_wl_free": (None, [c_void_p]),
    "sparkalog_wl_random_dag": (c_int, [c_uint64, c_uint32, c_uint64, c_void_p, c_void_p]),
    "sparkalog_wl_edge_text": (c_uint64, [c_uint64, c_uint64, c_uint32, c_uint32, POINTER(c_void_p), u32p, u32p]),
    "sparkalog_wl_component_graph": (c_uint64, [c_uint64, c_uint32, c_double, c_uint32, c_uint32, c_double, POINTER(u32p), POINTER(u32p), u32p]),
    "sparkalog_wl_batch_graphs": (c_uint64, [c_uint64, c_uint32, c_uint32, c_uint32, c_double, c_double, POINTER(u32p), POINTER(u32p), POINTER(u64p), POINTER(u32p)]),
    "sparkalog_wl_pairs": (None, [c_uint64, c_uint64, c_int, c_uint32, c_double, c_void_p, c_void_p]),
    "sparkalog_wl_distinct_input": (None, [c_uint64, c_uint64, c_double, c_int, c_uint32, c_double, c_void_p, c_void_p]),
    "sparkalog_wl_overlapping_sets": (None, [c_uint64, c_uint64, c_double, c_int, c_uint32, c_double, c_void_p,
                                             c_void_p, c_void_p, c_void_p]),
    "sparkalog_wl_pair_digest": (c_uint64, [c_void_p, c_void_p, c_uint64]),
    "sparkalog_wl_packed_digest": (c_uint64, [c_void_p, c_uint64]),
    "sparkalog_wl_mix64": (c_uint64, [c_uint64]),
}

_hip = None
_workload = None


def _bind(lib: ctypes.CDLL, signatures: dict) -> ctypes.CDLL:
    for name, (restype, argtypes) in signatures.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def hip() -> ctypes.CDLL:
    """The HIP operator library; raises if it has not been built."""
    global _hip
    if _hip is None:
        if not HIP_LIB.exists():
            raise RuntimeError(
                f"{HIP_LIB} is missing: build it with `python -m sparkalog_amd.build`")
        _hip = _bind(ctypes.CDLL(str(HIP_LIB)), _HIP_SIGNATURES)
    return _hip


def workload() -> ctypes.CDLL:
    global _workload
    if _workload is None:
        if not WORKLOAD_LIB.exists():
            raise RuntimeError(
                f"{WORKLOAD_LIB} is missing: build it with `python -m sparkalog_amd.build`")
        _workload = _bind(ctypes.CDLL(str(WORKLOAD_LIB)), _WORKLOAD_SIGNATURES)
    return _workload


def hip_exported_symbols() -> list[str]:
    return list(_HIP_SIGNATURES)


def workload_exported_symbols() -> list[str]:
    return list(_WORKLOAD_SIGNATURES)
