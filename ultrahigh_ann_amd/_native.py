"""ctypes binding of the engine's C ABI (include/uhann.h).

The shared library is built in-tree (``ultrahigh_ann_amd/lib/libuhann.so``)
by ``__graft_entry__.build()`` or ``make -C ultrahigh_ann_amd/csrc``. There is
no fallback: if the library is missing every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int64, c_size_t, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libuhann.so")

# Status codes of include/uhann.h and the Python exceptions standing in for
# the reference's C++ exception types.
OK = 0


class UhannError(RuntimeError):
    """std::runtime_error: device errors, no device."""


class UhannInvalidArgument(ValueError):
    """std::invalid_argument."""


class UhannLengthError(ValueError):
    """std::length_error."""


class UhannOverflowError(ArithmeticError):
    """std::overflow_error: non-finite results."""


class UhannDomainError(ValueError):
    """std::domain_error: invalid probabilities."""


class UhannOutOfRange(IndexError):
    """std::out_of_range."""


class UhannLogicError(RuntimeError):
    """std::logic_error."""


_STATUS = {
    1: UhannInvalidArgument,
    2: UhannLengthError,
    3: UhannError,
    4: UhannOverflowError,
    5: UhannDomainError,
    6: UhannOutOfRange,
    7: UhannLogicError,
}

class ServerConfig(ctypes.Structure):
    """uhann_server_config."""

    _fields_ = [("device", c_int), ("hbm_budget_bytes", c_uint64), ("max_tick_queries", ctypes.c_uint32),
                ("max_tick_requests", ctypes.c_uint32), ("scratch_floats", c_uint64), ("staging_floats", c_uint64),
                ("use_graph", c_int), ("pairs_tq", c_int)]


class IndexSpec(ctypes.Structure):
    """uhann_index_spec."""

    _fields_ = [("kind", c_int), ("metric", c_int), ("strategy", c_int), ("host_refs", c_void_p),
                ("rows", c_uint64), ("dim", c_uint64), ("sample", c_void_p), ("sample_count", c_uint64),
                ("projection", c_void_p), ("projection_dim", c_uint64)]


class ServingStats(ctypes.Structure):
    """uhann_serving_stats."""

    _fields_ = [("requests", c_uint64), ("queries", c_uint64), ("ticks", c_uint64), ("p50_ms", c_double),
                ("p99_ms", c_double), ("mean_ms", c_double), ("max_ms", c_double), ("wall_s", c_double),
                ("queries_per_s", c_double), ("gpu_busy_s", c_double), ("mean_tick_queries", c_double),
                ("resident_indexes", c_uint64), ("pending_indexes", c_uint64), ("resident_bytes", c_uint64)]


class TraceEntry(ctypes.Structure):
    """uhann_trace_entry."""

    _fields_ = [("arrival_s", c_double), ("index_id", c_uint64), ("queries", c_void_p), ("ld_queries", c_uint64),
                ("count", c_uint64), ("on_device", c_int), ("out_idx", c_void_p), ("out_dist", c_void_p)]


_lib = None


def _declare(lib):
    u64p = POINTER(c_uint64)
    f32p = POINTER(c_float)
    f64p = POINTER(c_double)
    sigs = {
        "uhann_last_error": (c_char_p, []),
        "uhann_device_available": (c_int, []),
        "uhann_device_name": (c_int, [c_int, ctypes.c_char_p, c_size_t]),
        "uhann_scan_create": (c_int, [c_int, f32p, c_uint64, c_uint64, f32p, c_int, POINTER(c_void_p)]),
        "uhann_scan_destroy": (None, [c_void_p]),
        "uhann_scan_info": (c_int, [c_void_p, u64p, u64p, u64p, POINTER(c_int)]),
        "uhann_scan_workspace_create": (c_int, [c_void_p, c_uint64, POINTER(c_void_p)]),
        "uhann_workspace_destroy": (None, [c_void_p]),
        "uhann_workspace_info": (c_int, [c_void_p, u64p, u64p]),
        "uhann_scan_query_batch": (c_int, [c_void_p, c_void_p, f32p, c_uint64, u64p, f32p, c_int]),
        "uhann_scan_query_device_batch": (
            c_int,
            [c_void_p, c_void_p, c_void_p, c_uint64, u64p, f32p, c_int],
        ),
        "uhann_scan_query_device_async": (
            c_int,
            [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_void_p, c_int, c_void_p],
        ),
        "uhann_scan_query_device_tile": (
            c_int,
            [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_void_p, c_int, c_int,
             c_int, c_int, c_void_p],
        ),
        "uhann_scan_enable_mixed": (c_int, [c_void_p, c_int]),
        "uhann_scan_set_fp8_shortlist": (c_int, [c_void_p, ctypes.c_uint32]),
        "uhann_workspace_mixed_stats": (c_int, [c_void_p, u64p, u64p, u64p]),
        "uhann_workspace_set_timing": (c_int, [c_void_p, c_int]),
        "uhann_workspace_stage_times": (c_int, [c_void_p, f32p, c_int]),
        "uhann_exact_create": (c_int, [c_int, f32p, c_uint64, c_uint64, c_int, POINTER(c_void_p)]),
        "uhann_flat_create": (c_int, [c_int, f32p, c_uint64, c_uint64, c_void_p, c_uint64, c_int,
                                      POINTER(c_void_p)]),
        "uhann_hierarchical_create": (c_int, [c_int, f32p, c_uint64, c_uint64, c_void_p, c_uint64, f64p,
                                              c_uint64, c_int, POINTER(c_void_p)]),
        "uhann_index_destroy": (None, [c_void_p]),
        "uhann_index_info": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), u64p, u64p, u64p, u64p, u64p]),
        "uhann_index_enable_mixed": (c_int, [c_void_p, c_int]),
        "uhann_index_workspace_create": (c_int, [c_void_p, c_uint64, POINTER(c_void_p)]),
        "uhann_index_query_batch": (c_int, [c_void_p, c_void_p, f32p, c_uint64, u64p, f32p, c_int]),
        "uhann_index_query_device_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, u64p, f32p, c_int]),
        "uhann_index_query_device_async": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_void_p,
                                                   c_void_p, c_int, c_void_p]),
        "uhann_workspace_status": (c_int, [c_void_p]),
        "uhann_rng_create": (c_int, [c_uint64, POINTER(c_void_p)]),
        "uhann_rng_destroy": (None, [c_void_p]),
        "uhann_rng_equal": (c_int, [c_void_p, c_void_p, POINTER(c_int)]),
        "uhann_rng_next": (c_int, [c_void_p, u64p]),
        "uhann_build_coordinate_sample": (c_int, [c_void_p, f64p, c_uint64, c_uint64, c_void_p, c_uint64, u64p]),
        "uhann_build_projection": (c_int, [c_void_p, c_int, c_void_p, c_uint64, c_uint64, f64p]),
        "uhann_importance_probabilities": (c_int, [c_int, f32p, c_uint64, c_uint64, f64p]),
        "uhann_sampling_mass": (c_int, [f64p, c_uint64, f64p]),
        "uhann_sha256_file": (c_int, [c_char_p, c_void_p]),
        "uhann_scan_distance_table": (c_int, [c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_uint64, c_void_p]),
        "uhann_save_probabilities": (c_int, [c_char_p, c_int, c_uint64, c_void_p, f64p, c_uint64]),
        "uhann_load_probabilities": (c_int, [c_char_p, POINTER(c_int), u64p, c_void_p, f64p, c_uint64, u64p]),
        "uhann_uniform_probabilities": (c_int, [c_uint64, c_double, f64p]),
        "uhann_importance_probabilities_device": (c_int, [c_int, f32p, c_uint64, c_uint64, c_int, f64p]),
        "uhann_scan_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, POINTER(c_int), c_int,
                                        POINTER(c_int)]),
        "uhann_index_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, POINTER(c_int), c_int,
                                         POINTER(c_int)]),
        "uhann_scan_gram_estimates": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_uint64,
                                              c_int, c_void_p]),
        "uhann_scan_fp8_estimates": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_uint64,
                                             c_void_p]),
        "uhann_workspace_shortlist_stats": (c_int, [c_void_p, u64p]),
        "uhann_scan_int8_estimates": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_uint64,
                                              c_int, c_void_p]),
        "uhann_int8_rows": (c_int, [c_void_p, c_uint64, c_uint64, c_uint64, c_void_p, c_void_p, c_void_p]),
        "uhann_scan_query_device_noalloc": (c_int, [c_void_p, c_void_p, c_uint64, c_uint64, c_void_p, c_void_p,
                                                    c_void_p]),
        "uhann_projection_create": (c_int, [f32p, c_uint64, c_uint64, c_int, POINTER(c_void_p)]),
        "uhann_projection_destroy": (None, [c_void_p]),
        "uhann_projection_workspace_create": (c_int, [c_void_p, c_uint64, POINTER(c_void_p)]),
        "uhann_projection_project": (c_int, [c_void_p, c_void_p, f32p, c_uint64, POINTER(c_void_p)]),
        "uhann_projection_info": (c_int, [c_void_p, u64p, u64p, u64p, POINTER(c_int)]),
        "uhann_median_create": (c_int, [f32p, c_uint64, c_uint64, c_int, POINTER(c_void_p)]),
        "uhann_median_destroy": (None, [c_void_p]),
        "uhann_median_workspace_create": (c_int, [c_void_p, c_uint64, POINTER(c_void_p)]),
        "uhann_median_query_device_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_uint64, u64p]),
        "uhann_median_info": (c_int, [c_void_p, u64p, u64p, u64p, POINTER(c_int)]),
        "uhann_device_properties": (c_int, [c_int, c_char_p, ctypes.c_size_t, POINTER(c_int), POINTER(c_int), u64p,
                                            POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "uhann_server_create": (c_int, [POINTER(ServerConfig), POINTER(c_void_p)]),
        "uhann_server_destroy": (None, [c_void_p]),
        "uhann_server_register": (c_int, [c_void_p, POINTER(IndexSpec), u64p]),
        "uhann_server_index_state": (c_int, [c_void_p, c_uint64, POINTER(c_int), u64p]),
        "uhann_server_remove": (c_int, [c_void_p, c_uint64]),
        "uhann_server_submit": (c_int, [c_void_p, c_uint64, c_void_p, c_int, c_uint64, c_uint64, c_void_p, c_void_p,
                                        u64p]),
        "uhann_server_wait": (c_int, [c_void_p, c_uint64, c_double, POINTER(c_int), POINTER(c_int)]),
        "uhann_server_run_trace": (c_int, [c_void_p, POINTER(TraceEntry), c_uint64, u64p]),
        "uhann_server_drain": (c_int, [c_void_p]),
        "uhann_server_stats": (c_int, [c_void_p, POINTER(ServingStats)]),
        "uhann_server_reset_stats": (c_int, [c_void_p]),
        "uhann_server_autotune": (c_int, [c_void_p, c_void_p, c_uint64, ctypes.c_uint32, POINTER(c_int)]),
    }
    optional = {}
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in optional.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
    _ = (f64p, c_int64)
    return lib


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UhannError(
                f"HIP engine library not built: {LIB_PATH} is missing "
                "(run __graft_entry__.build() or make -C ultrahigh_ann_amd/csrc)"
            )
        # torch carries its own HIP runtime. When both it and the engine's
        # (/opt/rocm) live in one process, torch's must initialise first:
        # started second it finds no device. Initialise it before loading
        # the engine, whatever order the caller imports things in.
        try:
            import torch

            torch.cuda.is_available()
        except ImportError:
            pass
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def check(status: int) -> None:
    if status != OK:
        msg = lib().uhann_last_error().decode(errors="replace")
        raise _STATUS.get(status, UhannError)(msg)


def f32ptr(arr):
    return arr.ctypes.data_as(POINTER(c_float)) if arr is not None else None


def f64ptr(arr):
    return arr.ctypes.data_as(POINTER(c_double)) if arr is not None else None


def u64ptr(arr):
    return arr.ctypes.data_as(POINTER(c_uint64)) if arr is not None else None
