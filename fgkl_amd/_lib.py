"""ctypes binding of the engine's C-ABI (include/fgkl.h).

The library is built in-tree (``python -m fgkl_amd.build``) and loaded from
``fgkl_amd/libfgkl.so``; there is no fallback implementation, so a missing library is an error.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfgkl.so"

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class RunStats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("primary_ms", C.c_double),
        ("rescue_ms", C.c_double),
        ("rescued_pairs", C.c_int64),
        ("bucket_ms", C.c_double * 8),
        ("bucket_cells", C.c_int64 * 8),
        ("bucket_work", C.c_int64 * 8),
        ("dense_rescue_ms", C.c_double),
    ]


class TuneReport(C.Structure):
    _fields_ = [
        ("rows", C.c_int32 * 8),
        ("block", C.c_int32 * 8),
        ("boost", C.c_int32 * 8),
        ("default_ms", C.c_double * 8),
        ("tuned_ms", C.c_double * 8),
        ("candidates", C.c_int32),
        ("seconds", C.c_double),
    ]


class SwStats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("bucket_ms", C.c_double * 8),
        ("bucket_pairs", C.c_int64 * 8),
        ("bucket_cells", C.c_int64 * 8),
        ("substring_pairs", C.c_int64),
    ]


# name -> (restype, argtypes); every symbol include/fgkl.h declares.
SIGNATURES = {
    "fgkl_last_error": (C.c_char_p, []),
    "fgkl_backend_name": (C.c_char_p, []),
    "fgkl_device_count": (C.c_int, []),
    "fgkl_pairhmm_compute_native": (
        C.c_int,
        [u8p, C.c_int64] * 5 + [i32p, C.c_int64, u8p, C.c_int64, i32p, C.c_int64, C.c_int,
                                 f64p, C.c_int64],
    ),
    "fgkl_pdhmm_compute_native": (
        C.c_int,
        [u8p, C.c_int64] * 5 + [i32p, C.c_int64, u8p, C.c_int64, u8p, C.c_int64, i32p,
                                 C.c_int64, C.c_int, f64p, C.c_int64],
    ),
    "fgkl_pairhmm_compute_native_noalloc": (
        C.c_int,
        [u8p, C.c_int64] * 5 + [i32p, C.c_int64, u8p, C.c_int64, i32p, C.c_int64, C.c_int,
                                 f64p, C.c_int64],
    ),
    "fgkl_pairhmm_new": (C.c_void_p, [C.c_int, C.c_int, C.c_int]),
    "fgkl_pairhmm_free": (None, [C.c_void_p]),
    "fgkl_pairhmm_fallback_pairs": (C.c_uint64, [C.c_void_p]),
    "fgkl_pairhmm_compute": (
        C.c_int,
        [C.c_void_p, C.c_int64, i64p, u8p, u8p, u8p, u8p, u8p, C.c_int64, i64p, u8p, u8p,
         f64p, C.c_int64],
    ),
    "fgkl_batch_create": (
        C.c_int,
        [C.POINTER(C.c_void_p), C.c_int, C.c_int64, i64p, i64p, i64p, u8p, u8p, u8p, u8p, u8p,
         i64p, u8p, u8p],
    ),
    "fgkl_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(RunStats)]),
    "fgkl_batch_run_pairs": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(RunStats)]),
    "fgkl_batch_results": (C.c_int, [C.c_void_p, f64p, C.c_int64]),
    "fgkl_batch_pairs": (C.c_int64, [C.c_void_p]),
    "fgkl_batch_work": (C.c_int, [C.c_void_p, i64p]),
    "fgkl_batch_set_lds_records": (C.c_int, [C.c_void_p, C.c_int]),
    "fgkl_batch_set_uniform_gaps": (C.c_int, [C.c_void_p, C.c_int]),
    "fgkl_batch_snapshots": (C.c_int, [C.c_void_p, i64p]),
    "fgkl_plan_snapshots": (C.c_int, [C.c_int, C.c_int64, i64p, i64p, i64p, i64p, u8p, u8p,
                                      i64p]),
    "fgkl_batch_tune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TuneReport)]),
    "fgkl_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "fgkl_batch_get_launch": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fgkl_batch_free": (None, [C.c_void_p]),
    "fgkl_synth_region": (C.c_void_p, [C.c_uint64, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                       C.c_int]),
    "fgkl_synth_workload": (C.c_void_p, [C.c_int, C.c_uint64, C.c_int64]),
    "fgkl_synth_sizes": (None, [C.c_void_p, i64p, i64p, i64p, i64p, i64p]),
    "fgkl_synth_copy": (None, [C.c_void_p, i64p, i64p, i64p, u8p, u8p, u8p, u8p, u8p, i64p,
                               u8p, u8p]),
    "fgkl_synth_free": (None, [C.c_void_p]),
    "fgkl_replay_load": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, C.POINTER(C.c_void_p)]),
    "fgkl_synth_expected": (C.c_int64, [C.c_void_p, f64p, C.c_int64]),
    "fgkl_replay_stats": (None, [C.c_void_p, i64p, i64p, i64p]),
    "fgkl_replay_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, i64p, i64p, i64p, u8p, u8p,
                                    u8p, u8p, u8p, i64p, u8p, u8p, f64p, C.c_int64]),
    "fgkl_synth_splitmix": (None, [C.c_uint64, C.c_int64, u64p]),
    "fgkl_measure_valu_rate": (C.c_int, [C.c_int, f64p, f64p]),
    "fgkl_selftest_dpp": (C.c_int, [i32p]),
    "fgkl_selftest_dpp_shl": (C.c_int, [i32p]),
    "fgkl_grad_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64, i64p, i64p, i64p, u8p, i64p,
                                   u8p]),
    "fgkl_grad_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "fgkl_grad_pairs": (C.c_int64, [C.c_void_p]),
    "fgkl_grad_read_bases": (C.c_int64, [C.c_void_p]),
    "fgkl_grad_forward": (C.c_int, [C.c_void_p, f64p, f64p, f64p, C.c_double, f64p, f64p]),
    "fgkl_grad_backward": (C.c_int, [C.c_void_p, f64p, f64p, f64p, C.c_double, f64p, f64p, f64p,
                                     f64p, f64p, f64p, f64p]),
    "fgkl_grad_free": (None, [C.c_void_p]),
    "fgkl_aggregate_config": (C.c_int, [C.c_int, C.c_int]),
    "fgkl_aggregate_stats": (None, [i64p, i64p]),
    "fgkl_release_idle": (C.c_int64, []),
    "fgkl_pool_stats": (C.c_int, [C.c_int, i64p, i64p, i64p]),
    "fgkl_selftest_pool": (C.c_int, [C.c_int, C.c_int, i64p]),
    "fgkl_sw_backend_name": (C.c_char_p, []),
    "fgkl_sw_align_native": (C.c_int, [u8p, C.c_int64, u8p, C.c_int64, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int32, C.c_int32, i32p, C.c_char_p,
                                       C.c_int64]),
    "fgkl_sw_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int64, i64p, u8p, i64p, u8p]),
    "fgkl_sw_batch_run": (C.c_int, [C.c_void_p, i32p, C.c_int, C.POINTER(SwStats)]),
    "fgkl_sw_batch_results": (C.c_int, [C.c_void_p, i32p, i64p, C.POINTER(C.c_uint32),
                                        C.c_int64]),
    "fgkl_sw_batch_pairs": (C.c_int64, [C.c_void_p]),
    "fgkl_sw_batch_cells": (C.c_int64, [C.c_void_p]),
    "fgkl_sw_batch_cigar_bound": (C.c_int64, [C.c_void_p]),
    "fgkl_sw_batch_free": (None, [C.c_void_p]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"fgkl engine library {LIB_PATH} is not built; run `python -m fgkl_amd.build`"
            )
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().fgkl_last_error()
    return msg.decode() if msg else ""
