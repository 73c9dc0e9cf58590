"""ctypes binding of libqs_amd.so (include/qs_amd.h).

The library is loaded from this package directory only; a missing library is a
hard error (there is no CPU fallback anywhere in the product).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# The library next to this file; QS_AMD_LIB names another build of it.
LIB_PATH = os.environ.get("QS_AMD_LIB") or os.path.join(HERE, "libqs_amd.so")

QS_MAX_DIMS = 16

QS_OK, QS_ERR_ARGUMENT, QS_ERR_DIMENSION, QS_ERR_BOUNDS, QS_ERR_OVERFLOW = 0, 1, 2, 3, 4
QS_ERR_DEVICE, QS_ERR_MEMORY, QS_ERR_UNSUPPORTED = 5, 6, 7

QS_F32, QS_F64, QS_C32, QS_C64 = 0, 1, 2, 3
QS_MEM_DEVICE, QS_MEM_HOST = 0, 1
QS_PLAIN, QS_SPLIT_INT8 = 0, 1
QS_FAULT_DEVICE_ALLOC, QS_FAULT_PINNED_ALLOC = 0, 1

PATH_NAMES = {0: "empty", 1: "scale", 2: "gemm", 3: "dot", 4: "outer", 5: "skinny", 6: "split"}


class QSGroup(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("_pad", C.c_int32),
        ("lengths", C.c_int64 * QS_MAX_DIMS),
        ("strides", (C.c_int64 * QS_MAX_DIMS) * 2),
    ]


class QSContractDesc(C.Structure):
    _fields_ = [
        ("a_type", C.c_int32), ("b_type", C.c_int32), ("c_type", C.c_int32),
        ("compute_type", C.c_int32),
        ("conj_a", C.c_int32), ("conj_b", C.c_int32),
        ("memory", C.c_int32), ("precision", C.c_int32),
        ("m", QSGroup), ("n", QSGroup), ("k", QSGroup),
        ("a", C.c_void_p), ("a_base", C.c_int64), ("a_len", C.c_int64),
        ("b", C.c_void_p), ("b_base", C.c_int64), ("b_len", C.c_int64),
        ("c", C.c_void_p), ("c_base", C.c_int64), ("c_len", C.c_int64),
        ("target_rel_err", C.c_double),
        ("tile", C.c_int32),
        ("path_mode", C.c_int32),
        ("k_splits", C.c_int32), ("_reserved", C.c_int32),
    ]


class QSPlanInfo(C.Structure):
    _fields_ = [
        ("path", C.c_int32), ("tile_m", C.c_int32), ("tile_n", C.c_int32), ("tile_k", C.c_int32),
        ("k_splits", C.c_int32), ("a_kfast", C.c_int32), ("b_kfast", C.c_int32), ("moduli", C.c_int32),
        ("m", C.c_int64), ("n", C.c_int64), ("k", C.c_int64), ("workspace_bytes", C.c_int64),
        ("paired_loads", C.c_int32), ("mixed_domain", C.c_int32),
    ]


QS_ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t)
QS_FREE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_size_t)


class QSAllocator(C.Structure):
    _fields_ = [("alloc", QS_ALLOC_FN), ("free", QS_FREE_FN), ("ctx", C.c_void_p)]


class QSTuneResult(C.Structure):
    _fields_ = [("tile", C.c_int32), ("k_splits", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double),
                ("candidates", C.c_int32), ("_pad", C.c_int32)]


class QSServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64), ("batches", C.c_int64),
                ("mean_batch", C.c_double), ("p50_ms", C.c_double), ("p90_ms", C.c_double), ("p99_ms", C.c_double),
                ("max_ms", C.c_double), ("mean_ms", C.c_double), ("span_s", C.c_double),
                ("throughput_rps", C.c_double), ("tflops", C.c_double), ("peak_bytes", C.c_int64),
                ("budget_bytes", C.c_int64)]


class QSErrorReport(C.Structure):
    _fields_ = [("moduli", C.c_int32), ("samples", C.c_int32), ("rel_fro", C.c_double), ("max_rel", C.c_double),
                ("target", C.c_double), ("exceeded", C.c_int32), ("escalations", C.c_int32),
                ("fallback", C.c_int32), ("_pad", C.c_int32)]


class QSTargetProfile(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("compute_units", C.c_int32), ("wavefront_size", C.c_int32),
        ("lds_bytes_per_cu", C.c_int32), ("lds_bytes_per_block", C.c_int32),
        ("clock_khz", C.c_int32), ("memory_clock_khz", C.c_int32), ("memory_bus_bits", C.c_int32),
        ("l2_bytes", C.c_int32), ("hbm_bytes", C.c_int64), ("xcds", C.c_int32),
        ("mfma_f64_m", C.c_int32), ("mfma_f64_n", C.c_int32), ("mfma_f64_k", C.c_int32),
        ("mfma_f32_m", C.c_int32), ("mfma_f32_n", C.c_int32), ("mfma_f32_k", C.c_int32),
        ("mfma_i8_m", C.c_int32), ("mfma_i8_n", C.c_int32), ("mfma_i8_k", C.c_int32),
        ("peak_fp64_tflops", C.c_double), ("peak_fp32_tflops", C.c_double),
        ("peak_int8_tops", C.c_double), ("peak_hbm_gbs", C.c_double),
    ]


class QSGroupTuneResult(C.Structure):
    _fields_ = [("classes", C.c_int32), ("candidates", C.c_int32), ("members_changed", C.c_int32),
                ("_pad", C.c_int32), ("before_ms", C.c_double), ("after_ms", C.c_double)]


class QSAllocStats(C.Structure):
    _fields_ = [("device_allocs", C.c_int64), ("device_frees", C.c_int64), ("device_bytes", C.c_int64),
                ("pinned_allocs", C.c_int64), ("pinned_frees", C.c_int64),
                ("bridge_allocs", C.c_int64), ("bridge_frees", C.c_int64)]


class QSAddDesc(C.Structure):
    _fields_ = [
        ("a_type", C.c_int32), ("c_type", C.c_int32), ("compute_type", C.c_int32), ("conj_a", C.c_int32),
        ("m", QSGroup), ("n", QSGroup), ("k", QSGroup),
        ("a", C.c_void_p), ("a_base", C.c_int64), ("a_len", C.c_int64),
        ("c", C.c_void_p), ("c_base", C.c_int64), ("c_len", C.c_int64),
    ]


# Every symbol include/qs_amd.h declares: name -> (restype, argtypes).
SIGNATURES = {
    "qs_last_error": (C.c_char_p, []),
    "qs_version": (C.c_char_p, []),
    "qs_target_profile_get": (C.c_int, [C.POINTER(QSTargetProfile)]),
    "qs_target_measure": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p]),
    "qs_plan_create": (C.c_int, [C.POINTER(QSContractDesc), C.POINTER(C.c_void_p)]),
    "qs_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qs_plan_get_info": (C.c_int, [C.c_void_p, C.POINTER(QSPlanInfo)]),
    "qs_plan_create_with_allocator": (C.c_int, [C.c_void_p, C.POINTER(QSAllocator), C.POINTER(C.c_void_p)]),
    "qs_plan_release": (C.c_int, [C.c_void_p]),
    "qs_pool_allocator_create": (C.c_int, [C.c_int, C.c_uint64, C.c_void_p]),
    "qs_pool_allocator_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_int64)]),
    "qs_pool_allocator_destroy": (C.c_int, [C.c_void_p]),
    "qs_plan_rebind": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                 C.c_void_p, C.c_int64, C.c_int64]),
    "qs_plan_destroy": (C.c_int, [C.c_void_p]),
    "qs_add_plan_create": (C.c_int, [C.POINTER(QSAddDesc), C.POINTER(C.c_void_p)]),
    "qs_add_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "qs_add_plan_destroy": (C.c_int, [C.c_void_p]),
    "qs_add_plan_rebind": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int64]),
    "qs_plan_set_error_report": (C.c_int, [C.c_void_p, C.c_int32]),
    "qs_plan_get_error_report": (C.c_int, [C.c_void_p, C.POINTER(QSErrorReport)]),
    "qs_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(QSTuneResult)]),
    "qs_autotune_save": (C.c_int, [C.c_char_p]),
    "qs_autotune_load": (C.c_int, [C.c_char_p]),
    "qs_autotune_clear": (C.c_int, []),
    "qs_group_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_void_p)]),
    "qs_group_create_batched": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_void_p)]),
    "qs_group_execute": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p]),
    "qs_group_flops": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "qs_group_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "qs_group_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(QSGroupTuneResult)]),
    "qs_group_destroy": (C.c_int, [C.c_void_p]),
    "qs_group_capture": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_void_p,
                                   C.POINTER(C.c_void_p)]),
    "qs_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qs_graph_destroy": (C.c_int, [C.c_void_p]),
    "qs_server_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "qs_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                   C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "qs_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "qs_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
    "qs_server_drain": (C.c_int, [C.c_void_p]),
    "qs_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(QSServerStats)]),
    "qs_server_reset_stats": (C.c_int, [C.c_void_p]),
    "qs_server_destroy": (C.c_int, [C.c_void_p]),
    "qs_fill_offsets": (C.c_int, [C.POINTER(QSGroup), C.c_int64, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]),
    "qs_alloc_stats_get": (C.c_int, [C.POINTER(QSAllocStats)]),
    "qs_fault_inject": (C.c_int, [C.c_int32, C.c_int64]),
    "qs_contract_noalloc": (C.c_int, [C.POINTER(QSContractDesc), C.c_void_p, C.c_void_p, C.c_void_p]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libqs_amd.so not found at {LIB_PATH}: build it first "
                "(python quasistrided.jl_amd/build.py or __graft_entry__.build())")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().qs_last_error().decode()


def make_group(lengths, strides0, strides1) -> QSGroup:
    g = QSGroup()
    D = len(lengths)
    if D > QS_MAX_DIMS:
        raise ValueError(f"a group of {D} dims exceeds the library limit of {QS_MAX_DIMS}")
    g.ndim = D
    for d in range(D):
        g.lengths[d] = int(lengths[d])
        g.strides[0][d] = int(strides0[d])
        g.strides[1][d] = int(strides1[d])
    return g
