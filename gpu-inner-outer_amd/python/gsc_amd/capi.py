"""ctypes mirror of include/gsc_amd.h and the loader of libgsc_amd.so.

The library is the product: if it is missing the import fails loudly — there
is no Python or CPU fallback for any computation.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char, c_char_p, c_double, c_int, c_int32, c_size_t, c_uint32, c_uint64, c_void_p

PKG_DIR = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(PKG_DIR, "libgsc_amd.so")

GSC_OK = 0
GSC_ERR_INVALID_ARGUMENT = 1
GSC_ERR_CONFIG = 2
GSC_ERR_NO_DEVICE = 3
GSC_ERR_DEVICE = 4
GSC_ERR_OUT_OF_MEMORY = 5
GSC_ERR_PARSE = 6
GSC_ERR_BUSY = 7

INVALID_INPUT = 0xFFFFFFFF
UNREACHABLE_STEP = 0xFFFFFFFF
MESSAGE_LEN = 256

CONSTRAINT_NONE = 0
CONSTRAINT_HYPERBOLIC = 1
CONSTRAINT_ELLIPTIC = 2


class Rect4(ctypes.Structure):
    _fields_ = [("lb", c_double * 4), ("ub", c_double * 4)]


class CaseDesc(ctypes.Structure):
    _fields_ = [
        ("name", c_char_p),
        ("state_lb", c_double * 4),
        ("state_ub", c_double * 4),
        ("state_eta", c_double * 4),
        ("wrap_dim", c_int32),
        ("input_lb", c_double * 2),
        ("input_ub", c_double * 2),
        ("input_eta", c_double * 2),
        ("candidate", Rect4),
        ("map", Rect4),
        ("obstacles", POINTER(Rect4)),
        ("obstacle_count", c_uint32),
        ("constraint_type", c_int32),
        ("hyperbolic_a", c_double),
        ("hyperbolic_b", c_double),
        ("hyperbolic_c", c_double),
        ("elliptic_a", c_double),
        ("elliptic_b", c_double),
        ("disturbance_half_width", c_double * 4),
        ("sample_time", c_double),
        ("integration_substeps", c_uint32),
        ("max_iterations", c_uint32),
        ("verbose", c_int32),
        ("extra_targets", POINTER(Rect4)),
        ("extra_target_count", c_uint32),
        ("target_band_enabled", c_int32),
        ("target_band_c", c_double),
        ("target_band_w", c_double),
        ("state_allow", POINTER(ctypes.c_uint8)),
    ]


class CaseShape(ctypes.Structure):
    _fields_ = [
        ("state_shape", c_uint32 * 4),
        ("state_strides", c_uint64 * 4),
        ("state_count", c_uint64),
        ("input_shape", c_uint32 * 2),
        ("input_strides", c_uint64 * 2),
        ("input_count", c_uint64),
        ("pair_count", c_uint64),
        ("prefix_shape", c_uint32 * 4),
        ("prefix_count", c_uint64),
    ]


class IterationStats(ctypes.Structure):
    _fields_ = [
        ("iteration", c_uint32),
        ("reserved", c_uint32),
        ("newly_reachable", c_uint64),
        ("newly_certified", c_uint64),
        ("total_reachable", c_uint64),
        ("total_certified", c_uint64),
        ("iteration_ms", c_double),
        ("prefix_build_ms", c_double),
        ("satisfaction_check_ms", c_double),
        ("reduction_ms", c_double),
    ]


KERNEL_TIMING_FIELDS = [
    "abstraction_kernel_ms",
    "abstraction_memcpy_h2d_ms",
    "scatter_mask_ms",
    "scan_axis0_ms",
    "scan_axis1_ms",
    "scan_axis2_ms",
    "scan_axis3_ms",
    "prefix_build_total_ms",
    "pair_satisfaction_ms",
    "reduce_inputs_ms",
    "iteration_memcpy_ms",
    "result_memcpy_d2h_ms",
    "total_kernel_ms",
    "total_memcpy_ms",
    "total_compute_ms",
]


class KernelTimings(ctypes.Structure):
    _fields_ = [(f, c_double) for f in KERNEL_TIMING_FIELDS]


class MemoryStats(ctypes.Structure):
    _fields_ = [
        ("total_allocated_bytes", c_uint64),
        ("peak_usage_bytes", c_uint64),
        ("abstraction_bytes", c_uint64),
        ("prefix_bytes", c_uint64),
        ("iteration_bytes", c_uint64),
        ("result_bytes", c_uint64),
    ]


BUDGET_FIELDS = [
    "pair_min_bytes",
    "pair_max_bytes",
    "pair_valid_bytes",
    "pair_satisfied_bytes",
    "reachable_prev_bytes",
    "reachable_next_bytes",
    "valid_mask_bytes",
    "candidate_mask_bytes",
    "controller_bytes",
    "reach_step_bytes",
    "prefix_reachable_bytes",
    "prefix_valid_bytes",
    "total_bytes",
]


class MemoryBudget(ctypes.Structure):
    _fields_ = [(f, c_uint64) for f in BUDGET_FIELDS]


class RunReport(ctypes.Structure):
    _fields_ = [
        ("compiled_with_gpu", c_int32),
        ("runtime_available", c_int32),
        ("executed", c_int32),
        ("converged", c_int32),
        ("message", c_char * MESSAGE_LEN),
        ("result_message", c_char * MESSAGE_LEN),
        ("abstraction_ms", c_double),
        ("solve_ms", c_double),
        ("total_ms", c_double),
        ("iterations", c_uint32),
        ("iteration_stats_written", c_uint32),
        ("reachable_states", c_uint64),
        ("certified_candidate_states", c_uint64),
        ("candidate_states", c_uint64),
        ("pair_count", c_uint64),
        ("shape", CaseShape),
        ("budget", MemoryBudget),
        ("kernel_timings", KernelTimings),
        ("memory_stats", MemoryStats),
    ]


class ResultBuffers(ctypes.Structure):
    _fields_ = [
        ("valid_mask", c_void_p),
        ("candidate_mask", c_void_p),
        ("reachable_mask", c_void_p),
        ("controller", c_void_p),
        ("reach_step", c_void_p),
        ("abstraction_min_flat", c_void_p),
        ("abstraction_max_flat", c_void_p),
        ("abstraction_valid", c_void_p),
        ("iteration_stats", POINTER(IterationStats)),
        ("iteration_stats_capacity", c_uint32),
    ]


class SolverOptions(ctypes.Structure):
    _fields_ = [
        ("device", c_int32),
        ("keep_export", c_int32),
        ("time_kernels", c_int32),
        ("sweep_all_pairs", c_int32),
        ("verbose", c_int32),
        ("recompute_abstraction", c_int32),
        ("tile_wg", c_int32),
        ("tile_ib", c_int32),
    ]


class SolveSummary(ctypes.Structure):
    _fields_ = [
        ("iterations", c_uint32),
        ("converged", c_int32),
        ("reachable_states", c_uint64),
        ("certified_candidate_states", c_uint64),
        ("candidate_states", c_uint64),
        ("abstraction_ms", c_double),
        ("solve_ms", c_double),
        ("prefix_ms", c_double),
        ("satisfaction_ms", c_double),
        ("reduction_ms", c_double),
        ("pairs_tested", c_uint64),
        ("pair_bytes_read", c_uint64),
        ("boxes_counted", c_uint64),
        ("test_slots", c_uint64),
        ("message", c_char * MESSAGE_LEN),
        ("frontier", c_uint32),
        ("reserved", c_uint32),
    ]


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("device", c_int32),
        ("tile_wg", c_int32),
        ("tile_ib", c_int32),
        ("iterations_per_launch", c_int32),
        ("order", c_int32),
        ("use_graph", c_int32),
        ("reserved", c_int32 * 2),
    ]


class BatchSummary(ctypes.Structure):
    _fields_ = [
        ("cases", c_uint32),
        ("replays", c_uint32),
        ("max_iterations", c_uint32),
        ("stopped", c_uint32),
        ("pairs", c_uint64),
        ("abstraction_ms", c_double),
        ("solve_ms", c_double),
        ("gpu_ms", c_double),
    ]


class BatchTuneEntry(ctypes.Structure):
    _fields_ = [
        ("tile_wg", c_int32),
        ("tile_ib", c_int32),
        ("iterations_per_launch", c_int32),
        ("order", c_int32),
        ("solve_ms", c_double),
    ]


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("device", c_int32),
        ("max_active", c_int32),
        ("iterations_per_step", c_int32),
        ("tile_wg", c_int32),
        ("tile_ib", c_int32),
        ("result_pool_mib", c_int32),
        ("reserved", c_int32 * 2),
        ("hbm_budget_bytes", c_uint64),
    ]


REQUEST_QUEUED, REQUEST_RUNNING, REQUEST_DONE, REQUEST_FAILED = 0, 1, 2, 3


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("id", c_uint64),
        ("state", c_int32),
        ("status", c_int32),
        ("state_count", c_uint64),
        ("pair_count", c_uint64),
        ("device_bytes", c_uint64),
        ("queued_ms", c_double),
        ("latency_ms", c_double),
        ("summary", SolveSummary),
        ("error", c_char * MESSAGE_LEN),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", c_uint64),
        ("completed", c_uint64),
        ("failed", c_uint64),
        ("steps", c_uint64),
        ("active", c_uint32),
        ("queued", c_uint32),
        ("max_active_seen", c_uint32),
        ("reserved", c_uint32),
        ("budget_bytes", c_uint64),
        ("budget_in_use", c_uint64),
        ("budget_peak", c_uint64),
    ]


# Every exported symbol of include/gsc_amd.h: (name, restype, argtypes).
SYMBOLS = [
    ("gsc_backend_compiled", c_int, []),
    ("gsc_device_count", c_int, []),
    ("gsc_case_shape_of", c_int, [POINTER(CaseDesc), POINTER(CaseShape), c_char_p, c_size_t]),
    ("gsc_run_case_u32", c_int, [POINTER(CaseDesc), POINTER(ResultBuffers), POINTER(RunReport)]),
    ("gsc_run_case_u32_static", c_int, [POINTER(CaseDesc), POINTER(ResultBuffers), POINTER(RunReport)]),
    ("gsc_solver_default_options", None, [POINTER(SolverOptions)]),
    ("gsc_solver_create", c_int, [POINTER(CaseDesc), POINTER(SolverOptions), POINTER(c_void_p), c_char_p, c_size_t]),
    ("gsc_solver_shape", c_int, [c_void_p, POINTER(CaseShape)]),
    ("gsc_solver_build_abstraction", c_int, [c_void_p, POINTER(c_double)]),
    ("gsc_solver_solve", c_int, [c_void_p, POINTER(SolveSummary)]),
    ("gsc_solver_download", c_int, [c_void_p, POINTER(ResultBuffers), POINTER(RunReport)]),
    ("gsc_solver_last_error", c_char_p, [c_void_p]),
    ("gsc_solver_destroy", None, [c_void_p]),
    ("gsc_config_load", c_int, [c_char_p, POINTER(c_void_p), c_char_p, c_size_t]),
    ("gsc_config_parse", c_int, [c_char_p, POINTER(c_void_p), c_char_p, c_size_t]),
    ("gsc_config_desc", POINTER(CaseDesc), [c_void_p]),
    ("gsc_config_free", None, [c_void_p]),
    ("gsc_config_builtin", c_int, [c_int, POINTER(c_void_p)]),
    ("gsc_batch_default_options", None, [POINTER(BatchOptions)]),
    ("gsc_batch_create", c_int, [POINTER(POINTER(CaseDesc)), c_uint32, POINTER(BatchOptions), POINTER(c_void_p),
                                 c_char_p, c_size_t]),
    ("gsc_batch_build_abstraction", c_int, [c_void_p, POINTER(c_double)]),
    ("gsc_batch_solve", c_int, [c_void_p, POINTER(BatchSummary)]),
    ("gsc_batch_tune", c_int, [c_void_p, POINTER(BatchTuneEntry), c_uint32, POINTER(c_uint32)]),
    ("gsc_batch_get_options", c_int, [c_void_p, POINTER(BatchOptions)]),
    ("gsc_batch_set_options", c_int, [c_void_p, POINTER(BatchOptions)]),
    ("gsc_batch_case_summary", c_int, [c_void_p, c_uint32, POINTER(SolveSummary)]),
    ("gsc_batch_download", c_int, [c_void_p, c_uint32, POINTER(ResultBuffers), POINTER(RunReport)]),
    ("gsc_batch_last_error", c_char_p, [c_void_p]),
    ("gsc_batch_destroy", None, [c_void_p]),
    ("gsc_server_default_options", None, [POINTER(ServerOptions)]),
    ("gsc_server_create", c_int, [POINTER(ServerOptions), POINTER(c_void_p), c_char_p, c_size_t]),
    ("gsc_server_submit", c_int, [c_void_p, POINTER(CaseDesc), POINTER(c_uint64)]),
    ("gsc_server_wait", c_int, [c_void_p, c_uint64, c_double, POINTER(RequestInfo)]),
    ("gsc_server_fetch", c_int, [c_void_p, c_uint64, POINTER(ResultBuffers), POINTER(RunReport)]),
    ("gsc_server_get_stats", c_int, [c_void_p, POINTER(ServerStats)]),
    ("gsc_server_destroy", None, [c_void_p]),
]

_lib = None


def load() -> ctypes.CDLL:
    """Load libgsc_amd.so from the package directory (built by `make` there)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libgsc_amd.so not found at {LIB_PATH}; build it with `make -C {PKG_DIR}` "
            "(or __graft_entry__.build()) — there is no fallback path"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, restype, argtypes in SYMBOLS:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib
