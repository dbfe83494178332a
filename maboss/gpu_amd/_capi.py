"""ctypes binding of the engine's C-ABI (include/maboss_gpu.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
maboss/gpu_amd/csrc``) into ``maboss/gpu_amd/lib/libmaboss_gpu.so``.  Nothing
here falls back to another implementation: if the library is missing, loading
fails with an error that says how to build it.
"""
from __future__ import annotations

import ctypes
import os
import threading

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmaboss_gpu.so")
CLI_PATH = os.path.join(LIB_DIR, "MaBoSS.GPU")

MBG_OK = 0
ERROR_NAMES = {
    1: "ARG", 2: "IO", 3: "PARSE", 4: "MODEL", 5: "LIMIT", 6: "COMPILE",
    7: "HIP", 8: "OVERFLOW", 9: "INTERNAL",
}

# Every symbol declared in include/maboss_gpu.h.
EXPORTED = [
    "mbg_last_error", "mbg_version",
    "mbg_model_load", "mbg_model_load_strings", "mbg_model_free", "mbg_model_get_info",
    "mbg_model_node_name", "mbg_model_node_internal", "mbg_model_set_constant",
    "mbg_model_set_variable", "mbg_model_param_name", "mbg_model_param_value",
    "mbg_model_eval_rate",
    "mbg_generate_code", "mbg_free_string", "mbg_program_compile", "mbg_program_free",
    "mbg_program_compile_seconds",
    "mbg_sim_params_default", "mbg_simulate", "mbg_simulate_async", "mbg_window_probs_device",
    "mbg_batch_params_default", "mbg_simulate_batched",
    "mbg_grad_params_default", "mbg_grad_check", "mbg_backward", "mbg_backward_noalloc",
    "mbg_result_free", "mbg_result_wait", "mbg_result_get_info", "mbg_result_window_size", "mbg_result_window",
    "mbg_result_window_probs", "mbg_result_entropies", "mbg_result_error_th", "mbg_result_window_errors",
    "mbg_result_final_size",
    "mbg_result_final", "mbg_result_fixed_size", "mbg_result_fixed",
    "mbg_result_write_csv", "mbg_result_print", "mbg_result_write_raw",
    "mbg_result_save", "mbg_result_load", "mbg_result_merge", "mbg_simulate_checkpointed",
]


class MbgError(RuntimeError):
    """A non-zero status from the C-ABI, with mbg_last_error()'s message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{ERROR_NAMES.get(code, code)}] {message}")
        self.code = code
        self.message = message


class ModelInfo(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32), ("n_words", ctypes.c_int32),
        ("n_observed", ctypes.c_int32), ("windows", ctypes.c_int32),
        ("discrete_time", ctypes.c_int32), ("max_time", ctypes.c_float),
        ("time_tick", ctypes.c_float), ("sample_count", ctypes.c_int64),
        ("seed", ctypes.c_uint64), ("n_params", ctypes.c_int32),
        ("total_literals", ctypes.c_int32),
    ]


class CompileOpts(ctypes.Structure):
    _fields_ = [
        ("entropy_formula", ctypes.c_int32), ("no_cache", ctypes.c_int32),
        ("with_gradient", ctypes.c_int32), ("lane_occupancy", ctypes.c_int32),
        ("grad_occupancy", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]


class SimParams(ctypes.Structure):
    _fields_ = [
        ("n_trajectories", ctypes.c_int64), ("first_id", ctypes.c_uint64),
        ("seed", ctypes.c_int64), ("variant", ctypes.c_uint32),
        ("max_steps", ctypes.c_uint32), ("window_mode", ctypes.c_int32),
        ("final_mode", ctypes.c_int32), ("window_capacity", ctypes.c_uint64),
        ("final_capacity", ctypes.c_uint64), ("fixed_capacity", ctypes.c_uint64),
        ("block_size", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
        ("normalization", ctypes.c_double), ("flags", ctypes.c_int32), ("kernel", ctypes.c_int32),
        ("rates", ctypes.POINTER(ctypes.c_float)), ("moment_states", ctypes.c_int32),
        ("refill_lanes", ctypes.c_int32), ("reserved", ctypes.c_int32 * 2),
    ]


class Variant(ctypes.Structure):
    _fields_ = [
        ("n_trajectories", ctypes.c_int64), ("max_time", ctypes.c_float),
        ("pad", ctypes.c_int32), ("rates", ctypes.POINTER(ctypes.c_float)),
        ("mutations", ctypes.POINTER(ctypes.c_int8)),
    ]


class BatchParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_int64), ("max_steps", ctypes.c_uint32),
        ("block_size", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
        ("traj_per_lane", ctypes.c_int32), ("use_graph", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("lds_cells", ctypes.c_int32), ("reserved", ctypes.c_int32 * 7),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("transitions", ctypes.c_uint64), ("block_size", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32), ("traj_per_lane", ctypes.c_int32),
        ("used_graph", ctypes.c_int32), ("lds_cells", ctypes.c_int32), ("tuned_cached", ctypes.c_int32),
    ]


class GradParams(ctypes.Structure):
    _fields_ = [
        ("n_trajectories", ctypes.c_int64), ("first_id", ctypes.c_uint64),
        ("seed", ctypes.c_int64), ("variant", ctypes.c_uint32),
        ("max_steps", ctypes.c_uint32), ("block_size", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32), ("fold_lanes", ctypes.c_int32), ("reserved", ctypes.c_int32 * 5),
    ]


class GradInfo(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double), ("transitions", ctypes.c_uint64),
        ("block_size", ctypes.c_int32), ("grid", ctypes.c_int32),
    ]


class ResultInfo(ctypes.Structure):
    _fields_ = [
        ("windows", ctypes.c_int32), ("n_observed", ctypes.c_int32),
        ("n_words", ctypes.c_int32), ("discrete_time", ctypes.c_int32),
        ("time_tick", ctypes.c_float), ("max_time", ctypes.c_float),
        ("normalization", ctypes.c_double), ("n_trajectories", ctypes.c_uint64),
        ("transitions", ctypes.c_uint64), ("finished", ctypes.c_uint64),
        ("truncated", ctypes.c_uint64), ("n_fixed_points", ctypes.c_uint64),
        ("n_final_states", ctypes.c_uint64), ("n_window_cells", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("window_mode", ctypes.c_int32), ("final_mode", ctypes.c_int32),
        ("window_capacity", ctypes.c_uint64), ("kernel", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("simulate_ms", ctypes.c_double), ("window_pairs", ctypes.c_uint64),
        ("moment_spills", ctypes.c_uint64), ("has_errors", ctypes.c_int32), ("pad2", ctypes.c_int32),
        ("window_spills", ctypes.c_uint64),
        ("window_record_bytes", ctypes.c_uint32),
        ("pad3", ctypes.c_int32),
        ("window_escapes", ctypes.c_uint64),
    ]


_lock = threading.Lock()
_lib = None


def _declare(lib):
    P = ctypes.c_void_p
    i32, i64, u32, u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "mbg_last_error": ([], ctypes.c_char_p),
        "mbg_version": ([], ctypes.c_char_p),
        "mbg_model_load": ([ctypes.c_char_p, ctypes.c_char_p, pp], ctypes.c_int),
        "mbg_model_load_strings": ([ctypes.c_char_p, ctypes.c_char_p, pp], ctypes.c_int),
        "mbg_model_free": ([P], None),
        "mbg_model_get_info": ([P, ctypes.POINTER(ModelInfo)], ctypes.c_int),
        "mbg_model_node_name": ([P, i32], ctypes.c_char_p),
        "mbg_model_node_internal": ([P, i32], ctypes.c_int),
        "mbg_model_set_constant": ([P, ctypes.c_char_p, ctypes.c_double], ctypes.c_int),
        "mbg_model_set_variable": ([P, ctypes.c_char_p, ctypes.c_double], ctypes.c_int),
        "mbg_model_param_name": ([P, i32], ctypes.c_char_p),
        "mbg_model_param_value": ([P, i32], ctypes.c_float),
        "mbg_model_eval_rate": ([P, i32, ctypes.POINTER(u32), ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "mbg_generate_code": ([P, ctypes.POINTER(CompileOpts), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "mbg_free_string": ([P], None),
        "mbg_program_compile": ([P, ctypes.POINTER(CompileOpts), pp], ctypes.c_int),
        "mbg_program_free": ([P], None),
        "mbg_program_compile_seconds": ([P], ctypes.c_double),
        "mbg_sim_params_default": ([ctypes.POINTER(SimParams)], None),
        "mbg_simulate": ([P, ctypes.POINTER(SimParams), P, pp], ctypes.c_int),
        "mbg_simulate_async": ([P, ctypes.POINTER(SimParams), P, pp], ctypes.c_int),
        "mbg_window_probs_device": ([P, ctypes.POINTER(SimParams), P, P, P], ctypes.c_int),
        "mbg_batch_params_default": ([ctypes.POINTER(BatchParams)], None),
        "mbg_grad_params_default": ([ctypes.POINTER(GradParams)], None),
        "mbg_grad_check": ([P], ctypes.c_int),
        "mbg_backward": ([P, ctypes.POINTER(GradParams), P, P, P, P, ctypes.POINTER(GradInfo)], ctypes.c_int),
        "mbg_backward_noalloc": ([P, ctypes.POINTER(GradParams), P, P, P, P], ctypes.c_int),
        "mbg_simulate_batched": ([P, ctypes.POINTER(BatchParams), i32, ctypes.POINTER(Variant), P, pp,
                                  ctypes.POINTER(BatchInfo)], ctypes.c_int),
        "mbg_result_free": ([P], None),
        "mbg_result_wait": ([P], ctypes.c_int),
        "mbg_result_get_info": ([P, ctypes.POINTER(ResultInfo)], ctypes.c_int),
        "mbg_result_window_size": ([P, i32], i64),
        "mbg_result_window": ([P, i32, P, P], ctypes.c_int),
        "mbg_result_window_probs": ([P, i32, P, P], ctypes.c_int),
        "mbg_result_entropies": ([P, P, P], ctypes.c_int),
        "mbg_result_error_th": ([P, P], ctypes.c_int),
        "mbg_result_window_errors": ([P, ctypes.c_int32, P], ctypes.c_int),
        "mbg_result_final_size": ([P], i64),
        "mbg_result_final": ([P, P, P], ctypes.c_int),
        "mbg_result_fixed_size": ([P], i64),
        "mbg_result_fixed": ([P, P, P], ctypes.c_int),
        "mbg_result_write_csv": ([P, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "mbg_result_print": ([P, ctypes.c_char_p], ctypes.c_int),
        "mbg_result_write_raw": ([P, ctypes.c_char_p], ctypes.c_int),
        "mbg_result_save": ([P, ctypes.c_char_p], ctypes.c_int),
        "mbg_result_load": ([P, ctypes.c_char_p, pp], ctypes.c_int),
        "mbg_result_merge": ([P, P], ctypes.c_int),
        "mbg_simulate_checkpointed": ([P, ctypes.POINTER(SimParams), ctypes.c_uint64, ctypes.c_char_p, P, pp],
                                      ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _ = (i64, u64)


def _framework_hip_first():
    """Load torch (when installed) before the engine.

    torch bundles its own HIP runtime and code-object manager under the same
    sonames as the system ROCm the engine is built against.  Whichever copy
    a process loads first is the one every later library binds to, and torch
    does not see the GPU through the system copy, so torch goes first; the
    engine works with either (its JIT compiles from source when the loaded
    code-object manager cannot read its embedded bitcode)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                    "or `make -C maboss/gpu_amd/csrc`")
            _framework_hip_first()
            handle = ctypes.CDLL(LIB_PATH)
            _declare(handle)
            _lib = handle
        return _lib


def check(rc: int):
    if rc != MBG_OK:
        raise MbgError(rc, lib().mbg_last_error().decode(errors="replace"))
