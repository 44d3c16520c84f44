"""ctypes binding of libdense_hodlr_amd.so (declarations in include/dense_hodlr.h).

The library is loaded from dense_hodlr_amd/lib/ only.  There is no fallback:
if it is missing, or no GPU is usable, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdense_hodlr_amd.so")

HODLR_OK = 0
ERRORS = {
    -1: "HODLR_ERR_ARG",
    -2: "HODLR_ERR_STATE",
    -3: "HODLR_ERR_DEVICE",
    -4: "HODLR_ERR_NOMEM",
    -5: "HODLR_ERR_NUMERIC",
    -6: "HODLR_ERR_UNSUPPORTED",
    -7: "HODLR_ERR_NOT_CONVERGED",
}
HODLR_ERR_NOT_CONVERGED = -7

KERNELS = {"gaussian": 0, "matern32": 1, "se": 2, "imq": 3, "quadratic": 4, "exponential": 5}
LR_METHODS = {"partialPiv_ACA": 0, "SVD": 1}
PRECISIONS = {"fp64": 0, "mixed": 1, "fp64_refined": 2}
FACTOR_SM, FACTOR_LU = 0, 1

# every symbol include/dense_hodlr.h declares
EXPORTS = [
    "hodlr_create_kernel", "hodlr_create_kernel_callback", "hodlr_create_dense", "hodlr_destroy",
    "hodlr_set_lr_method", "hodlr_set_tolerance", "hodlr_set_min_pivot", "hodlr_set_precision",
    "hodlr_set_refinement", "hodlr_set_profiling", "hodlr_set_graphs", "hodlr_autotune", "hodlr_set_launch_shapes", "hodlr_compress", "hodlr_factorize", "hodlr_solve_sm",
    "hodlr_solve_lu", "hodlr_solve_device", "hodlr_logabsdet", "hodlr_determinant",
    "hodlr_matvec", "hodlr_matvec_device_noalloc", "hodlr_residual", "hodlr_get_stats", "hodlr_rows", "hodlr_num_nodes", "hodlr_get_nodes",
    "hodlr_get_factor", "hodlr_block", "hodlr_block_device", "hodlr_last_error", "hodlr_version",
    "hodlr_create_batch", "hodlr_batch_count", "hodlr_batch_offsets", "hodlr_logabsdet_batch",
    "hodlr_rank_model_create", "hodlr_rank_model_destroy", "hodlr_rank_model_num_params",
    "hodlr_rank_model_get_params", "hodlr_rank_model_set_params", "hodlr_rank_model_forward",
    "hodlr_rank_model_predict_bf16", "hodlr_rank_model_gradient", "hodlr_rank_model_train",
    "hodlr_block_features", "hodlr_set_rank_predictor",
    "hodlr_memory_get_stats", "hodlr_memory_trim", "hodlr_memory_cache",
    "hodlr_measure_device_peaks", "hodlr_read_binary_matrix", "hodlr_write_binary_matrix", "hodlr_read_text_matrix",
    "hodlr_create_dense_from_file", "hodlr_create_kernel_tree", "hodlr_create_dense_tree",
    "hodlr_kd_sort", "hodlr_kd_tree", "hodlr_diag_block", "hodlr_split_at_top",
    "hodlr_iterative_solve", "hodlr_iterative_solve_device", "hodlr_exact_apply_device",
]


class KernelDesc(C.Structure):
    _fields_ = [
        ("family", C.c_int),
        ("dim", C.c_int),
        ("points", C.POINTER(C.c_double)),
        ("length_scale", C.c_double),
        ("diag", C.c_double),
        ("diag_mode", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("num_levels", C.c_int),
        ("num_nodes", C.c_int),
        ("num_leaves", C.c_int),
        ("aca_iterations", C.c_int),
        ("compress_time", C.c_double),
        ("factor_time", C.c_double),
        ("solve_time", C.c_double),
        ("refine_time", C.c_double),
        ("refine_iterations", C.c_int),
        ("last_residual", C.c_double),
        ("max_rank", C.c_int * 64),
        ("mean_rank", C.c_double * 64),
        ("device_bytes", C.c_longlong),
        ("aca_row_time", C.c_double),
        ("aca_col_time", C.c_double),
        ("aca_launches", C.c_int),
        ("aca_bytes", C.c_double),
        ("pred_blocks", C.c_int),
        ("pred_misses", C.c_int),
        ("pred_log_err", C.c_double),
        ("graph_captures", C.c_int),
        ("graph_launches", C.c_int),
        ("tune_aca_rows_per_thread", C.c_int),
        ("tune_gemm_tile_m", C.c_int),
        ("tune_gemm_split_target", C.c_int),
        ("tune_lu_panel", C.c_int),
        ("tuned", C.c_int),
        ("tune_gemm_mfma", C.c_int),
        ("factor_stream_time", C.c_double),
        ("tune_leaf_solve", C.c_int),
        ("iter_iterations", C.c_int),
        ("iter_precond_residual", C.c_double),
        ("iter_residual", C.c_double),
        ("fac_time", C.c_double * 4),
        ("fac_flops", C.c_double * 4),
        ("fac_bytes", C.c_double * 4),
        ("fac_bound_time", C.c_double * 4),
        ("fac_launches", C.c_int * 4),
        ("aca_panel_grows", C.c_int),
        ("aca_panel_bytes", C.c_double),
    ]


class DevicePeaks(C.Structure):
    _fields_ = [(name, C.c_double) for name in ("hbm_read_gbs", "hbm_copy_gbs", "fp64_mfma_tflops",
                                                 "fp64_valu_tflops", "fp32_mfma_tflops")]


class MemoryStats(C.Structure):
    _fields_ = [
        ("live_bytes", C.c_longlong),
        ("cached_bytes", C.c_longlong),
        ("hits", C.c_longlong),
        ("misses", C.c_longlong),
        ("malloc_seconds", C.c_double),
        ("free_seconds", C.c_double),
    ]


CALLBACK = C.CFUNCTYPE(C.c_double, C.c_int, C.c_int, C.c_void_p)

_lib = None


class HodlrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class HodlrNotConverged(HodlrError):
    """Refinement ended above its target; ``x`` holds the best iterate
    (set by the solve that raised it)."""

    x = None


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with `python -m dense_hodlr_amd.build`")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    fp = C.POINTER(C.c_float)
    h = C.c_void_p
    hp = C.POINTER(C.c_void_p)
    sig = {
        "hodlr_create_kernel": (C.c_int, [hp, C.c_int, C.POINTER(KernelDesc), C.c_int]),
        "hodlr_create_kernel_callback": (C.c_int, [hp, C.c_int, CALLBACK, C.c_void_p, C.c_int]),
        "hodlr_create_dense": (C.c_int, [hp, C.c_int, dp, C.c_int, C.c_int, C.c_int]),
        "hodlr_destroy": (C.c_int, [h]),
        "hodlr_set_lr_method": (C.c_int, [h, C.c_int]),
        "hodlr_set_tolerance": (C.c_int, [h, C.c_double]),
        "hodlr_set_min_pivot": (C.c_int, [h, C.c_double]),
        "hodlr_set_precision": (C.c_int, [h, C.c_int]),
        "hodlr_set_refinement": (C.c_int, [h, C.c_double, C.c_int]),
        "hodlr_set_profiling": (C.c_int, [h, C.c_int]),
        "hodlr_set_graphs": (C.c_int, [h, C.c_int]),
        "hodlr_autotune": (C.c_int, [h]),
        "hodlr_set_launch_shapes": (C.c_int, [h, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_int]),
        "hodlr_compress": (C.c_int, [h]),
        "hodlr_factorize": (C.c_int, [h, C.c_int]),
        "hodlr_solve_sm": (C.c_int, [h, dp, C.c_int, C.c_int, dp, C.c_int]),
        "hodlr_solve_lu": (C.c_int, [h, dp, C.c_int, C.c_int, dp, C.c_int]),
        "hodlr_solve_device": (C.c_int, [h, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]),
        "hodlr_logabsdet": (C.c_int, [h, dp]),
        "hodlr_determinant": (C.c_int, [h, dp]),
        "hodlr_matvec": (C.c_int, [h, dp, C.c_int, C.c_int, dp, C.c_int]),
        "hodlr_matvec_device_noalloc": (C.c_int, [h, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                                  C.c_int]),
        "hodlr_residual": (C.c_int, [h, dp, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int]),
        "hodlr_get_stats": (C.c_int, [h, C.POINTER(Stats)]),
        "hodlr_rows": (C.c_int, [h]),
        "hodlr_num_nodes": (C.c_int, [h]),
        "hodlr_get_nodes": (C.c_int, [h, ip, ip, ip, ip, ip, ip]),
        "hodlr_get_factor": (C.c_int, [h, C.c_int, C.c_int, dp]),
        "hodlr_block": (C.c_int, [h, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int]),
        "hodlr_block_device": (C.c_int, [h, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_void_p, C.c_int]),
        "hodlr_last_error": (C.c_char_p, [h]),
        "hodlr_version": (C.c_char_p, []),
        "hodlr_create_batch": (C.c_int, [hp, C.c_int, ip, C.POINTER(KernelDesc), C.c_int]),
        "hodlr_batch_count": (C.c_int, [h]),
        "hodlr_batch_offsets": (C.c_int, [h, ip]),
        "hodlr_logabsdet_batch": (C.c_int, [h, dp]),
        "hodlr_rank_model_create": (C.c_int, [hp, C.c_ulonglong]),
        "hodlr_rank_model_destroy": (C.c_int, [h]),
        "hodlr_rank_model_num_params": (C.c_int, []),
        "hodlr_rank_model_get_params": (C.c_int, [h, fp]),
        "hodlr_rank_model_set_params": (C.c_int, [h, fp]),
        "hodlr_rank_model_forward": (C.c_int, [h, fp, C.c_int, fp]),
        "hodlr_rank_model_predict_bf16": (C.c_int, [h, fp, C.c_int, fp]),
        "hodlr_rank_model_gradient": (C.c_int, [h, fp, fp, C.c_int, fp, fp]),
        "hodlr_rank_model_train": (C.c_int, [h, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float,
                                             C.c_ulonglong, fp]),
        "hodlr_block_features": (C.c_int, [h, C.c_double, fp]),
        "hodlr_set_rank_predictor": (C.c_int, [h, h, C.c_double]),
        "hodlr_read_binary_matrix": (C.c_int, [C.c_char_p, ip, ip, dp, C.c_int]),
        "hodlr_write_binary_matrix": (C.c_int, [C.c_char_p, C.c_int, C.c_int, dp, C.c_int]),
        "hodlr_read_text_matrix": (C.c_int, [C.c_char_p, ip, ip, dp, C.c_int]),
        "hodlr_create_dense_from_file": (C.c_int, [hp, C.c_char_p, C.c_int, C.c_int]),
        "hodlr_create_kernel_tree": (C.c_int, [hp, C.c_int, C.POINTER(KernelDesc), C.c_int, ip,
                                               C.c_int]),
        "hodlr_create_dense_tree": (C.c_int, [hp, C.c_int, dp, C.c_int, C.c_int, ip, C.c_int]),
        "hodlr_kd_sort": (C.c_int, [dp, C.c_int, C.c_int, ip]),
        "hodlr_kd_tree": (C.c_int, [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ip, ip, ip]),
        "hodlr_iterative_solve": (C.c_int, [h, dp, C.c_int, C.c_int, dp, C.c_int, C.c_int,
                                            C.c_double, C.c_double, C.c_int, C.c_int]),
        "hodlr_iterative_solve_device": (C.c_int, [h, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                                   C.c_int, C.c_int, C.c_double, C.c_double,
                                                   C.c_int, C.c_int]),
        "hodlr_exact_apply_device": (C.c_int, [h, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                               C.c_int]),
        "hodlr_diag_block": (C.c_int, [h, C.c_int, hp]),
        "hodlr_split_at_top": (C.c_int, [h, hp, hp]),
        "hodlr_memory_get_stats": (C.c_int, [C.POINTER(MemoryStats)]),
        "hodlr_measure_device_peaks": (C.c_int, [C.POINTER(DevicePeaks)]),
        "hodlr_memory_trim": (C.c_int, []),
        "hodlr_memory_cache": (C.c_int, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, handle=None):
    if rc < 0:
        msg = lib().hodlr_last_error(handle)
        cls = HodlrNotConverged if rc == HODLR_ERR_NOT_CONVERGED else HodlrError
        raise cls(rc, msg.decode() if msg else "")
    return rc
