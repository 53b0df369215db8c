"""ctypes binding of libgraeng.so (the C ABI of include/graeng.h).

The library is loaded from the package directory.  There is no CPU fallback:
if the library is missing or no GPU is visible, every engine call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgraeng.so")

c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)

GRA_OK = 0
STATUS_NAMES = {0: "GRA_OK", 1: "GRA_BAD_ARG", 2: "GRA_OOM", 3: "GRA_HIP_ERROR",
                4: "GRA_NUMERIC_BREAKDOWN", 5: "GRA_DEGENERATE_GEOMETRY", 6: "GRA_NO_DEVICE"}

F64, F32, BF16 = 0, 1, 2
DTYPES = {"f64": F64, "f32": F32, "bf16": BF16}

EXPORTS = [
    "gra_status_string", "gra_last_error", "gra_ctx_create", "gra_ctx_destroy", "gra_sync",
    "gra_ctx_set_stream", "gra_ctx_stream",
    "gra_assemble", "gra_get_dims", "gra_copy_G", "gra_G_vec_mul", "gra_GT_vec_mul",
    "gra_col_norms", "gra_expand_cols", "gra_assemble_batch", "gra_assemble_batch_corners",
    "gra_assemble_wavelet",
    "gra_wavelet_info", "gra_assemble_rectilinear", "gra_assemble_corners", "gra_G_mat_mul", "gra_GT_mat_mul",
    "gra_forward", "gra_forward_rectilinear", "gra_set_regularization", "gra_solve_cg", "gra_barrier_P",
    "gra_barrier_update", "gra_misfit", "gra_irls_weights", "gra_gn_setup_problem",
    "gra_gn_step", "gra_gn_get_model", "gra_timing_enable", "gra_timing_reset",
    "gra_timing_get", "gra_tune_set", "gra_tune_get", "gra_autotune", "gra_mixed_precision",
    "gra_G_vec_mul_noalloc", "gra_gn_step_batch", "gra_forward_rectilinear_batch",
]

TUNE = {"n_rows": 0, "n_unroll": 1, "n_splits": 2, "t_chunks": 3, "t_unroll": 4, "nontemporal": 5,
        "fused_wgs": 6, "fused_depth": 7, "n_block": 8, "node_tile": 9, "fused_fail": 10,
        "fused_fallbacks": 11, "fused_math": 12, "last_refreshes": 13, "cg_graph": 14, "fused_min_blocks": 15}


class GraError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)} {msg}".strip())


class GraCsr(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("row_ptr", c_i64p),
                ("col_idx", c_i32p), ("val", c_dp)]


class GraBatchItem(ctypes.Structure):
    _fields_ = [("prisms", c_dp), ("Nm", ctypes.c_int64), ("pts", c_dp), ("N_obs", ctypes.c_int64),
                ("field_flag", ctypes.c_uint32)]


class GraFwdRectItem(ctypes.Structure):
    _fields_ = [("xs", c_dp), ("nx", ctypes.c_int64), ("ys", c_dp), ("ny", ctypes.c_int64),
                ("zs", c_dp), ("nz", ctypes.c_int64), ("rho", c_dp), ("pts", c_dp),
                ("N_obs", ctypes.c_int64), ("field_flag", ctypes.c_uint32), ("out", c_dp)]


class GraCgArgs(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxit", ctypes.c_int), ("lambda_", ctypes.c_double),
                ("mk", c_dp), ("mref", c_dp), ("Wd", c_dp), ("dobs", c_dp), ("P", c_dp),
                ("WL_s", c_dp), ("WL_x", c_dp), ("WL_y", c_dp), ("WL_z", c_dp),
                ("WL_tx", c_dp), ("WL_ty", c_dp), ("WL_tz", c_dp),
                ("use_cross_gradient", ctypes.c_int), ("a_crg", ctypes.c_double),
                ("S_crg", c_dp), ("quirk_mode", ctypes.c_int), ("fused", ctypes.c_int)]


class GraGnSetup(ctypes.Structure):
    _fields_ = [("dobs", c_dp), ("Wd", c_dp), ("m_min", c_dp), ("m_max", c_dp),
                ("mref", c_dp), ("m_ini", c_dp), ("lp", ctypes.c_int), ("eps2", ctypes.c_double),
                ("fused", ctypes.c_int), ("use_cross_gradient", ctypes.c_int),
                ("a_crg", ctypes.c_double), ("S_crg", c_dp)]


_lib = None


def load() -> ctypes.CDLL:
    """Loads libgraeng.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m graada3d_amd.build` "
                           "(the engine has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    sig = {
        "gra_status_string": ([ctypes.c_int], ctypes.c_char_p),
        "gra_last_error": ([vp], ctypes.c_char_p),
        "gra_ctx_create": ([ctypes.c_int, ctypes.POINTER(vp)], ctypes.c_int),
        "gra_ctx_destroy": ([vp], None),
        "gra_sync": ([vp], ctypes.c_int),
        "gra_ctx_set_stream": ([vp, vp], ctypes.c_int),
        "gra_ctx_stream": ([vp], vp),
        "gra_assemble": ([vp, c_dp, ctypes.c_int64, c_dp, ctypes.c_int64, ctypes.c_uint32,
                          ctypes.c_int], ctypes.c_int),
        "gra_assemble_wavelet": ([vp, c_dp, ctypes.c_int64, c_dp, ctypes.c_int64,
                                  ctypes.c_uint32, c_i64p, ctypes.c_double], ctypes.c_int),
        "gra_wavelet_info": ([vp, c_i64p, c_i64p], ctypes.c_int),
        "gra_assemble_rectilinear": ([vp, c_dp, ctypes.c_int64, c_dp, ctypes.c_int64, c_dp,
                                      ctypes.c_int64, c_dp, ctypes.c_int64, ctypes.c_uint32,
                                      ctypes.c_int], ctypes.c_int),
        "gra_assemble_corners": ([vp, c_dp, ctypes.c_int64, c_dp, ctypes.c_int64,
                                  ctypes.c_uint32, ctypes.c_int], ctypes.c_int),
        "gra_get_dims": ([vp, c_i64p, c_i64p, c_i64p, ctypes.POINTER(ctypes.c_uint32),
                          ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "gra_copy_G": ([vp, c_dp, ctypes.c_int], ctypes.c_int),
        "gra_G_vec_mul": ([vp, c_dp, c_dp], ctypes.c_int),
        "gra_G_mat_mul": ([vp, c_dp, ctypes.c_int64, c_dp], ctypes.c_int),
        "gra_GT_mat_mul": ([vp, c_dp, ctypes.c_int64, c_dp], ctypes.c_int),
        "gra_GT_vec_mul": ([vp, c_dp, c_dp], ctypes.c_int),
        "gra_col_norms": ([vp, c_dp], ctypes.c_int),
        "gra_expand_cols": ([vp, c_i64p, ctypes.c_int64, c_dp], ctypes.c_int),
        "gra_assemble_batch": ([ctypes.POINTER(vp), ctypes.c_int64,
                                ctypes.POINTER(GraBatchItem), ctypes.c_int], ctypes.c_int),
        "gra_assemble_batch_corners": ([ctypes.POINTER(vp), ctypes.c_int64,
                                        ctypes.POINTER(GraBatchItem), ctypes.c_int],
                                       ctypes.c_int),
        "gra_forward": ([vp, c_dp, c_dp, ctypes.c_int64, c_dp, ctypes.c_int64, ctypes.c_uint32,
                         c_dp], ctypes.c_int),
        "gra_forward_rectilinear": ([vp, c_dp, ctypes.c_int64, c_dp, ctypes.c_int64, c_dp,
                                     ctypes.c_int64, c_dp, c_dp, ctypes.c_int64, ctypes.c_uint32,
                                     c_dp], ctypes.c_int),
        "gra_set_regularization": ([vp] + [ctypes.POINTER(GraCsr)] * 7, ctypes.c_int),
        "gra_solve_cg": ([vp, ctypes.POINTER(GraCgArgs), c_dp, c_dp, ctypes.POINTER(ctypes.c_int)],
                         ctypes.c_int),
        "gra_barrier_P": ([vp, c_dp, c_dp, c_dp, c_dp], ctypes.c_int),
        "gra_barrier_update": ([vp, c_dp, c_dp, c_dp, c_dp, c_dp], ctypes.c_int),
        "gra_misfit": ([vp, c_dp, c_dp, c_dp, c_dp], ctypes.c_int),
        "gra_irls_weights": ([vp, ctypes.POINTER(GraCsr), c_dp, ctypes.c_int, ctypes.c_double,
                              c_dp], ctypes.c_int),
        "gra_gn_setup_problem": ([vp, ctypes.POINTER(GraGnSetup)], ctypes.c_int),
        "gra_gn_step": ([vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, c_dp, c_dp,
                         ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "gra_gn_get_model": ([vp, c_dp], ctypes.c_int),
        "gra_timing_enable": ([vp, ctypes.c_int], ctypes.c_int),
        "gra_timing_reset": ([vp], ctypes.c_int),
        "gra_timing_get": ([vp, ctypes.c_int, c_dp, c_i64p], ctypes.c_int),
        "gra_tune_set": ([vp, ctypes.c_int, ctypes.c_int64], ctypes.c_int),
        "gra_tune_get": ([vp, ctypes.c_int, c_i64p], ctypes.c_int),
        "gra_autotune": ([vp, ctypes.c_int, c_dp, c_dp], ctypes.c_int),
        "gra_mixed_precision": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "gra_G_vec_mul_noalloc": ([vp, vp, vp], ctypes.c_int),
        "gra_forward_rectilinear_batch": ([vp, ctypes.c_int64, ctypes.POINTER(GraFwdRectItem)],
                                          ctypes.c_int),
        "gra_gn_step_batch": ([ctypes.POINTER(vp), ctypes.c_int64, c_dp, ctypes.c_double,
                               ctypes.c_int, c_dp, c_dp, ctypes.POINTER(ctypes.c_int)],
                              ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L
