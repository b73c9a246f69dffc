"""ctypes binding of libmbrec.so (the HIP library built from ../csrc).

The library is the only compute path of this package: every operator and
solver call goes through it, and importing a function that needs it raises
when the shared object is missing or cannot be loaded.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbrec.so")

_lib = None

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_int32_p = C.POINTER(C.c_int32)
c_void_p = C.c_void_p

# (name, restype, argtypes) of every symbol declared in include/mbrec.h
SIGNATURES = [
    ("mbrec_last_error", C.c_char_p, []),
    ("mbrec_last_error_id", C.c_char_p, []),
    ("mbrec_version", C.c_int, []),
    ("mbrec_init", C.c_int, [C.c_int]),
    ("mbrec_forwardMex", C.c_int,
     [c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, C.c_int, c_void_p, C.c_int, C.c_int,
      C.c_int32, c_void_p, C.c_double, c_void_p]),
    ("mbrec_transposeMex", C.c_int,
     [c_void_p, C.c_int, C.c_int, C.c_int, c_void_p, c_void_p, C.c_int, c_void_p, C.c_int, C.c_int,
      c_void_p, C.c_double, c_void_p]),
    ("mbrec_forwardDelayAndSumMex", C.c_int,
     [c_void_p, C.c_int, C.c_int, c_void_p, C.c_int, C.c_int32, C.c_double, c_void_p]),
    ("mbrec_transposeDelayAndSumMex", C.c_int,
     [c_void_p, C.c_int, C.c_int, C.c_int, c_void_p, C.c_int, C.c_double, c_void_p]),
    ("mbrec_mex_type_error", C.c_int, [C.c_char_p, C.c_int, C.c_int]),
    ("mbrec_mex_cache_clear", None, []),
    ("mbrec_plan_create", C.c_int,
     [C.POINTER(c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, c_void_p]),
    ("mbrec_plan_set_sir", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, C.c_int, c_void_p]),
    ("mbrec_plan_set_ir", C.c_int, [c_void_p, c_void_p, C.c_int, C.c_int]),
    ("mbrec_plan_set_prefilter", C.c_int, [c_void_p, c_void_p, C.c_int, c_void_p, C.c_int]),
    ("mbrec_plan_set_sigma", C.c_int, [c_void_p, C.c_double]),
    ("mbrec_plan_destroy", C.c_int, [c_void_p]),
    ("mbrec_plan_device_bytes", C.c_size_t, [c_void_p]),
    ("mbrec_plan_tau_frac_bits", C.c_int, [c_void_p]),
    ("mbrec_plan_autotune", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p, c_int32_p]),
    ("mbrec_plan_set_launch", C.c_int, [c_void_p, C.c_int, C.c_int, C.c_int]),
    ("mbrec_plan_set_fir_kernel", C.c_int, [c_void_p, C.c_int]),
    ("mbrec_num_gather_variants", C.c_int, []),
    ("mbrec_plan_time_stage", C.c_int, [c_void_p, C.c_int, C.c_int, C.c_int, c_void_p, c_float_p]),
    ("mbrec_mex_forward_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p]),
    ("mbrec_mex_transpose_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p]),
    ("mbrec_model_forward_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p]),
    ("mbrec_model_transpose_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p]),
    ("mbrec_model_forward_window_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p, c_void_p]),
    ("mbrec_model_transpose_window_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p, c_void_p]),
    ("mbrec_model_transpose_noalloc_dev", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, c_void_p]),
    ("mbrec_rec_nn_with_L2_reg_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, c_void_p, c_void_p, c_void_p,
      c_void_p]),
    ("mbrec_travel_times_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, c_void_p, c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, c_void_p,
      C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, c_void_p, c_void_p, c_void_p]),
    ("mbrec_spmr_grid_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, C.c_int, C.c_double, C.c_double, c_void_p,
      c_void_p, c_void_p]),
    ("mbrec_spmr_index_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, C.c_int, C.c_double, C.c_double, C.c_double,
      C.c_int, c_void_p, c_void_p]),
    ("mbrec_spmr_table_dev", C.c_int,
     [c_void_p, C.c_int, c_void_p, C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, c_void_p, C.c_int,
      C.c_int, C.c_int, C.c_double, c_void_p, c_void_p, c_void_p]),
    ("mbrec_filter_butter_zero_phase_dev", C.c_int,
     [c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, c_void_p, c_void_p]),
    ("mbrec_apply_butterworth_window_dev", C.c_int,
     [c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, c_void_p]),
    ("mbrec_interpolate_broken_transducers_dev", C.c_int,
     [c_void_p, C.c_int, C.c_int, C.c_int, c_void_p, C.c_int, c_void_p, c_void_p]),
    ("mbrec_del2_dev", C.c_int, [c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, c_void_p]),
    ("mbrec_rec_nn_with_L2_reg_terms_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, c_void_p, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mbrec_rec_nn_with_L2_reg_batch_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p]),
    ("mbrec_rec_nn_with_L1_eye_reg_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, c_void_p, c_void_p]),
    ("mbrec_rec_nn_with_TV_reg_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, c_void_p, c_void_p,
      c_void_p]),
    ("mbrec_sb_itv2_dev", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mbrec_dmas_cf_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
      C.c_double, C.c_double, c_void_p, c_void_p]),
    ("mbrec_backproject_waveeq_dev", C.c_int,
     [c_void_p, c_void_p, C.c_int, C.c_int, c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
      C.c_int, C.c_double, C.c_double, c_void_p, C.c_int, C.c_int, c_void_p, C.c_int, c_void_p, c_void_p]),
    ("mbrec_bp_kernel_dev", C.c_int, [c_void_p, C.c_int, c_void_p, c_void_p]),
    ("mbrec_upsample_detectors_dev", C.c_int,
     [c_void_p, C.c_int, C.c_int, c_void_p, C.c_int, C.c_int, c_void_p, c_void_p]),
    ("mbrec_batch_create", C.c_int,
     [C.POINTER(c_void_p), C.c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
      c_void_p, C.c_int]),
    ("mbrec_batch_set_lambda", C.c_int, [c_void_p, c_void_p]),
    ("mbrec_batch_run", C.c_int, [c_void_p, C.c_int, C.c_int, c_void_p]),
    ("mbrec_batch_destroy", C.c_int, [c_void_p]),
    ("mbrec_unet_create", C.c_int, [C.POINTER(c_void_p), C.c_int, C.c_ulonglong]),
    ("mbrec_unet_destroy", C.c_int, [c_void_p]),
    ("mbrec_unet_num_params", C.c_size_t, [c_void_p]),
    ("mbrec_unet_layout", C.c_int,
     [c_void_p, C.c_int, c_int32_p, c_int32_p, c_int32_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("mbrec_unet_get_params", C.c_int, [c_void_p, c_void_p, C.c_size_t]),
    ("mbrec_unet_set_params", C.c_int, [c_void_p, c_void_p, C.c_size_t]),
    ("mbrec_unet_forward", C.c_int, [c_void_p, c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, c_void_p]),
    ("mbrec_unet_gradients", C.c_int,
     [c_void_p, c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, c_void_p, c_double_p, c_void_p]),
    ("mbrec_unet_set_relative_loss", C.c_int, [c_void_p, C.c_int]),
    ("mbrec_unet_train_step", C.c_int,
     [c_void_p, c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
      c_double_p, c_void_p]),
    ("mbrec_conv3x3_bf16", C.c_int,
     [c_void_p, c_void_p, c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, c_void_p]),
]


REG_APPLY = C.CFUNCTYPE(C.c_int, c_void_p, c_void_p, c_void_p, C.c_int, C.c_int, C.c_int, c_void_p)


class L2Term(C.Structure):
    """mbrec_l2_term of include/mbrec.h."""
    _fields_ = [("lambda_", C.c_double), ("kind", C.c_int), ("apply", REG_APPLY), ("ctx", c_void_p)]


class MbrecError(RuntimeError):
    """Error raised from the native library; ``identifier`` mirrors MATLAB error IDs."""

    def __init__(self, identifier: str, message: str, code: int = -1):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier
        self.code = code


def lib():
    """Loads libmbrec.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        L = lib()
        raise MbrecError(L.mbrec_last_error_id().decode(), L.mbrec_last_error().decode(), rc)


def ptr(a) -> int:
    """Raw address of a numpy array or torch tensor (must be contiguous)."""
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def stream_handle(device=None) -> int:
    import torch

    return torch.cuda.current_stream(device).cuda_stream
