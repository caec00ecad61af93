"""ctypes binding of libsaigehip.so (include/saige_hip.h).

The library is built in-tree (saige-gpu_amd/libsaigehip.so, see the
Makefile there).  There is no fallback: if the library is missing or cannot
be loaded, every entry point raises.
"""
import ctypes
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libsaigehip.so")

SAIGE_OK = 0
SAIGE_E_ARG = -1
SAIGE_E_STATE = -2
SAIGE_E_NOMEM = -3
SAIGE_E_IO = -4
SAIGE_E_DEVICE = -5
SAIGE_E_NOTCONV = -6
SAIGE_PENDING = 1

GRM_FP32 = 0
GRM_BF16 = 1
GRM_BF16_PANEL = 2
GRM_BF16_SPLIT = 3

_c_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_int = ctypes.c_int
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_u64 = ctypes.c_uint64

_pd = ctypes.POINTER(_f64)


class FitOptions(ctypes.Structure):
    """saige_fit_options"""
    _fields_ = [("maxiter", _int), ("tol", _f64), ("nrun", _int), ("tol_pcg", _f64), ("maxiter_pcg", _int),
                ("trace_cv_cutoff", _f64), ("probe_seed", _u64)]


class FitSpec(ctypes.Structure):
    """saige_fit_spec"""
    _fields_ = [("trait", _int), ("n", _i64), ("subset", ctypes.POINTER(_i64)), ("y", _pd), ("X", _pd),
                ("p", _int), ("offset", _pd)]


class FitResult(ctypes.Structure):
    """saige_fit_result"""
    _fields_ = [("theta", _f64 * 2), ("coefficients", _pd), ("linear_predictors", _pd), ("fitted_values", _pd),
                ("working_y", _pd), ("cov", _pd), ("converged", _int), ("iterations", _int), ("status", _int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "saige_fit_options_default": (None, [ctypes.POINTER(FitOptions)]),
    "saige_fit_null_batch": (_int, [_c_p, _int, ctypes.POINTER(FitSpec), ctypes.POINTER(FitOptions), _int,
                                    ctypes.POINTER(FitResult)]),
    "saige_server_start": (_int, [_c_p, ctypes.POINTER(FitOptions), _int, _i64, ctypes.POINTER(_c_p)]),
    "saige_submit": (_int, [_c_p, ctypes.POINTER(FitSpec), ctypes.POINTER(_i64)]),
    "saige_poll": (_int, [_c_p, _i64, _int, ctypes.POINTER(FitResult), _pd]),
    "saige_server_stop": (_int, [_c_p]),
    "saige_covariate_transform": (_int, [_pd, _i64, _int, _pd, _pd, ctypes.POINTER(_i32), ctypes.POINTER(_int)]),
    "saige_score_test_null": (_int, [_pd, _pd, _pd, _pd, _i64, _int, _pd, _pd, _pd, _pd, _pd, _pd]),
    "saige_glm_start": (_int, [_int, _i64, _pd, _pd, _int, _pd, _pd, _pd, _pd]),
    "saige_glm_working": (_int, [_int, _i64, _pd, _pd, _pd, _pd, _pd, _pd]),
    "saige_coefficients": (_int, [_i64, _int, _pd, _pd, _pd, _f64, _pd, _pd, _pd, _pd]),
    "saige_create": (_int, [_int, ctypes.POINTER(_c_p)]),
    "saige_destroy": (_int, [_c_p]),
    "saige_last_error": (ctypes.c_char_p, []),
    "saige_set_min_maf_for_grm": (_int, [_c_p, _f32]),
    "saige_set_max_missing_rate_for_grm": (_int, [_c_p, _f32]),
    "saige_set_min_mac_variance_ratio": (_int, [_c_p, _f32, _f32, _int]),
    "saige_set_vr_random_markers": (_int, [_c_p, _c_p, _i64]),
    "saige_setgeno": (_int, [_c_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _c_p, _i64, _c_p, _int]),
    "saige_setgeno_host": (_int, [_c_p, _c_p, _i64, _i64, _c_p, _i64, _c_p, _int]),
    "saige_setgeno_device": (_int, [_c_p, _c_p, _i64, _i64, _c_p, _i64, _c_p, _int]),
    "saige_get_nnomissing": (_i64, [_c_p]),
    "saige_get_total_markers": (_i64, [_c_p]),
    "saige_get_mqc": (_i64, [_c_p]),
    "saige_get_n_vr_markers": (_i64, [_c_p]),
    "saige_get_allele_freq": (_int, [_c_p, _c_p]),
    "saige_get_mac": (_int, [_c_p, _c_p]),
    "saige_get_inv_std": (_int, [_c_p, _c_p]),
    "saige_get_qc_marker_index": (_int, [_c_p, _c_p]),
    "saige_get_mac_vr": (_int, [_c_p, _c_p]),
    "saige_get_index_vr": (_int, [_c_p, _c_p]),
    "saige_get_marker_stats": (_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "saige_get_one_snp_geno": (_int, [_c_p, _i64, _c_p]),
    "saige_get_one_snp_geno_vr": (_int, [_c_p, _i64, _c_p]),
    "saige_get_one_snp_std_geno": (_int, [_c_p, _i64, _c_p]),
    "saige_get_diag_std": (_int, [_c_p, _c_p]),
    "saige_get_diag_of_kin": (_int, [_c_p, _c_p]),
    "saige_build_grm": (_int, [_c_p, _int]),
    "saige_get_grm_block": (_int, [_c_p, _i64, _i64, _i64, _i64, _c_p]),
    "saige_last_grm_seconds": (_f64, [_c_p]),
    "saige_grm_precision": (_int, [_c_p]),
    "saige_probes_device_rows": (_int, [_c_p, _c_p, _i64, ctypes.c_uint64, _i64, _int, _c_p, _i64]),
    "saige_alloc_host": (_int, [_c_p, _i64, ctypes.POINTER(_c_p)]),
    "saige_free_host": (_int, [_c_p, _c_p]),
    "saige_crossprod_mat_and_kin": (_int, [_c_p, _c_p, _c_p]),
    "saige_grm_matvec": (_int, [_c_p, _c_p, _i64, _c_p]),
    "saige_get_diag_of_sigma": (_int, [_c_p, _c_p, _c_p, _c_p]),
    "saige_get_crossprod": (_int, [_c_p, _c_p, _c_p, _c_p, _c_p]),
    "saige_pcg1": (_int, [_c_p, _c_p, _c_p, _c_p, _int, _f32, _c_p, _c_p]),
    "saige_set_start_end_index": (_int, [_c_p, _i64, _i64, _int]),
    "saige_loco_range": (_int, [_c_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_int),
                                ctypes.POINTER(_i64)]),
    "saige_crossprod_mat_and_kin_loco": (_int, [_c_p, _c_p, _c_p]),
    "saige_get_diag_of_sigma_loco": (_int, [_c_p, _c_p, _c_p, _c_p]),
    "saige_get_crossprod_loco": (_int, [_c_p, _c_p, _c_p, _c_p, _c_p]),
    "saige_pcg1_loco": (_int, [_c_p, _c_p, _c_p, _c_p, _int, _f32, _c_p, _c_p]),
    "saige_pcg_block": (_int, [_c_p, _c_p, _c_p, _c_p, _i64, _int, _f32, _c_p, _c_p]),
    "saige_engine_reserve": (_int, [_c_p, _int, _int]),
    "saige_engine_set_system": (_int, [_c_p, _int, _c_p, _c_p, _c_p]),
    "saige_engine_add": (_int, [_c_p, _int, _int, _c_p, _int, _f32]),
    "saige_engine_add_batch": (_int, [_c_p, _int, _c_p, _c_p, _c_p, _int, _f32]),
    "saige_engine_add_products": (_int, [_c_p, _int, _c_p, _c_p, _c_p]),
    "saige_engine_take_batch": (_int, [_c_p, _int, _c_p, _c_p]),
    "saige_engine_step": (_int, [_c_p, _int, _c_p, _c_p]),
    "saige_engine_take": (_int, [_c_p, _int, _c_p]),
    "saige_engine_use_loco": (_int, [_c_p, _int]),
    "saige_engine_add_batch_device": (_int, [_c_p, _int, _c_p, _c_p, _c_p, _int, _f32]),
    "saige_engine_add_products_device": (_int, [_c_p, _int, _c_p, _c_p, _c_p]),
    "saige_engine_take_batch_ex": (_int, [_c_p, _int, _c_p, _c_p, _c_p, _i64, _c_p, _c_p, _c_p, _i64, _c_p,
                                          _c_p, _int, _i64, _c_p]),
    "saige_dots_device": (_int, [_c_p, _c_p, _i64, _int, _c_p, _i64, _int, _c_p]),
    "saige_dots_device_noalloc": (_int, [_c_p, _c_p, _i64, _int, _c_p, _i64, _int, _c_p]),
    "saige_probes_device": (_int, [_c_p, _c_p, _i64, ctypes.c_uint64, _i64, _int]),
    "saige_copy_host_to_device": (_int, [_c_p, _c_p, _c_p, _i64]),
    "saige_autotune": (_int, [_c_p, _int, _c_p]),
    "saige_autotune_candidates": (_int, [_c_p, _c_p, _c_p]),
    "saige_set_gemm_config": (_int, [_c_p, _int, _int, _int]),
    "saige_get_gemm_config": (_int, [_c_p, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int),
                                      ctypes.POINTER(_f64)]),
    "saige_get_gemm_config_half": (_int, [_c_p, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_f64)]),
    "saige_get_gemm_config_quarter": (_int, [_c_p, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_f64)]),
    "saige_set_gemm_mode": (_int, [_c_p, _int]),
    "saige_get_gemm_mode": (_int, [_c_p, ctypes.POINTER(_int)]),
    "saige_set_gemv_mode": (_int, [_c_p, _int]),
    "saige_get_gemv_mode": (_int, [_c_p, ctypes.POINTER(_int)]),
    "saige_engine_stats": (_int, [_c_p, _int, ctypes.POINTER(_f64), ctypes.POINTER(_i64)]),
    "saige_device_memory": (_int, [_c_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "saige_engine_system_bytes": (_i64, [_c_p]),
    "saige_synth_bed_device": (_int, [_c_p, _c_p, _i64, _i64, _f32, _f32, _f32, _f32, _u64]),
    "saige_alloc_device": (_int, [_c_p, _i64, ctypes.POINTER(_c_p)]),
    "saige_free_device": (_int, [_c_p, _c_p]),
    "saige_copy_device": (_int, [_c_p, _c_p, _c_p, _i64]),
    "saige_copy_device_to_host": (_int, [_c_p, _c_p, _i64, _c_p]),
}


class SaigeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NotConverged(SaigeError):
    pass


_LIB = None


def load():
    """Load the in-tree library (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `make -C saige-gpu_amd` (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(rc, allow_notconv=False, allow_pending=False):
    if rc == SAIGE_OK or (allow_pending and rc == SAIGE_PENDING):
        return rc
    msg = load().saige_last_error().decode("utf-8", "replace")
    if rc == SAIGE_E_NOTCONV:
        if allow_notconv:
            return rc
        raise NotConverged(rc, msg)
    raise SaigeError(rc, msg)


def ptr(a):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def cols_f32(B):
    """N x k array -> column-major float32 buffer (column c contiguous)."""
    B = np.asarray(B, dtype=np.float32)
    if B.ndim == 1:
        return np.ascontiguousarray(B)
    return np.asfortranarray(B)
