"""ctypes binding of libmsamd.so (the C ABI declared in include/msamd.h).

The library is loaded from ``mountainsort_amd/lib/libmsamd.so``.  There is no
fallback: when the library is missing every entry point raises
``MsamdLibraryMissing``, so a GPU run can never silently take another path.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libmsamd.so")

_lib = None

i64, i32, f64, vp, cint = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p, ctypes.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "msamd_version": (cint, []),
    "msamd_last_error": (ctypes.c_char_p, []),
    "msamd_bandpass_chunk_size": (i64, [i64, i64]),
    "msamd_bandpass_kernel": (cint, [i64, f64, f64, f64, f64, vp]),
    "msamd_bandpass_supported": (cint, [i64, i64]),
    "msamd_bandpass_workspace_size": (i64, [i64, i64, i64, i64]),
    "msamd_bandpass_filter": (cint, [vp, cint, i64, i64, f64, f64, f64, f64, i64, i64, cint, f64,
                                     vp, vp, i64, vp]),
    "msamd_bandpass_windows": (i64, [i64, i64]),
    "msamd_bandpass_filter_windows": (cint, [vp, cint, i64, i64, i64, i64, i64, i64, f64, f64,
                                             f64, f64, i64, i64, cint, f64, vp, i64, i64, vp, i64,
                                             vp]),
    "msamd_detect_chunk_size": (i64, [i64, i64, i64]),
    "msamd_detect_workspace_size": (i64, [i64, i64, i32, f64]),
    "msamd_detect_events": (cint, [vp, i64, i64, i64, i64, i64, i64, vp, i32, f64, f64, cint, vp,
                                   i64, ctypes.POINTER(i64), vp]),
    "msamd_detect_fetch": (cint, [vp, i32, vp, vp, vp]),
    "msamd_extract_clips": (cint, [vp, i64, i64, vp, i64, vp, i32, i32, vp, vp]),
    "msamd_extract_clips_batched": (cint, [vp, i64, i64, vp, vp, i64, vp, vp, i32, i32, vp, vp,
                                           vp]),
    "msamd_pca_workspace_size": (i64, [i32, vp, vp, i32, i32, i32, i64]),
    "msamd_pca_features": (cint, [vp, i32, vp, vp, i32, i32, i32, i64, vp, vp, vp, i64, vp]),
    "msamd_pca_plan_workspace_size": (i64, [i32, vp, vp, i32, i32, i32, i64, vp]),
    "msamd_pca_plan_create": (cint, [i32, vp, vp, i32, i32, i32, i64, vp, vp, i64, vp, vp]),
    "msamd_pca_plan_run": (cint, [vp, vp, vp, vp, vp]),
    "msamd_pca_plan_launch_graph": (cint, [vp, vp, vp, vp, vp]),
    "msamd_pca_plan_run_timeseries": (cint, [vp, vp, i64, i64, vp, vp, vp, vp, vp, vp]),
    "msamd_pca_plan_tiles": (None, [vp, vp]),
    "msamd_pca_plan_destroy": (None, [vp]),
    "msamd_pca_plan_workspace_bytes": (i64, [vp]),
    "msamd_pca_plan_attach": (cint, [vp, vp, i64, vp]),
    "msamd_pca_subsets_workspace_size": (i64, [i32, vp, i32, i32, i64]),
    "msamd_pca_subsets": (cint, [vp, i32, i64, vp, vp, i32, i32, i64, vp, vp, i64, vp]),
    "msamd_pca_autotune_workspace_size": (i64, [i32, vp, vp, i32, i32, i32, i64]),
    "msamd_pca_autotune": (cint, [vp, i32, vp, vp, i32, i32, i32, i64, vp, vp, vp, i64, vp, vp,
                                  vp]),
    "msamd_pca_autotune_timeseries": (cint, [vp, i64, i64, vp, vp, vp, i32, vp, vp, i32, i32,
                                             i32, i64, vp, vp, vp, i64, vp, vp, vp]),
    "msamd_whiten_covariance_workspace_size": (i64, [i64, i64]),
    "msamd_whiten_covariance": (cint, [vp, i64, i64, vp, vp, i64, vp]),
    "msamd_whiten_covariance_rows": (cint, [vp, i64, i64, vp, vp, vp, i64, vp]),
    "msamd_whiten_covariance_accumulate": (cint, [vp, i64, i64, vp, vp, cint, vp, i64, vp]),
    "msamd_whiten_covariance_finish": (cint, [vp, i64, i64, vp]),
    "msamd_whitening_matrix_workspace_size": (i64, [i64]),
    "msamd_whitening_matrix": (cint, [vp, i64, vp, vp, i64, vp]),
    "msamd_whitening_matrix_exact_workspace_size": (i64, [i64]),
    "msamd_whitening_matrix_exact": (cint, [vp, i64, vp, vp, i64, vp, vp, vp]),
    "msamd_whiten_apply_workspace_size": (i64, [i64]),
    "msamd_whiten_apply": (cint, [vp, i64, i64, vp, vp, f64, vp, vp, i64, vp]),
    "msamd_whiten_apply_rows": (cint, [vp, i64, i64, vp, vp, vp, f64, vp, vp, i64, vp]),
    "msamd_whiten_clips": (cint, [vp, i64, i64, i64, vp, vp, f64, vp, vp, i64, vp]),
    "msamd_whiten_apply_cov": (cint, [vp, i64, i64, vp, vp, vp, vp, f64, vp, vp, i64, vp]),
    "msamd_whiten_apply_cov_noalloc": (cint, [vp, i64, i64, vp, vp, vp, vp, f64, vp, vp]),
    "msamd_mask_workspace_size": (i64, [i64, i64, i64]),
    "msamd_mask_out_artifacts": (cint, [vp, i64, i64, f64, i64, vp, vp, vp, i64, vp]),
    "msamd_mask_rows": (cint, [vp, i64, i64, f64, i64, vp, vp, vp, i64, vp]),
    "msamd_templates_workspace_size": (i64, [i64, i32]),
    "msamd_compute_templates": (cint, [vp, i64, i64, vp, vp, i64, i32, i32, vp, vp, i64, vp]),
    "msamd_fit_stage_workspace_size": (i64, [i64, i64, vp, i64, i32, i32, i64, i64]),
    "msamd_fit_stage": (cint, [vp, i64, i64, vp, vp, i64, vp, i32, i32, f64, i64, i64, vp, vp,
                               vp, i64, vp]),
    "msamd_time_chunk_size": (i64, [i64, i64, i64, i64]),
    "msamd_tcn_num_params": (i32, []),
    "msamd_tcn_init": (cint, [vp, ctypes.c_uint64, vp]),
    "msamd_tcn_labels": (cint, [vp, vp, i32, i64, vp, vp, vp]),
    "msamd_tcn_train_workspace_size": (i64, [i32, i32]),
    "msamd_tcn_grad": (cint, [vp, i64, i64, vp, vp, vp, vp, i32, i32, f64, vp, vp, vp, i64, vp]),
    "msamd_tcn_train": (cint, [vp, i64, i64, vp, vp, i64, vp, vp, i32, i32, i32, i32, f64, f64,
                               ctypes.c_uint64, vp, vp, i64, vp]),
    "msamd_tcn_score_workspace_size": (i64, []),
    "msamd_tcn_score": (cint, [vp, i64, i64, vp, vp, vp, i64, vp]),
    "msamd_tcn_detect_workspace_size": (i64, [i64, i64, i64, i32, f64]),
    "msamd_tcn_detect": (cint, [vp, i64, i64, i64, i64, i64, i64, vp, f64, f64, cint, f64, f64,
                                vp, i64, ctypes.POINTER(i64), vp, vp]),
}


class MsamdLibraryMissing(RuntimeError):
    pass


ERR_TIMEOUT = 6  # MSAMD_ERR_TIMEOUT


class MsamdError(RuntimeError):
    def __init__(self, name, code, msg):
        super().__init__(f"{name} failed with code {code}: {msg}")
        self.code = code


def lib():
    """Load libmsamd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MsamdLibraryMissing(
                f"{LIB_PATH} not found: build it with `python -m mountainsort_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call an int-status entry point and raise MsamdError on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.msamd_last_error()
        raise MsamdError(name, rc, msg.decode() if msg else "")
    return rc
