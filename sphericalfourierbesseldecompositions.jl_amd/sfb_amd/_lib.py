"""ctypes binding of libsfb_mi355x.so (declarations in include/sfb.h).

The library must have been built (``python sphericalfourierbesseldecompositions.jl_amd/build.py``);
there is no fallback path: a missing library raises at first use.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsfb_mi355x.so")

SFB_OK = 0
SFB_ERR_ARG = 1
SFB_ERR_LMAX = 2
SFB_ERR_NONFINITE = 3
SFB_ERR_SETUP = 4
SFB_ERR_DEVICE = 5
SFB_ERR_NOMEM = 6

SFB_PREC_FP64 = 0
SFB_PREC_MIXED = 1
SFB_PREC_BF16X3 = 2
SFB_PREC_BF16X2 = 3
SFB_WIN_DENSE = 0
SFB_WIN_SEPARABLE = 1

# every symbol include/sfb.h declares, with (restype, argtypes)
_i64, _i32, _dbl, _P = ctypes.c_int64, ctypes.c_int32, ctypes.c_double, ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "sfb_last_error": (ctypes.c_char_p, []),
    "sfb_version": (ctypes.c_char_p, []),
    "sfb_modes_create_kmax": (ctypes.c_int, [_PP, _dbl, _dbl, _dbl, _i64]),
    "sfb_modes_create_nl": (ctypes.c_int, [_PP, _i64, _i64, _dbl, _dbl, _i64]),
    "sfb_modes_sizes": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "sfb_modes_tables": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "sfb_modes_destroy": (None, [_P]),
    "sfb_spline_build": (ctypes.c_int, [_i64, _P, _P, _P]),
    "sfb_spline_eval": (ctypes.c_int, [_i64, _P, _P, _P, _i64, _P, _P, _i32]),
    "sfb_plan_create": (ctypes.c_int, [_PP, _i64, _i64, _i64, _P, _P, _P, _P, _P, _i32, _i32, _i32]),
    "sfb_plan_destroy": (None, [_P]),
    "sfb_plan_sizes": (ctypes.c_int, [_P, _P, _P, _P]),
    "sfb_plan_set_precision": (ctypes.c_int, [_P, _i32]),
    "sfb_cat2amln": (ctypes.c_int, [_P, _i64, _P, _P, _dbl, _i32, _P, _P, _P]),
    "sfb_cat2amln_device": (ctypes.c_int, [_P, _i64, _P, _P, _dbl, _i32, _P, _P, _P, _P]),
    "sfb_cat2amln_batch": (ctypes.c_int, [_P, _i32, _P, _P, _P, _P, _i32, _P, _P, _P]),
    "sfb_cat2amln_batch_device": (ctypes.c_int, [_P, _i32, _P, _P, _P, _P, _i32, _P, _P, _P, _P]),
    "sfb_plan_set_batch_graph": (ctypes.c_int, [_P, _i32, _P]),
    "sfb_win_rhat_ln": (ctypes.c_int, [_P, _i64, _P, _dbl, _i32, _P, _P]),
    "sfb_win_rhat_ln_device": (ctypes.c_int, [_P, _i64, _P, _dbl, _i32, _P, _P, _P]),
    "sfb_field2anlm": (ctypes.c_int, [_P, _i64, _P, _dbl, _P, _P]),
    "sfb_field2anlm_device": (ctypes.c_int, [_P, _i64, _P, _dbl, _P, _P, _P]),
    "sfb_anlm2field": (ctypes.c_int, [_P, _i64, _P, _P, _P]),
    "sfb_anlm2field_device": (ctypes.c_int, [_P, _i64, _P, _P, _P, _P]),
    "sfb_xyz2rtp": (ctypes.c_int, [_i32, _i64, _P, _P]),
    "sfb_winweights2galweights": (ctypes.c_int, [_i32, _i64, _i64, _P, _P, _i64, _P, _P]),
    "sfb_apply_window": (ctypes.c_int, [_i32, _i64, _i64, _dbl, _dbl, _dbl, _P, _i64, _P, _P, _P, _P]),
    "sfb_apply_window_noalloc": (ctypes.c_int, [_i32, _i64, _i64, _dbl, _dbl, _dbl, _P, _i64, _P, _P, _P, _P, _P]),
    "sfb_amln2clnn": (ctypes.c_int, [_P, _i64, _P, _P, _P, _P]),
    "sfb_amln2clnn_cross": (ctypes.c_int, [_P, _P, _i64, _P, _P, _P, _P]),
    "sfb_pixwin": (ctypes.c_int, [_i64, _i64, _i32, _i32, _P]),
    "sfb_cat2amln_vjp": (ctypes.c_int, [_P, _i64, _P, _P, _dbl, _P, _P, _P, _P]),
    "sfb_cat2amln_vjp_device": (ctypes.c_int, [_P, _i64, _P, _P, _dbl, _P, _P, _P, _P]),
    "sfb_ang2pix_ring": (ctypes.c_int, [_i32, _i64, _i64, _P, _P, _P]),
    "sfb_delta_map": (ctypes.c_int, [_P, _i64, _P, _P, _dbl, _i32, _P, _P, _i64, _P]),
    "sfb_map2alm": (ctypes.c_int, [_P, _i64, _P, _P]),
    "sfb_gemm_f64": (ctypes.c_int, [_i32, _i64, _i64, _i64, _P, _P, _P]),
    "sfb_gemm_f32": (ctypes.c_int, [_i32, _i64, _i64, _i64, _P, _P, _P]),
    "sfb_gemm_bf16": (ctypes.c_int, [_i32, _i64, _i64, _i64, _P, _P, _P, _i32]),
    "sfb_gemm_f32_stored": (ctypes.c_int, [_i32, _i64, _i64, _i64, _i64, _P, _P, _P, _P]),
    "sfb_phi_transform": (ctypes.c_int, [_P, _i64, _P, _P]),
    "sfb_ring_kernel_shapes": (ctypes.c_int, [_i64, _i64, _i64, _P]),
    "sfb_plan_profile": (ctypes.c_int, [_P, _i32]),
    "sfb_plan_phase_times": (ctypes.c_int, [_P, _P, _P, _P]),
    "sfb_plan_tiles": (ctypes.c_int, [_P, _P, _P]),
}

PHASES = ["catalog", "scatter", "phi", "analysis", "synthesis", "residual", "pack", "segment", "monopole", "adjoint", "phi_polar"]

_lib = None


class SFBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[sfb error {code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python sphericalfourierbesseldecompositions.jl_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != SFB_OK:
        msg = lib().sfb_last_error().decode(errors="replace")
        if rc == SFB_ERR_NONFINITE:
            raise AssertionError(msg)
        raise SFBError(rc, msg)


def ptr(a):
    """Host data pointer of a contiguous numpy array (None stays NULL)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)
