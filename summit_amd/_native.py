"""ctypes binding of libsummit_gpe.so (C ABI in include/summit_gpe.h).

The library is built in-tree by ``summit_amd.build`` (``__graft_entry__.build``).
There is no CPU fallback: a missing library raises at first use.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsummit_gpe.so")

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_char_p = ctypes.c_char_p

# name -> (restype, argtypes); every symbol of include/summit_gpe.h
SIGNATURES = {
    "gpe_last_error": (c_char_p, []),
    "gpe_version": (c_char_p, []),
    "gpe_set_device": (ctypes.c_int, [ctypes.c_int]),
    "gpe_geno_create": (ctypes.c_int, [c_i64, c_i64, ctypes.POINTER(c_vp)]),
    "gpe_geno_upload_bed": (ctypes.c_int, [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp]),
    "gpe_geno_synthesize": (ctypes.c_int, [c_vp, c_u64, c_dbl, c_dbl, c_dbl, c_i32, c_dbl, c_vp]),
    "gpe_geno_prepare": (ctypes.c_int, [c_vp, c_i32, c_i32, c_u64, c_i64, c_vp, c_vp]),
    "gpe_geno_stats": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "gpe_geno_download": (ctypes.c_int, [c_vp, c_i32, c_vp, c_i64]),
    "gpe_geno_download_rows": (ctypes.c_int, [c_vp, c_i32, c_i64, c_i64, c_vp, c_i64]),
    "gpe_geno_shape": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "gpe_geno_free": (ctypes.c_int, [c_vp]),
    "gpe_gz_i32": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp, c_i64, c_i32, c_vp]),
    "gpe_gtu_i64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp]),
    "gpe_resvar_inv": (ctypes.c_int, [c_vp, c_i64, c_i64, c_dbl, c_vp, c_vp]),
    "gpe_phase1_xz": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i32, c_vp, c_i32, c_i32, c_i32, c_u64,
                                     c_vp, c_i64, c_vp, c_i32, c_vp]),
    "gpe_phase1_xz_tiled": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i32, c_vp, c_i32, c_i32, c_i32, c_i32,
                                           c_u64, c_vp, c_i64, c_vp, c_i32, c_vp]),
    "gpe_phase1_xz_indirect": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i64, c_vp, c_i32, c_vp, c_vp, c_i32, c_i32,
                                              c_u64, c_vp, c_i64, c_vp, c_i32, c_vp]),
    "gpe_add_i32": (ctypes.c_int, [c_vp, c_i32, c_vp]),
    "gpe_phase2_accum": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i32, c_vp, c_i64, c_i32, c_i32, c_vp, c_i64,
                                        c_vp, c_vp, c_dbl, c_i32, c_i32, c_vp, c_vp]),
    "gpe_phase2_accum_colsums": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i32, c_vp, c_i64, c_i32, c_vp, c_i32,
                                                c_vp, c_i64, c_vp, c_vp, c_dbl, c_i32, c_i32, c_vp, c_vp]),
    "gpe_probes": (ctypes.c_int, [c_i32, c_u64, c_vp, c_i32, c_i64, c_i32, c_i32, c_vp]),
    "gpe_philox_rademacher": (ctypes.c_int, [c_vp, c_vp, c_i32, c_i64, c_i32, c_vp, c_i64, c_vp]),
    "gpe_phase1_xz_z8": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i32, c_vp, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp,
                                        c_i32, c_vp]),
    "gpe_phase1_xz_f64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i32, c_vp, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp,
                                         c_i32, c_vp]),
    "gpe_project_f64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_i32, c_vp]),
    "gpe_project_f64_noalloc": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32, c_i32,
                                               c_vp]),
    "gpe_col_sums_f64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i32, c_vp, c_vp]),
    "gpe_gxe_direct_features": (ctypes.c_int, [c_vp, c_i64, c_vp, c_i64, c_i32, c_vp, c_vp, c_vp,
                                               ctypes.c_double, c_i64, c_vp, c_i64, c_vp]),
    "gpe_basis_coef_f64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_i32, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp,
                                          c_vp]),
    "gpe_weighted_sumsq": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i32, c_i64, c_vp, c_i64, c_vp]),
    "gpe_sq_rowsums_f64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp, c_i64, c_vp, c_i64, c_i32,
                                          c_vp, c_i64, c_vp]),
    "gpe_windowed_ld": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_dbl, c_i64, c_vp, c_i32, c_vp, c_vp]),
    "gpe_windowed_ld_multi": (ctypes.c_int, [c_vp, c_i32, c_vp, c_vp, c_dbl, c_i64, c_vp, c_i32, c_vp, c_vp]),
    "gpe_windowed_ld_side": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_dbl, c_i64, c_i32, c_vp, c_i32, c_vp, c_vp]),
    "gpe_windowed_ld_cov": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_dbl, c_i64, c_vp, c_i32, c_vp, c_i64, c_vp,
                                           c_i64, c_i32, c_vp, c_vp]),
    "gpe_xc_f64":(ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp, c_i64, c_i32, c_i32, c_vp]),
    "gpe_xtv_f64": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_i64, c_i32, c_vp, c_i64, c_i32, c_vp]),
    "gpe_mlp_n_params": (ctypes.c_int, [ctypes.POINTER(c_i64)]),
    "gpe_mlp_create": (ctypes.c_int, [c_i32, c_i64, ctypes.POINTER(c_vp)]),
    "gpe_mlp_free": (ctypes.c_int, [c_vp]),
    "gpe_mlp_set_params": (ctypes.c_int, [c_vp, c_vp, c_i64]),
    "gpe_mlp_get_params": (ctypes.c_int, [c_vp, c_vp, c_i64]),
    "gpe_mlp_get_grad": (ctypes.c_int, [c_vp, c_vp, c_i64]),
    "gpe_mlp_step_count": (ctypes.c_int, [c_vp, ctypes.POINTER(c_i32)]),
    "gpe_mlp_forward": (ctypes.c_int, [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp]),
    "gpe_mlp_train_step": (ctypes.c_int, [c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, c_i32,
                                          c_vp, c_vp]),
    "gpe_schedule_set": (ctypes.c_int, [c_i32, c_i64, c_i64, c_i32, c_i32, c_i32]),
    "gpe_schedule_get": (ctypes.c_int, [c_i32, c_i64, c_i64, c_i32, c_i32, ctypes.POINTER(c_i32)]),
    "gpe_schedule_split_rule": (ctypes.c_int, [c_i32, c_i64, c_i64, c_i64, c_i64, c_vp]),
    "gpe_schedule_clear": (ctypes.c_int, []),
    "gpe_profile_enable": (ctypes.c_int, [c_i32]),
    "gpe_profile_read": (ctypes.c_int, [c_i32, ctypes.POINTER(c_dbl), ctypes.POINTER(c_i64)]),
    "gpe_profile_reset": (ctypes.c_int, []),
}

_LIB = None


class GpeError(RuntimeError):
    """Raised when a libsummit_gpe entry point reports failure."""


def load() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: the HIP engine has not been built "
            "(run `python -c 'import __graft_entry__ as g; g.build()'`); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().gpe_last_error().decode("utf-8", "replace")
        raise GpeError(msg or f"libsummit_gpe status {status}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))
