"""ctypes binding of liblcf_hip.so (the C ABI declared in include/lcf_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = _PKG_DIR / "liblcf_hip.so"

_i32, _i64, _u32, _u64, _f64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
_pd = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)
_pi64 = C.POINTER(C.c_int64)
_vp = C.c_void_p
_ppv = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/lcf_hip.h.
SIGNATURES = {
    "lcf_ctx_create": (_i32, [_i32, _ppv]),
    "lcf_ctx_destroy": (None, [_vp]),
    "lcf_last_error": (C.c_char_p, [_vp]),
    "lcf_ctx_set_precision": (_i32, [_vp, _i32]),
    "lcf_ctx_set_schedule": (_i32, [_vp, _i32]),
    "lcf_ctx_num_cus": (_i32, [_vp]),
    "lcf_ctx_synchronize": (_i32, [_vp]),
    "lcf_ctx_autotune_pso": (_i32, [_vp, _vp, _pi32, _i32, _i32, _i32, _i32, _u64]),
    "lcf_ctx_get_tuning": (_i32, [_vp, _pi32, _i32]),
    "lcf_ctx_set_tuning": (_i32, [_vp, _pi32, _i32]),
    "lcf_batch_create": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _ppv]),
    "lcf_batch_create_current": (_i32, [_i64, _pi64, _pd, _pd, _pd, _pi32, _pd, _ppv]),
    "lcf_batch_destroy": (None, [_vp]),
    "lcf_batch_n_sources": (_i64, [_vp]),
    "lcf_batch_n_obs": (_i64, [_vp, _i64]),
    "lcf_eval_batch": (_i32, [_vp, _i32, _pd, _i64, _pd, _i64, _pd]),
    "lcf_batch_pso": (_i32, [_vp, _i32, _vp, _i32, _i32, _i32, _u64, _pd, _pd]),
    "lcf_batch_model_select": (_i32, [_vp, _vp, _i32, _i32, _i32, _f64, _pi32, _pd, _pd]),
    "lcf_batch_all_models": (_i32, [_vp, _vp, _i32, _i32, _i32, _pd, _pd]),
    "lcf_batch_fit_models": (_i32, [_vp, _vp, _pi32, _i32, _i32, _f64, _i32, _i32, _i32, _u64,
                                    _u32, _pd, _pd, _pi32, _pi32, _pd, _pd]),
    "lcf_batch_fit_models_noalloc": (_i32, [_vp, _vp, _pi32, _i32, _i32, _f64, _i32, _i32, _i32, _u64,
                                            _u32, _pd, _pd, _pi32, _pi32, _pd, _pd]),
    "lcf_batch_multi_bazin": (_i32, [_vp, _vp, _i32, _i32, _i32, _u64, _pi32, _pd, _pd, _pd, _pd, _pd]),
    "lcf_batch_bb_temperature": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _pi32]),
    "lcf_batch_thermal": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _i32,
                                 _u64, _pi32, _pd, _pi32]),
    "lcf_batch_nonparametric": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _i32, _pd, _i32,
                                       _i32, _i32, _i32, _pi32, _pd, _pd, _pd, _pd, _pd]),
    "lcf_batch_svi_fit": (_i32, [_vp, _vp, _pi32, _pd, _pi32, _pd, _pd, _i32, _i32, _f64, _pd, _pd, _pd]),
    "lcf_batch_svi_fit_ex": (_i32, [_vp, _vp, _pi32, _pd, _pd, _pd, _i32, _i32, _f64, _u64, _u32,
                                    _pd, _pd, _pd, _pi32]),
    "lcf_svi_elbo_grad": (_i32, [_vp, _vp, _pi32, _pd, _pd, _pd, _pd, _i32, _u64, _i32, _u32,
                                 _pd, _pd, _pd]),
    "lcf_batch_gp_fit": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _i32, _pd, _i32, _pd, _i32, _i32,
                                _pd, _pd, _pd, _pd]),
    "lcf_batch_gp_fit_ex": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _i32, _i32, _pd, _i32, _pd, _i32,
                                   _i32, _i32, _i32, _pd, _pd, _pd, _pd, _pd]),
    "lcf_batch_laplace": (_i32, [_vp, _vp, _pi32, _pd, _i32, _pd, _pd, _pd]),
    "lcf_batch_fit_gp_2d": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _pd, _pi32, _pd,
                                   _pd]),
    "lcf_batch_gp2d_train": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _i32, _f64, _pd, _pd, _i32, _i32,
                                    _pd, _pd, _pd, _pi32, _pi32, _pd, _pd, _pi32]),
    "lcf_batch_gp2d_train_ex": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _i32, _f64, _pd, _pd, _i32,
                                       _i32, _u32, _pd, _pd, _pd, _pi32, _pi32, _pd, _pd, _pi32]),
    "lcf_gp2d_nlml_grad": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi32]),
    "lcf_cpu_pso": (_i32, [_vp, _vp, _vp, _i64, _pd, _pd, _pi32]),
    "lcf_batch_finalize": (_i32, [_vp, _vp, _pi32, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _pd, _pd]),
    "lcf_batch_gp2d_fit": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _i32, _pd, _i32, _pd, _i32,
                                  _pd, _i32, _i32, _pd, _pd, _pd, _pd]),
    "lcf_batch_gp2d_fit_ex": (_i32, [_vp, _i64, _pi64, _pd, _pd, _pd, _pd, _pd, _pd, _i32, _i32, _pd, _i32,
                                     _pd, _i32, _pd, _i32, _i32, _i32, _i32, _pd, _pd, _pd, _pd, _pd]),
    "lcf_last_kernel_ms": (_f64, [_vp]),
    "lcf_last_model_ms": (_f64, [_vp, _i32]),
}

_LIB = None


def load(path: Path | str | None = None) -> C.CDLL:
    """Load (once) and return the native library with typed signatures."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"{p} is not built; run `python lightcurve-fitting_amd/build.py` "
            "(there is no CPU fallback)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array's data (None for None)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))
