"""Loader for the oracle (oracle/_build/libgforacle.so).  Test infrastructure."""
import ctypes as C
import os
import subprocess

import numpy as np

from galfast_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.environ.get("GFO_LIB") or os.path.join(ROOT, "oracle", "_build", "libgforacle.so")

_o = None


def lib():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
        _o = C.CDLL(ORACLE_SO)
        P = C.POINTER
        vp = C.c_void_p
        _o.gfo_seed_from_pipeline_seed.restype = C.c_uint32
        _o.gfo_mwc_seed.restype = C.c_int
        _o.gfo_model_rho.restype = C.c_float
        _o.gfo_model_rho.argtypes = [P(A.gf_model), C.c_float, C.c_float, C.c_float]
        _o.gfo_lf_norm_f.restype = C.c_float
        _o.gfo_lf_norm_f.argtypes = [A.gf_model, C.c_float, C.c_float, C.c_float, C.c_float]
        _o.gfo_spline.restype = C.c_double
        _o.gfo_spline.argtypes = [vp, vp, C.c_int, C.c_double]
        _o.gfo_tex1d.restype = C.c_float
        _o.gfo_tex1d.argtypes = [vp, C.c_int, vp, C.c_float]
        _o.gfo_draw_compat.restype = C.c_int64
        _o.gfo_draw_compat.argtypes = [P(A.gf_skygen_params), P(A.gf_model), vp, P(A.gf_extinction),
                                       P(A.gf_footprint), vp, C.c_uint32, P(A.gf_skygen_cols), C.c_uint64]
        _o.gfo_count_crawl.argtypes = [P(A.gf_skygen_params), P(A.gf_model), vp, P(A.gf_extinction),
                                       P(A.gf_count_result), vp]
        _o.gfo_count_bins.argtypes = [P(A.gf_skygen_params), P(A.gf_model), vp, P(A.gf_extinction),
                                      P(A.gf_count_result), vp, vp]
        _o.gfo_plan.argtypes = [vp, C.c_uint64, vp, C.c_uint32, vp]
        _o.gfo_plan_advance.argtypes = [vp, C.c_uint32, C.c_uint64]
        _o.gfo_mwc_jump.argtypes = [vp, C.c_uint32, C.c_uint64]
        _o.gfo_mwc_step.argtypes = [vp, C.c_uint32, C.c_uint64]
        _o.gfo_emit.argtypes = [P(A.gf_skygen_params), P(A.gf_model), vp, P(A.gf_extinction),
                                P(A.gf_footprint), vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, C.c_uint32,
                                P(A.gf_skygen_cols)]
        _o.gfo_draw_window.restype = C.c_uint64
        _o.gfo_draw_window.argtypes = [P(A.gf_skygen_params), P(A.gf_model), vp, P(A.gf_extinction),
                                       P(A.gf_footprint), vp, C.c_uint32, C.c_uint64, C.c_uint64,
                                       P(A.gf_skygen_cols)]
        _o.gfo_order.restype = C.c_uint64
        _o.gfo_order.argtypes = [P(A.gf_skygen_cols), C.c_uint64, P(A.gf_skygen_cols)]
        _o.gfo_clip.argtypes = [P(A.gf_skygen_params), P(A.gf_footprint), P(A.gf_skygen_cols), C.c_uint64]
        _o.gfo_countsmap_bin.restype = C.c_uint64
        _o.gfo_countsmap_bin.argtypes = [P(A.gf_countsmap_params), A.gf_col, A.gf_col, A.gf_col, C.c_uint64, vp, vp]
        _o.gfo_module_chain.restype = C.c_int
        _o.gfo_module_chain.argtypes = [P(A.gf_rows), P(A.gf_chain), vp, C.c_uint32, P(A.gf_table)]
        _o.gfo_galequ.argtypes = [C.c_double, C.c_double, P(C.c_double), P(C.c_double)]
        _o.gfo_outside_footprint.argtypes = [P(A.gf_footprint), C.c_double, C.c_double]
        _o.gfo_footprint_cover.argtypes = [P(A.gf_footprint), C.c_double, C.c_int, C.c_void_p]
        _o.gfo_vcyl2pm.argtypes = [P(A.gf_vel2pm_data)] + [C.c_float] * 8 + [vp]
        _o.gfo_deproject.argtypes = [P(A.gf_lambert), C.c_float, C.c_float, P(C.c_double), P(C.c_double)]
        _o.gfo_resample_clip.argtypes = [vp, vp, C.c_int, C.c_float, C.c_int, vp, vp]
        _o.gfo_resample.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp]
        _o.gfo_photometry_tables.restype = C.c_int
        _o.gfo_photometry_tables.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, vp, vp, C.c_int,
                                             P(C.c_int), P(C.c_int), vp]
        for fn in ("gfo_math_f", "gfo_libm_f"):
            getattr(_o, fn).argtypes = [C.c_int, vp, vp, vp, C.c_size_t]
        _o.gfo_mwc_sample.argtypes = [vp, C.c_uint32, C.c_int, C.c_float, C.c_uint32, vp]
        _o.gfo_mt19937.argtypes = [C.c_uint32, C.c_uint32, vp]
        _o.gfo_rand48.argtypes = [C.c_uint32, C.c_uint32, vp, vp]
        _o.gfo_mwc_seed.argtypes = [C.c_uint32, vp, C.c_uint32, C.c_uint32, vp]
    return _o


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def math_f(fn, x, y=None, libm=False):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), dtype=np.float32)
    out = np.empty_like(x)
    f = lib().gfo_libm_f if libm else lib().gfo_math_f
    f(fn, ptr(x), ptr(y), ptr(out), x.size)
    return out
