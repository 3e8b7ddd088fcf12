"""MI355X-native catalog hot path of galfast.

The compute lives in ``libgalfast_hip.so`` (HIP kernels for gfx950 plus the
C++ host layer) and is reached through the C ABI of ``include/galfast_hip.h``.
This package loads that library, mirrors its parameter blocks in ctypes and
offers small helpers for structure-of-arrays column buffers.  There is no
CPU fallback: importing :func:`lib` without the built library raises.
"""
import ctypes as C
import os

# A catalog run keeps four HIP streams busy at once (plan and draw passes,
# emission, two row-pass streams).  HIP maps a process's streams onto its
# hardware queues, four by default with one taken by the null stream; two
# streams that share a queue run strictly one after the other, so the
# emission of the next chunk would wait behind the kinematics passes of the
# current one.  Eight queues give every stream its own.  HIP reads this
# once, when it initialises in the process: it takes effect only when the
# package is imported before the first GPU call, and a value the
# environment already sets is left as it is.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

from . import _abi
from ._abi import *  # noqa: F401,F403

HERE = os.path.dirname(os.path.abspath(__file__))
# GALFAST_LIB names another build of the library (e.g. one built with
# different compile options) to load instead of the in-tree one
LIB_PATH = os.environ.get("GALFAST_LIB") or os.path.join(HERE, "libgalfast_hip.so")
DATA_DIR = os.path.join(HERE, "data")

_lib = None


class GalfastError(RuntimeError):
    pass


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GalfastError(
                "libgalfast_hip.so is not built; run `python -c \"import __graft_entry__ as g; g.build()\"`")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = C.POINTER
    vp = C.c_void_p
    L.gf_strerror.restype = C.c_char_p
    L.gf_strerror.argtypes = [C.c_int]
    L.gf_seed_from_pipeline_seed.restype = C.c_uint32
    L.gf_seed_from_pipeline_seed.argtypes = [C.c_uint32]
    L.gf_mwc_seed_host.argtypes = [C.c_uint32, P(C.c_uint32), C.c_uint32, C.c_uint32, P(C.c_uint32)]
    L.gf_mwc_seed_device.argtypes = [C.c_uint32, vp, C.c_uint32, C.c_uint32, vp, vp]
    L.gf_rng_alloc.argtypes = [P(gf_rng), C.c_uint32]
    L.gf_rng_upload.argtypes = [P(gf_rng), P(C.c_uint32), vp]
    L.gf_rng_download.argtypes = [P(gf_rng), P(C.c_uint32), vp]
    L.gf_rng_free.argtypes = [P(gf_rng)]
    L.gf_mwc_sample.argtypes = [P(gf_rng), C.c_int, C.c_float, C.c_uint32, vp, vp]
    L.gf_os_FeH_kernel.argtypes = [P(gf_rows), P(gf_FeH_data), P(gf_rng), gf_col, gf_col, gf_col, gf_col, vp]
    L.gf_os_GaussianFeH_kernel.argtypes = [P(gf_rows), gf_bitmap, P(gf_GaussianFeH_data), P(gf_rng),
                                           gf_col, gf_col, gf_col, gf_col, gf_col, vp]
    L.gf_os_Bond2010_kernel.argtypes = [P(gf_rows), P(gf_Bond2010_data), P(gf_rng), gf_col, gf_col, gf_col,
                                        gf_col, vp]
    L.gf_os_vel2pm_kernel.argtypes = [P(gf_rows), P(gf_vel2pm_data), gf_col, gf_col, gf_col, gf_col, gf_col, vp]
    L.gf_os_gal2other_kernel.argtypes = [P(gf_rows), C.c_int, gf_col, gf_col, gf_col, vp]
    L.gf_os_photometry_kernel.argtypes = [P(gf_rows), gf_bitmap, P(gf_photometry_data), gf_col, gf_col,
                                          gf_col, gf_col, C.c_int, gf_col, gf_col, gf_col, gf_col, vp]
    L.gf_os_unresolvedMultiples_kernel.argtypes = [P(gf_rows), gf_bitmap, P(gf_multiples_data), P(gf_rng),
                                                   C.c_int, gf_col, gf_col, gf_col, gf_col, gf_col, vp]
    L.gf_photometric_errors_kernel.argtypes = [P(gf_rows), P(gf_photoerr_data), P(gf_rng), gf_col, gf_col,
                                               gf_col, vp]
    L.gf_photoerr_records.argtypes = [C.c_int32, vp, vp, vp, vp, vp, vp]
    L.gf_module_chain.argtypes = [P(gf_rows), P(gf_chain), P(gf_rng), P(gf_table), vp]
    L.gf_skygen_count.argtypes = [P(gf_skygen_params), P(gf_model), vp, P(gf_extinction), C.c_int,
                                  P(gf_count_result), vp, vp, vp]
    L.gf_skygen_draw_compat.argtypes = [P(gf_skygen_params), P(gf_model), vp, P(gf_extinction),
                                        P(gf_footprint), P(gf_rng), P(gf_skygen_cols), C.c_uint64,
                                        P(C.c_uint64), vp]
    L.gf_skygen_order.argtypes = [P(gf_skygen_cols), C.c_uint64, P(gf_skygen_cols), P(C.c_uint64), vp]
    L.gf_clip_footprint.argtypes = [P(gf_skygen_params), P(gf_footprint), gf_col, gf_col, gf_col, C.c_uint64, vp]
    L.gf_footprint_cover.argtypes = [P(gf_footprint), C.c_double, C.c_int32, vp, vp]
    L.gf_countsmap_bin.argtypes = [P(gf_countsmap_params), gf_col, gf_col, gf_col, C.c_uint64, vp, vp, vp, vp]
    L.gf_countsmap_finish.argtypes = [P(gf_countsmap_params), vp, vp]
    L.gf_skygen_plan.argtypes = [P(gf_skygen_params), P(gf_model), vp, P(gf_extinction), vp, P(gf_rng),
                                 P(vp), P(C.c_uint64), vp]
    L.gf_skygen_emit.argtypes = [vp, C.c_uint64, C.c_uint64, P(gf_footprint), P(gf_rng),
                                 P(gf_skygen_cols), vp]
    L.gf_skygen_plan_free.argtypes = [vp]
    for name in ("gf_init", "gf_sync", "gf_malloc", "gf_free"):
        getattr(L, name).restype = C.c_int
    L.gf_init.argtypes = [C.c_int]
    L.gf_sync.argtypes = [vp]
    L.gf_malloc.argtypes = [P(vp), C.c_size_t]
    L.gf_free.argtypes = [vp]


def check(rc, what="galfast call"):
    """Raise GalfastError for a non-zero return code."""
    if rc != 0:
        msg = lib().gf_strerror(rc).decode()
        raise GalfastError("%s failed: %s (%d)" % (what, msg, rc))
    return rc


def exported_symbols():
    """Names of every function declared in include/galfast_hip.h."""
    import re
    hdr = os.path.join(HERE, "..", "include", "galfast_hip.h")
    text = open(hdr).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(gf_[a-zA-Z0-9_]+)\s*\(", text)
    return sorted(set(names))
