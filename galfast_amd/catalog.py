"""Python face of the catalog pipeline (``gf_catalog_*`` in the C ABI).

    cat = Catalog("cmd.conf", seed=42)
    expected = cat.count()
    stats = cat.run(on_chunk)     # on_chunk(table: gf_table, nrows, comp)

mirrors ``galfast catalog cmd.conf -s seed`` (pipeline.cpp:1313-1445): the
count pass, then draw + module chain chunk by chunk, rows delivered in
device memory.
"""
import ctypes as C

from . import lib, check, DATA_DIR, GalfastError
from ._abi import gf_table, gf_count_result

CHUNK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(gf_table), C.c_uint64, C.c_int)


class gf_catalog_info_t(C.Structure):
    _fields_ = [("npixels", C.c_uint64), ("nm", C.c_int32), ("nM", C.c_int32), ("nmodels", C.c_int32),
                ("nbands", C.c_int32), ("nabsmag", C.c_int32), ("pad", C.c_int32),
                ("footprint_covered_sr", C.c_double)]


class gf_run_stats(C.Structure):
    _fields_ = [("expected", C.c_double), ("generated", C.c_uint64), ("rows_chain", C.c_uint64),
                ("n_emit", C.c_uint64), ("n_chain", C.c_uint64), ("t_plan", C.c_double),
                ("t_emit", C.c_double), ("t_chain", C.c_double), ("t_total", C.c_double),
                ("visible", C.c_uint64), ("t_stage", C.c_double * 6),
                ("t_emit_stars", C.c_double), ("t_emit_units", C.c_double)]


def _declare():
    L = lib()
    if getattr(L, "_gf_catalog_declared", False):
        return L
    P = C.POINTER
    L.gf_last_error.restype = C.c_char_p
    L.gf_catalog_open.argtypes = [C.c_char_p, C.c_uint32, C.c_char_p, P(C.c_void_p)]
    L.gf_catalog_set.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
    L.gf_catalog_info.argtypes = [C.c_void_p, P(gf_catalog_info_t)]
    L.gf_catalog_count.argtypes = [C.c_void_p, P(gf_count_result), P(C.c_double)]
    L.gf_catalog_run.argtypes = [C.c_void_p, CHUNK_FN, C.c_void_p, P(gf_run_stats)]
    L.gf_catalog_close.argtypes = [C.c_void_p]
    L.gf_catalog_close.restype = None
    L.gf_catalog_component_ids.argtypes = [C.c_void_p, P(C.c_uint32), C.c_int]
    L.gf_writer_open.argtypes = [C.c_void_p, C.c_char_p, C.c_int, P(C.c_void_p)]
    L.gf_writer_close.argtypes = [C.c_void_p, P(C.c_uint64)]
    L._gf_catalog_declared = True
    return L


class Catalog:
    def __init__(self, cmd_conf, seed=42, datadir=DATA_DIR, chunk_rows=None):
        L = _declare()
        self._h = C.c_void_p()
        rc = L.gf_catalog_open(cmd_conf.encode(), seed, datadir.encode(), C.byref(self._h))
        if rc != 0:
            msg = L.gf_last_error().decode() if rc == -4 else L.gf_strerror(rc).decode()
            raise GalfastError("gf_catalog_open(%s): %s" % (cmd_conf, msg))
        if chunk_rows:
            self.set("chunk_rows", chunk_rows)

    def set(self, key, value):
        check(lib().gf_catalog_set(self._h, key.encode(), float(value)), "gf_catalog_set")

    def info(self):
        i = gf_catalog_info_t()
        check(lib().gf_catalog_info(self._h, C.byref(i)), "gf_catalog_info")
        return i

    def count(self):
        n = self.info().nmodels
        per = (gf_count_result * max(n, 1))()
        tot = C.c_double()
        check(lib().gf_catalog_count(self._h, per, C.byref(tot)), "gf_catalog_count")
        self.per_model = [per[i] for i in range(n)]
        return tot.value

    def run(self, on_chunk=None):
        st = gf_run_stats()
        if on_chunk is None:
            cb = CHUNK_FN()
        else:
            def _cb(user, table, n, comp):
                try:
                    on_chunk(table.contents, n, comp)
                    return 0
                except Exception as e:  # propagate as a failed run
                    self._cb_error = e
                    return -8
            cb = CHUNK_FN(_cb)
        self._cb_error = None
        rc = lib().gf_catalog_run(self._h, cb, None, C.byref(st))
        if self._cb_error is not None:
            raise self._cb_error
        check(rc, "gf_catalog_run")
        return st

    def component_ids(self):
        n = lib().gf_catalog_component_ids(self._h, None, 0)
        ids = (C.c_uint32 * max(n, 1))()
        lib().gf_catalog_component_ids(self._h, ids, n)
        return list(ids[:n])

    def write(self, path, kind="fits"):
        """Run the catalog into a file: the reference's output stage
        (kind "fits": os_fitsout, "text": os_textout, "countsmap": the job's
        os_countsMap stage) for the visible rows.  Returns (run statistics,
        rows written or binned)."""
        L = lib()
        w = C.c_void_p()
        check(L.gf_writer_open(self._h, path.encode(), {"text": 0, "fits": 1, "countsmap": 2}[kind], C.byref(w)), "gf_writer_open")
        st = gf_run_stats()
        rc = L.gf_catalog_run(self._h, C.cast(L.gf_writer_chunk, CHUNK_FN), w, C.byref(st))
        rows = C.c_uint64()
        rc2 = L.gf_writer_close(w, C.byref(rows))
        check(rc, "gf_catalog_run")
        check(rc2, "gf_writer_close")
        return st, rows.value

    def close(self):
        if self._h:
            lib().gf_catalog_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
