"""Python face of the request server (``gf_server_*`` in the C ABI)."""
import ctypes as C

from . import lib, check, DATA_DIR, GalfastError
from ._abi import gf_table


class gf_request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("l", C.c_double), ("b", C.c_double), ("radius", C.c_double),
                ("m1", C.c_float), ("seed", C.c_uint32), ("l1", C.c_double), ("b1", C.c_double)]


class gf_result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("batch_size", C.c_int32), ("npixels", C.c_uint64),
                ("generated", C.c_uint64), ("visible", C.c_uint64), ("expected", C.c_double),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
                ("device_bytes", C.c_uint64), ("batch", C.c_uint64)]


class gf_server_config(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_int32), ("workers", C.c_int32),
                ("plain", C.c_int32), ("linger_us", C.c_int32), ("eager_chain", C.c_int32), ("pad", C.c_int32)]


class gf_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("rows", C.c_uint64), ("visible", C.c_uint64), ("t_first_submit", C.c_double),
                ("t_last_done", C.c_double), ("t_prep", C.c_double), ("t_batch", C.c_double),
                ("t_wait", C.c_double), ("t_deliver", C.c_double)]


RESULT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(gf_result), C.POINTER(gf_table), C.c_uint64, C.c_uint64)


def _declare():
    L = lib()
    if getattr(L, "_gf_server_declared", False):
        return L
    P = C.POINTER
    L.gf_last_error.restype = C.c_char_p
    L.gf_server_create.argtypes = [C.c_char_p, C.c_char_p, P(gf_server_config), RESULT_FN, C.c_void_p,
                                   P(C.c_void_p)]
    L.gf_server_submit.argtypes = [C.c_void_p, P(gf_request)]
    L.gf_server_wait_idle.argtypes = [C.c_void_p]
    L.gf_server_stats_get.argtypes = [C.c_void_p, P(gf_server_stats)]
    L.gf_server_clock.restype = C.c_double
    L.gf_server_destroy.argtypes = [C.c_void_p]
    L.gf_server_destroy.restype = None
    L._gf_server_declared = True
    return L


class Server:
    """Continuous-batching request server.

    on_result(result: gf_result, table: gf_table or None, row0, nrows) runs
    on a server thread; the device columns are valid only during the call.
    """

    def __init__(self, cmd_conf, on_result=None, hbm_budget=64 << 30, max_batch=64, workers=2, plain=False,
                 linger_us=0, eager_chain=False, datadir=DATA_DIR):
        L = _declare()
        cfg = gf_server_config(hbm_budget, max_batch, workers, 1 if plain else 0, linger_us,
                               1 if eager_chain else 0, 0)
        self.results = []
        self._user_cb = on_result

        def _cb(user, res, table, row0, nrows):
            r = gf_result.from_buffer_copy(res.contents)
            try:
                if self._user_cb is not None:
                    self._user_cb(r, table.contents if table else None, row0, nrows)
            except Exception as e:  # keep the server alive; surface later
                self._error = e
            self.results.append(r)
            return 0

        self._error = None
        self._cb = RESULT_FN(_cb)
        self._h = C.c_void_p()
        rc = L.gf_server_create(cmd_conf.encode(), datadir.encode(), C.byref(cfg), self._cb, None, C.byref(self._h))
        if rc != 0:
            msg = L.gf_last_error().decode() if rc == -4 else L.gf_strerror(rc).decode()
            raise GalfastError("gf_server_create: " + msg)

    @staticmethod
    def clock():
        return _declare().gf_server_clock()

    def submit(self, rid, l, b, radius, m1, seed, l1=0.0, b1=0.0):
        """A beam (radius > 0) or, with radius 0, the Galactic rect
        [l, l1] x [b, b1] (degrees)."""
        r = gf_request(rid, l, b, radius, m1, seed, l1, b1)
        check(lib().gf_server_submit(self._h, C.byref(r)), "gf_server_submit")

    def wait(self):
        check(lib().gf_server_wait_idle(self._h), "gf_server_wait_idle")
        if self._error is not None:
            raise self._error

    def stats(self):
        s = gf_server_stats()
        check(lib().gf_server_stats_get(self._h, C.byref(s)), "gf_server_stats_get")
        return s

    def close(self):
        if self._h:
            lib().gf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
