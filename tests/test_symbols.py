"""The C-ABI library loads and exports every symbol the header declares."""
import ctypes as C

import galfast_amd as G


def test_library_exports_every_declared_symbol():
    L = G.lib()
    names = G.exported_symbols()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_seed_matches_oracle():
    import numpy as np
    import oracle as O
    from gfx import PRIMES
    L = G.lib()
    for s in (1, 42, 22, 0):
        assert L.gf_seed_from_pipeline_seed(s) == O.lib().gfo_seed_from_pipeline_seed(s)
    st = np.zeros(3 * 65536, dtype=np.uint32)
    n = L.gf_mwc_seed_host(1234, PRIMES.ctypes.data_as(C.POINTER(C.c_uint32)), len(PRIMES), 65536,
                           st.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert n == 65536
    ref = np.zeros_like(st)
    O.lib().gfo_mwc_seed(1234, O.ptr(PRIMES), len(PRIMES), 65536, O.ptr(ref))
    assert np.array_equal(st, ref)


def test_error_strings():
    L = G.lib()
    assert L.gf_strerror(0) == b"success"
    assert b"stream" in L.gf_strerror(-3)
