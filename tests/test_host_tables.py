"""Host-side table construction of the library against the oracle, and the
oracle's table builders against closed forms (CPU only)."""
import ctypes as C

import numpy as np

import galfast_amd as G
from galfast_amd import configs
import oracle as O


def _lib():
    L = G.lib()
    L.gf_lf_texture_host.argtypes = [C.c_void_p] * 2 + [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    L.gf_photometry_tables_host.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p,
                                            C.c_void_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.c_void_p]
    return L


def test_lf_texture_matches_oracle_and_closed_form():
    M, phi = configs.seeded_lf()
    tex = np.zeros(1024, dtype=np.float32)
    tc = np.zeros(2, dtype=np.float32)
    assert _lib().gf_lf_texture_host(O.ptr(M), O.ptr(phi), len(M), 1024, O.ptr(tex), O.ptr(tc)) == 0
    rtex = np.zeros(1024, dtype=np.float32)
    rtc = np.zeros(2, dtype=np.float32)
    O.lib().gfo_resample_clip(O.ptr(M), O.ptr(phi), len(M), C.c_float(0), 1024, O.ptr(rtex), O.ptr(rtc))
    assert np.array_equal(tex, rtex) and np.array_equal(tc, rtc)
    # closed form: texel i sits at X0 + i*dX on the piecewise-linear LF,
    # with zero-valued knots 1e-5 of the range outside both ends
    rng = M[-1] - M[0]
    X = np.r_[M[0] - 1e-5 * rng, M, M[-1] + 1e-5 * rng]
    Y = np.r_[0.0, phi, 0.0]
    dX = (X[-1] - X[0]) / 1023
    xs = X[0] + np.arange(1024) * dX
    np.testing.assert_allclose(rtex, np.interp(xs, X, Y).astype(np.float32), rtol=2e-7, atol=1e-12)
    assert rtc[0] == np.float32(X[0]) and rtc[1] == np.float32(1.0 / dX)


def test_photometry_tables_match_oracle_and_closed_form():
    rows = configs.seeded_photosys()
    nc = 4
    cap = 4 * 1000 * 2000 * 4
    out = [np.zeros(cap, dtype=np.float32) for _ in range(4)]
    nF, nM = C.c_int(), C.c_int()
    rnF, rnM = C.c_int(), C.c_int()
    tc, rtc = np.zeros(4, dtype=np.float32), np.zeros(4, dtype=np.float32)
    rows = np.ascontiguousarray(rows)
    nt = _lib().gf_photometry_tables_host(O.ptr(rows), rows.shape[0], nc, 0.01, 0.01, O.ptr(out[0]), O.ptr(out[1]),
                                         cap, C.byref(nF), C.byref(nM), O.ptr(tc))
    rt = O.lib().gfo_photometry_tables(O.ptr(rows), rows.shape[0], nc, 0.01, 0.01, O.ptr(out[2]), O.ptr(out[3]),
                                       cap, C.byref(rnF), C.byref(rnM), O.ptr(rtc))
    assert nt == rt == 1 and nF.value == rnF.value and nM.value == rnM.value
    n = nF.value * nM.value * 4
    assert np.array_equal(out[0][:n], out[2][:n]) and np.array_equal(out[1][:n], out[3][:n])
    assert np.array_equal(tc, rtc)
    # closed form: on the knot grid the table reproduces the input colours,
    # and nothing inside the knots' hull is flagged as extrapolated
    col = out[2][:n].reshape(nM.value, nF.value, 4)
    flg = out[3][:n].reshape(nM.value, nF.value, 4)
    FeH0, Mr0 = rows[:, 1].min(), rows[:, 0].min()
    for r in rows[::97]:
        f = int(round((r[1] - FeH0) / 0.01))
        m = int(round((r[0] - Mr0) / 0.01))
        if f < nF.value and m < nM.value:
            np.testing.assert_allclose(col[m, f], r[2:6].astype(np.float32), rtol=0, atol=2e-6)
            assert not flg[m, f].any()
