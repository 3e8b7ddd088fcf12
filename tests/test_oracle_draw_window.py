"""The oracle's capacity-limited reference draw: rows [r0, r1) of the crawl
drawn from the streams' starting states (gfo_draw_window), batch after
batch, concatenate to the unlimited draw (gfo_draw_compat) column for
column, and the count pass leaves the streams where the draw does
(skygen/skyconfig_impl.h:434-534, skygen/skygen.h:184-257)."""
import ctypes as C

import numpy as np

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O
import gfx


def setup():
    lf_M, lf_phi = gfx.seeded_lf()
    tex, tc = gfx.lf_texture(lf_M, lf_phi)
    pix = gfx.pixelize_rect(85.0, 95.0, 55.0, 65.0, 1.0)
    sp = gfx.skygen_params(15.0, 21.0, 0.05, 3.0, 15.0, len(pix), nthreads=1024)
    m = gfx.make_model(A.GF_MODEL_EXPDISK, 1, 0.12, tex, tc, l=3600.0, h=900.0)
    return sp, pix, m, (tex, tc)


def test_windows_concatenate_to_the_unlimited_draw():
    sp, pix, m, keep = setup()
    lib = O.lib()
    st0 = gfx.mwc_state(9)
    full_state = st0.copy()
    n = lib.gfo_draw_compat(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(full_state.copy()),
                            65536, None, 0)
    assert n > 20_000
    full = HostTable(n, catalog_spec())
    assert lib.gfo_draw_compat(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(full_state),
                               65536, C.byref(full.skygen_cols()), n) == n
    count_state = st0.copy()
    assert lib.gfo_draw_compat(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(count_state),
                               65536, None, 0) == n
    assert np.array_equal(count_state, full_state)      # the count pass moves the streams like the draw
    cap = 4999
    got = HostTable(n, catalog_spec())
    for r0 in range(0, n, cap):
        r1 = min(n, r0 + cap)
        part = HostTable(r1 - r0, catalog_spec())
        assert lib.gfo_draw_window(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(st0), 65536,
                                   r0, r1, C.byref(part.skygen_cols())) == r1 - r0
        for c in A.SKYGEN_COLUMNS:
            name = "absmag" if c == "M" else c
            got.arr(name)[:, r0:r1] = part.arr(name)[:, :r1 - r0]
    for c in A.SKYGEN_COLUMNS:
        name = "absmag" if c == "M" else c
        assert np.array_equal(got.arr(name).view(np.uint8), full.arr(name).view(np.uint8)), name
    # a window past the end writes nothing
    assert lib.gfo_draw_window(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(st0), 65536,
                               n, n + 10, C.byref(got.skygen_cols())) == 0
