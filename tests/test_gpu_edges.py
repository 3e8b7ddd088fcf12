"""Edge cases of the catalog path on the GPU: row ranges around the draw
pass's 16-row tiles and 256-row chunks (some starting inside the table), an
empty row range, and a footprint that holds no pixel."""
import ctypes as C
import os

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import _abi as A
from galfast_amd import configs
from galfast_amd.catalog import Catalog
import gfx
from devutil import stream_ptr
from test_gpu_modules import Fixture, _gfeh_params

pytestmark = pytest.mark.gpu

ALL = ["FeH", "FeH_comp", "absmag", "absmagSys", "absmagSysNcomp", "vcyl", "mags", "photoFlags", "obsMags", "pm",
       "radec"]


def _run_both(fx, rows, kind):
    """One chain on the device and on the oracle over `rows`: kind "phot"
    (FeH, Bond2010, photometry, photometric errors, vel2pm, gal2other) or
    "mult" (GaussianFeH, unresolvedMultiples, Bond2010, photometric
    errors)."""
    comps = A.bitmap(0, 1, 2)
    ch = A.gf_chain()
    bond = gfx.bond_params()
    pe_d, pe_h = fx.perr(True), fx.perr(False)
    ch.bond, ch.perr = C.pointer(bond), C.pointer(pe_d)
    if kind == "phot":
        feh, v2pm = gfx.feh_params(), gfx.vel2pm_params(1)
        ch.feh, ch.v2pm = C.pointer(feh), C.pointer(v2pm)
        ch.phot, ch.phot_comps = C.pointer(fx.phot_dev), comps
        ch.gal2other_coordsys = 1
        oracle = dict(feh=feh, bond=bond, phot=fx.phot_host, phot_comps=comps, perr=pe_h, v2pm=v2pm,
                      gal2other_coordsys=1)
    else:
        gfeh = _gfeh_params()
        md, mh = fx.mult(True, 1), fx.mult(False, 1)
        ch.gfeh, ch.gfeh_comps = C.pointer(gfeh), comps
        ch.mult, ch.mult_comps = C.pointer(md), comps
        oracle = dict(gfeh=gfeh, gfeh_comps=comps, mult=mh, mult_comps=comps, bond=bond, perr=pe_h)
    G.check(G.lib().gf_module_chain(C.byref(rows), C.byref(ch), C.byref(fx.drng.rng), C.byref(fx.dev.table()),
                                    stream_ptr()), "gf_module_chain")
    fx.rows = lambda: rows
    fx.oracle_chain(**oracle)


@pytest.mark.parametrize("kind", ["phot", "mult"])
@pytest.mark.parametrize("begin,n", [(0, 1), (0, 15), (0, 17), (3, 255), (0, 257), (100, 4099)])
def test_chain_small_and_ragged_row_ranges(kind, begin, n):
    """A single row, fewer rows than a tile, one more than a tile, a chunk
    short by one and longer by one, and a range of 16 chunks plus a ragged
    one, starting at row 0 and inside the table: every column of the chain
    equals the oracle's, the rows outside the range untouched on both
    sides."""
    fx = Fixture(n=begin + n + 5)
    _run_both(fx, A.gf_rows(begin, 256, begin + n, 0), kind)
    fx.compare(ALL)


def test_chain_empty_row_range():
    fx = Fixture(n=1000)
    before = fx.dev.to_host()
    rows = A.gf_rows(500, 256, 500, 0)
    ch = A.gf_chain()
    feh, bond = gfx.feh_params(), gfx.bond_params()
    ch.feh, ch.bond = C.pointer(feh), C.pointer(bond)
    G.check(G.lib().gf_module_chain(C.byref(rows), C.byref(ch), C.byref(fx.drng.rng), C.byref(fx.dev.table()),
                                    stream_ptr()), "gf_module_chain")
    after = fx.dev.to_host()
    for c in ALL:
        assert np.array_equal(before.arr(c), after.arr(c), equal_nan=True), c
    # the streams did not move
    assert np.array_equal(fx.drng.state(), fx.state)


def test_catalog_footprint_without_pixels(tmp_path):
    """A rect inside the zone of avoidance (|b| < bmin everywhere) holds no
    pixel: the job counts and draws nothing, without error."""
    d = str(tmp_path)
    conf = configs.write_c0(d)
    with open(os.path.join(d, "foot.conf"), "w") as f:
        f.write("module = footprint\ntype = rect\ncoordsys = gal\nrect = 85 95 -5 5\nbmin = 10\n")
    cat = Catalog(conf, seed=5)
    try:
        assert cat.info().npixels == 0
        assert cat.count() == 0
        st = cat.run()
        assert st.generated == 0 and st.visible == 0
    finally:
        cat.close()
