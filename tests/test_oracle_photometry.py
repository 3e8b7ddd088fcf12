"""Pins of the oracle's photometry (modules/photometry_gpu.cu.h:87-174) on
hand-built colour tables whose values the bilinear sampling reproduces
exactly: colours linear in (FeH, Mr).

  lone star    mag_b = DM + M_b + Am r_b, M_b = Mr + sum of the colours
               c_b .. c_(bidx-1) above the bootstrap band bidx,
               Mr - (c_bidx + .. + c_(b-1)) below it (:124-129, 159-172)
  system       mag_b = DM - 2.5 log10(sum of the components' 10^(-0.4 M_b))
               + Am r_b (:114-134, 166)
  flags        output bit (ncolors - b) set when band b was extrapolated:
               some colour between b and the bootstrap band has a nonzero
               flag-table sample, counting colour b itself on the faint
               side; never for the bootstrap band (:136-156)
  selection    hidden rows and components outside the module's set are
               left alone (:99-103)
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O
import gfx

F0, DF, NF = -3.0, 0.1, 36          # FeH grid
M0, DMR, NM = 1.0, 0.1, 171         # Mr grid


def colour_coeffs(nc, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.5, 0.5, nc), rng.uniform(-0.2, 0.2, nc), rng.uniform(0.05, 0.4, nc)


def tables(nc, coeffs, flag_fn=None):
    """float4 colour and flag tables [Mr][FeH][4] (FeH fastest), as many as
    nc colours need; flag_fn(i, F, M) gives colour i's flag texels."""
    a, bF, bM = coeffs
    F = F0 + DF * np.arange(NF)
    M = M0 + DMR * np.arange(NM)
    MM, FF = np.meshgrid(M, F, indexing="ij")
    ntex = (nc + 3) // 4
    col = np.zeros((ntex, NM, NF, 4), dtype=np.float32)
    flg = np.zeros((ntex, NM, NF, 4), dtype=np.float32)
    for i in range(nc):
        col[i // 4, :, :, i % 4] = a[i] + bF[i] * FF + bM[i] * MM
        if flag_fn is not None:
            flg[i // 4, :, :, i % 4] = flag_fn(i, FF, MM)
    return ntex, col, flg


def params(nc, bidx, red, ntex, col, flg):
    p = A.gf_photometry_data()
    p.ncolors, p.bidx = nc, bidx
    for i in range(17):
        p.reddening[i] = red[i] if i < len(red) else 1.0
    p.nFeH, p.nMr = NF, NM
    p.tcFeH[0], p.tcFeH[1] = F0, 1.0 / DF
    p.tcMr[0], p.tcMr[1] = M0, 1.0 / DMR
    for i in range(4):
        p.color[i] = col[i].ctypes.data if i < ntex else None
        p.cflags[i] = flg[i].ctypes.data if i < ntex else None
    return p


def colours(coeffs, FeH, Mr):
    a, bF, bM = coeffs
    return a[:, None] + bF[:, None] * FeH[None, :] + bM[:, None] * Mr[None, :]


def band_absmag(c, Mr, bidx, sign=1.0):
    """M_b of every band b = 0 .. nc from the colours c[i] = M_i - M_(i+1)."""
    nc = c.shape[0]
    out = np.zeros((nc + 1, Mr.size))
    for b in range(nc + 1):
        if b < bidx:
            out[b] = Mr + sign * c[b:bidx].sum(axis=0)
        elif b > bidx:
            out[b] = Mr - sign * c[bidx:b].sum(axis=0)
        else:
            out[b] = Mr
    return out


def star_table(n, nc, seed, Mr=(2.0, 16.0)):
    rng = np.random.default_rng(seed)
    t = HostTable(n, catalog_spec(nbands=nc + 1, nabsmag=2))
    t.arr("FeH")[0] = rng.uniform(F0 + 0.05, F0 + DF * (NF - 1) - 0.05, n)
    t.arr("absmag")[0] = rng.uniform(*Mr, n)
    t.arr("DM")[0] = rng.uniform(5.0, 15.0, n)
    t.arr("Am")[0] = rng.uniform(0.0, 0.5, n)
    t.arr("comp")[0] = rng.integers(0, 3, n)
    t.arr("hidden")[0] = (rng.uniform(size=n) < 0.1).astype(np.int32)
    t.arr("mags")[:] = 77.0
    return t


def run_chain(t, p, comps, mult=None, seed=5):
    ch = A.gf_chain()
    ch.phot = C.pointer(p)
    ch.phot_comps = comps
    if mult is not None:
        ch.mult = C.pointer(mult)
        ch.mult_comps = A.bitmap(0, 1, 2)
    st = gfx.mwc_state(seed)
    rows = A.gf_rows(0, 256, t.n, 0)
    assert O.lib().gfo_module_chain(C.byref(rows), C.byref(ch), O.ptr(st), 65536, C.byref(t.table())) == 0


@pytest.mark.parametrize("nc,bidx", [(4, 2), (4, 0), (4, 4), (6, 3), (5, 1)])
def test_lone_star_magnitudes(nc, bidx):
    n = 20_000
    coeffs = colour_coeffs(nc)
    ntex, col, flg = tables(nc, coeffs)
    red = np.random.default_rng(7).uniform(0.5, 3.0, nc + 1)
    p = params(nc, bidx, red, ntex, col, flg)
    t = star_table(n, nc, seed=nc * 10 + bidx)
    run_chain(t, p, A.bitmap(0, 2))
    FeH = t.arr("FeH")[0].astype(np.float64)
    Mr = t.arr("absmag")[0].astype(np.float64)
    DM = t.arr("DM")[0].astype(np.float64)
    Am = t.arr("Am")[0].astype(np.float64)
    c = colours(coeffs, FeH, Mr)
    want = DM + band_absmag(c, Mr, bidx) + Am * red[:, None]
    mags = t.arr("mags").astype(np.float64)
    sel = (t.arr("hidden")[0] == 0) & (t.arr("comp")[0] != 1)
    assert sel.sum() > n // 2 and (~sel).sum() > n // 10
    assert np.abs(mags[:, sel] - want[:, sel]).max() < 3e-5
    assert np.all(mags[:, ~sel] == 77.0)
    assert np.all(t.arr("photoFlags")[0][sel] == 0)
    # planted mistakes: the colour sums with the wrong sign, reddening of the neighbouring band
    if bidx > 0 or bidx < nc:
        bad = DM + band_absmag(c, Mr, bidx, sign=-1.0) + Am * red[:, None]
        assert np.abs(mags[:, sel] - bad[:, sel]).max() > 1e-2
    bad = DM + band_absmag(c, Mr, bidx) + Am * np.roll(red, 1)[:, None]
    assert np.abs(mags[:, sel] - bad[:, sel]).max() > 1e-2


@pytest.mark.parametrize("algo", [2, 3])
def test_system_is_the_luminosity_sum(algo):
    nc, bidx, n = 4, 2, 30_000
    coeffs = colour_coeffs(nc, seed=11)
    ntex, col, flg = tables(nc, coeffs)
    red = np.array([1.3, 1.0, 0.8, 0.6, 0.45])
    p = params(nc, bidx, red, ntex, col, flg)
    lf_M, lf_phi = gfx.seeded_lf()
    (sec, sec_tc), (cum, cum_tc), (inv, inv_tc) = gfx.multiples_tables(lf_M, lf_phi, fraction=0.5)
    mp = A.gf_multiples_data()
    mp.secProb = gfx.tex1d(sec.ctypes.data, len(sec), sec_tc)
    mp.cumLF = gfx.tex1d(cum.ctypes.data, len(cum), cum_tc)
    mp.invCumLF = gfx.tex1d(inv.ctypes.data, len(inv), inv_tc)
    mp.algo = algo
    t = star_table(n, nc, seed=21, Mr=(4.0, 12.0))
    run_chain(t, p, A.bitmap(0, 1, 2), mult=mp)
    sys = t.arr("absmagSys").astype(np.float64)
    ncomp = t.arr("absmagSysNcomp")[0]
    sel = t.arr("hidden")[0] == 0
    pair = sel & (ncomp == 2)
    lone = sel & (ncomp == 1)
    assert pair.sum() > n // 5 and lone.sum() > n // 5
    FeH = t.arr("FeH")[0].astype(np.float64)
    DM = t.arr("DM")[0].astype(np.float64)
    Am = t.arr("Am")[0].astype(np.float64)
    M1 = band_absmag(colours(coeffs, FeH, sys[0]), sys[0], bidx)
    M2 = band_absmag(colours(coeffs, FeH, sys[1]), sys[1], bidx)
    mags = t.arr("mags").astype(np.float64)
    want_pair = DM - 2.5 * np.log10(10 ** (-0.4 * M1) + 10 ** (-0.4 * M2)) + Am * red[:, None]
    want_lone = DM + M1 + Am * red[:, None]
    assert np.abs(mags[:, pair] - want_pair[:, pair]).max() < 5e-5
    assert np.abs(mags[:, lone] - want_lone[:, lone]).max() < 3e-5
    if algo == 3:
        # equal components: the system is 2.5 log10 2 brighter than either
        assert np.abs(mags[:, pair] - (want_lone[:, pair] - 2.5 * math.log10(2.0))).max() < 5e-5
    # planted mistake: magnitudes averaged instead of luminosities summed
    bad = DM + 0.5 * (M1 + M2) + Am * red[:, None]
    assert np.abs(mags[:, pair] - bad[:, pair]).max() > 0.1


def expected_flags(cf, nc, bidx):
    """The output flag word from per-colour flags cf[i] (bool)."""
    out = 0
    for b in range(nc + 1):
        if b == bidx:
            continue
        idx = range(b, bidx) if b < bidx else range(bidx, min(b, nc - 1) + 1)
        if any(cf[i] for i in idx):
            out |= 1 << (nc - b)
    return out


@pytest.mark.parametrize("nc", [4, 6])
def test_flag_propagation_rule(nc):
    coeffs = colour_coeffs(nc)
    n = 64
    red = np.ones(nc + 1)
    for bidx in range(nc + 1):
        for cf in itertools.product([0, 1], repeat=nc):
            ntex, col, flg = tables(nc, coeffs, flag_fn=lambda i, F, M: np.full_like(F, cf[i]))
            p = params(nc, bidx, red, ntex, col, flg)
            t = star_table(n, nc, seed=1)
            t.arr("hidden")[0] = 0
            run_chain(t, p, A.bitmap(0, 1, 2))
            got = set(t.arr("photoFlags")[0].tolist())
            assert got == {expected_flags(cf, nc, bidx)}, (bidx, cf, got)
    # planted mistake: without colour b itself on the faint side the rule
    # differs from the oracle for some pattern
    cf = [0] * nc
    cf[nc - 1] = 1
    bidx = 0
    alt = sum(1 << (nc - b) for b in range(1, nc + 1) if any(cf[i] for i in range(0, b)))
    assert alt != expected_flags(cf, nc, bidx)


def test_flag_follows_the_bilinear_footprint_of_flagged_texels():
    """A flag texel marks every sample whose bilinear stencil touches it
    with nonzero weight: Mr above the last unflagged row."""
    nc, bidx = 4, 1
    coeffs = colour_coeffs(nc)
    j0 = 80                                   # rows j >= j0 flagged for colour 3
    Mthr = M0 + DMR * j0
    ntex, col, flg = tables(nc, coeffs, flag_fn=lambda i, F, M: ((M >= Mthr - 1e-9) & (i == 3)).astype(np.float32))
    p = params(nc, bidx, np.ones(nc + 1), ntex, col, flg)
    t = star_table(20_000, nc, seed=2)
    t.arr("hidden")[0] = 0
    run_chain(t, p, A.bitmap(0, 1, 2))
    Mr = t.arr("absmag")[0].astype(np.float64)
    flagged = t.arr("photoFlags")[0] != 0
    below = Mthr - DMR                        # the last unflagged texel row
    clear = np.abs(Mr - below) > 1e-5
    assert np.array_equal(flagged[clear], (Mr > below)[clear])
    assert np.all(t.arr("photoFlags")[0][flagged] == expected_flags([0, 0, 0, 1], nc, bidx))
