"""The server's per-request pixelization of beams and Galactic rects (host
code of the library, no GPU) against the oracle's grid coverage of the same
region: every covered cell of both pole grids, with its sample count, in the
catalog's pixel order.  Regions that cross, touch or barely dip below the
equator are the cases where the two hemispheres' grids both hold cells of the
request."""
import ctypes as C
import math

import numpy as np
import pytest

import galfast_amd as GA
from galfast_amd import _abi as A
import oracle as O

D2R = math.pi / 180


def lib_pixels(l, b, r, dx):
    L = GA.lib()
    L.gf_beam_pixels_host.argtypes = [C.c_double] * 4 + [C.c_void_p, C.c_int]
    n = L.gf_beam_pixels_host(l, b, r, dx, None, 0)
    assert n >= 0
    out = (A.gf_pencil_beam * max(n, 1))()
    assert L.gf_beam_pixels_host(l, b, r, dx, out, n) == n
    return [(p.projIdx, p.X, p.Y, round(p.coveredFraction * 64)) for p in out[:n]]


def oracle_pixels(l, b, r, dx, rect=None):
    f = A.gf_footprint()
    f.nregion = 1
    g = f.r[0]
    if rect is None:
        g.kind = 2
        g.cl, g.sl = math.cos(l * D2R), math.sin(l * D2R)
        g.cb, g.sb = math.cos(b * D2R), math.sin(b * D2R)
        g.cosr = math.cos(r * D2R)
    else:
        g.kind = 1
        g.l0, g.l1, g.b0, g.b1 = rect
    d = dx * D2R
    n = int(math.ceil(2 * math.sqrt(2.0) / d))
    cov = np.zeros(2 * n * n, np.uint8)
    O.lib().gfo_footprint_cover(C.byref(f), d, n, cov.ctypes.data)
    out = []
    x0 = -n * d / 2
    for idx in np.flatnonzero(cov):
        h, rem = divmod(int(idx), n * n)
        i, j = divmod(rem, n)
        # cell (i, j) spans [x0 + i dx, + dx) x [x0 + j dx, + dx); its pixel
        # carries the centre as a float
        out.append((h, float(np.float32(x0 + i * d + 0.5 * d)), float(np.float32(x0 + j * d + 0.5 * d)),
                    int(cov[idx])))
    return out, n


@pytest.mark.parametrize("l,b,r", [
    (30.0, 0.0, 2.0),        # centred on the equator
    (120.0, 1.0, 1.02),      # dips 0.02 deg below the equator
    (200.0, -0.5, 0.52),     # reaches 0.02 deg above it
    (300.0, 3.0, 3.0),       # touches it
    (75.0, -5.0, 5.6),       # crosses it near the edge of the cap
    (10.0, 45.0, 3.0),       # one hemisphere only
    (100.0, 25.0, 30.0),     # a wide cap: the equator crossings lie between boundary samples
    (250.0, -28.0, 30.5),    # and from the south
    (140.0, 62.0, 5.642),    # 100 deg^2: most cells wholly inside, counted without samples
    (33.0, -71.0, 0.0564),   # 0.01 deg^2, smaller than a cell
    (5.0, 87.0, 6.0),        # over the north pole
    (181.0, -86.5, 9.0),     # over the south pole
])
def test_beam_pixels_equal_grid_coverage(l, b, r):
    dx = 0.25
    got = lib_pixels(l, b, r, dx)
    want, n = oracle_pixels(l, b, r, dx)
    # both lists are in (hemisphere, i, j) order
    assert len(got) == len(want)
    assert got == want
    if abs(b) < r - 0.01:
        assert {h for h, _, _, _ in want} == {0, 1}


def lib_rect_pixels(l0, l1, b0, b1, dx):
    L = GA.lib()
    L.gf_rect_pixels_host.argtypes = [C.c_double] * 5 + [C.c_void_p, C.c_int]
    n = L.gf_rect_pixels_host(l0, l1, b0, b1, dx, None, 0)
    assert n >= 0
    out = (A.gf_pencil_beam * max(n, 1))()
    assert L.gf_rect_pixels_host(l0, l1, b0, b1, dx, out, n) == n
    return [(p.projIdx, p.X, p.Y, round(p.coveredFraction * 64)) for p in out[:n]]


@pytest.mark.parametrize("rect,oracle_rect", [
    ((86.0, 89.0, 11.0, 14.0), (86.0, 89.0, 11.0, 14.0)),        # a config-4 tile
    ((358.0, 362.0, -30.0, -27.0), (-2.0, 2.0, -30.0, -27.0)),   # across l = 0, as footprint files reduce it
    ((10.0, 25.0, -3.0, 4.0), (10.0, 25.0, -3.0, 4.0)),          # across the equator
    ((0.0, 360.0, 80.0, 90.0), (0.0, 360.0, 80.0, 90.0)),        # the polar cap
    ((200.0, 230.0, -90.0, -60.0), (200.0, 230.0, -90.0, -60.0)),  # a wedge at the south pole
])
def test_rect_pixels_equal_grid_coverage(rect, oracle_rect):
    dx = 0.5
    got = lib_rect_pixels(*rect, dx)
    want, n = oracle_pixels(0, 0, 0, dx, rect=oracle_rect)
    assert len(got) == len(want) and got == want
    if rect[2] < 0 < rect[3]:
        assert {h for h, _, _, _ in want} == {0, 1}
