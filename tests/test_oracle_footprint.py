"""Footprint regions and sky coverage of the oracle, pinned by what the
geometry fixes (no GPU).

The reference builds footprints as gpc polygons from one configuration file
each (footprint.cpp:486-759) and pixelizes them on pole-centred Lambert grids
(skygen/os_skygen.cpp:235-283, 940-995).  The oracle tests the region
boundaries themselves and counts 8 x 8 sample points per grid cell; these
tests pin that restatement with
  * the J2000 Galactic pole and centre in equatorial coordinates (textbook
    values, IAU 1958 system as precessed to J2000),
  * closed-form areas of caps, rings, longitude/latitude rectangles and
    equatorial rectangles (the Lambert map is equal-area: a cell of side dx
    holds dx^2 sr),
  * exact sample-count identities (disjoint unions add, holes subtract),
each chosen so that a wrong sign, axis, hole test or zone-of-avoidance test
fails one of them.
"""
import ctypes as C
import math

import numpy as np
import pytest

import gfx as G
import oracle as O

D2R = math.pi / 180


def cover(f, dx_deg=1.0):
    dx = dx_deg * D2R
    n = int(math.ceil(2 * math.sqrt(2.0) / dx))
    out = np.zeros(2 * n * n, np.uint8)
    O.lib().gfo_footprint_cover(C.byref(f), dx, n, out.ctypes.data)
    return out.astype(np.int64), dx


def area(f, dx_deg=1.0):
    c, dx = cover(f, dx_deg)
    return c.sum() / 64.0 * dx * dx


def galequ(l, b):
    ra, dec = C.c_double(), C.c_double()
    O.lib().gfo_galequ(l * D2R, b * D2R, C.byref(ra), C.byref(dec))
    return ra.value / D2R, dec.value / D2R


def outside(f, l, b):
    return O.lib().gfo_outside_footprint(C.byref(f), l * D2R, b * D2R)


def test_galequ_known_directions():
    # North Galactic pole: ra 12h51m26.28s, dec +27d07m42.0s (J2000)
    ra, dec = galequ(123.0, 90.0)
    assert abs(ra - 192.85950) < 2e-4 and abs(dec - 27.12833) < 2e-4
    # Galactic centre: ra 17h45m37.2s, dec -28d56m10s (J2000)
    ra, dec = galequ(0.0, 0.0)
    assert abs(ra - 266.40500) < 2e-3 and abs(dec - -28.93617) < 2e-3
    # ascending node of the Galactic plane on the equator: l = 32.93 deg, ra = 282.86 deg
    ra, dec = galequ(32.932, 0.0)
    assert abs(dec) < 1e-9 and abs(ra - (192.859508333 + 90.0)) < 1e-9


def test_whole_sky_covers_4pi():
    f = G.footprint()
    # sampling of the grid edge at r^2 = 2 is the only error
    assert abs(area(f) / (4 * math.pi) - 1) < 2e-4
    assert outside(f, 10.0, -80.0) == 0


@pytest.mark.parametrize("r", [5.0, 30.0, 85.0])
def test_beam_cap_area(r):
    f = G.footprint(G.beam_region(40.0, 35.0, r))
    exact = 2 * math.pi * (1 - math.cos(r * D2R))
    assert abs(area(f) / exact - 1) < 2e-3 * (30.0 / r)
    assert outside(f, 40.0, 35.0) == 0 and outside(f, 40.0 + 180, -35.0) == 1


def test_beam_hole_is_the_ring():
    full = G.footprint(G.beam_region(200.0, -50.0, 20.0))
    ring = G.footprint(G.beam_region(200.0, -50.0, 20.0, hole=8.0))
    hole = G.footprint(G.beam_region(200.0, -50.0, 8.0))
    cf, cr, ch = cover(full)[0], cover(ring)[0], cover(hole)[0]
    # every sample of the cap is in the ring or in the hole, never both
    assert abs(cf.sum() - cr.sum() - ch.sum()) <= 2
    exact = 2 * math.pi * (math.cos(8 * D2R) - math.cos(20 * D2R))
    assert abs(area(ring) / exact - 1) < 5e-3
    assert outside(ring, 200.0, -50.0) == 1 and outside(ring, 200.0, -50.0 + 14.0) == 0


def test_bmin_cuts_the_plane_for_every_region_kind():
    # a cap of radius 85 deg around the north Galactic pole with |b| >= 10:
    # exactly the band 10 <= b <= 90, also as a rect
    cap = G.footprint(G.beam_region(0.0, 90.0, 85.0, bmin=10.0))
    band = G.rect_footprint(0, 360, 10, 90)
    exact = 2 * math.pi * (1 - math.sin(10 * D2R))
    assert abs(area(cap) / exact - 1) < 2e-3
    assert abs(area(band) / exact - 1) < 2e-3
    assert outside(cap, 77.0, 7.0) == 1 and outside(cap, 77.0, 12.0) == 0
    # the rect's own bmin
    r = G.rect_footprint(0, 360, -90, 90, bmin=30.0)
    assert abs(area(r) / (4 * math.pi * (1 - math.sin(30 * D2R))) - 1) < 2e-3


def test_rect_wrapping_longitudes():
    # l0 > l1 wraps through l = 0 (footprint.cpp:510-513): 350..10 is 20 deg wide
    f = G.rect_footprint(-10.0, 10.0, 20.0, 40.0)
    exact = 20 * D2R * (math.sin(40 * D2R) - math.sin(20 * D2R))
    assert abs(area(f, 0.5) / exact - 1) < 5e-3
    assert outside(f, 355.0, 30.0) == 0 and outside(f, 15.0, 30.0) == 1


def test_equatorial_rect():
    # ra 150..210, dec -10..+20 (J2000): area (a1 - a0)(sin d1 - sin d0)
    f = G.rect_footprint(150.0, 210.0, -10.0, 20.0, equ=1)
    exact = 60 * D2R * (math.sin(20 * D2R) - math.sin(-10 * D2R))
    assert abs(area(f, 0.5) / exact - 1) < 5e-3
    # the north Galactic pole is at ra 192.86, dec 27.13: outside; the
    # direction (l, b) = (300, 60) lies at ra ~ 193, dec ~ -3: inside
    assert outside(f, 0.0, 90.0) == 1
    ra, dec = galequ(300.0, 60.0)
    assert 150 < ra < 210 and -10 < dec < 20
    assert outside(f, 300.0, 60.0) == 0
    # the same rect in Galactic coordinates is elsewhere
    g = G.rect_footprint(150.0, 210.0, -10.0, 20.0)
    assert outside(g, 300.0, 60.0) == 1


def test_disjoint_union_adds_and_overlap_counts_once():
    a = G.beam_region(60.0, 40.0, 10.0)
    b = G.beam_region(240.0, -40.0, 12.0)
    ca, cb, cab = cover(G.footprint(a))[0], cover(G.footprint(b))[0], cover(G.footprint(a, b))[0]
    assert cab.sum() == ca.sum() + cb.sum()
    assert np.array_equal(cab, np.maximum(ca, cb))
    # a region inside another adds nothing
    inner = G.beam_region(60.0, 40.0, 4.0)
    assert np.array_equal(cover(G.footprint(a, inner))[0], ca)
    # overlapping regions: samples are counted once
    c = G.beam_region(70.0, 40.0, 10.0)
    cac = cover(G.footprint(a, c))[0]
    cc = cover(G.footprint(c))[0]
    assert max(ca.sum(), cc.sum()) < cac.sum() < ca.sum() + cc.sum()
