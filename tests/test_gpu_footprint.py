"""Footprints on the device: the pixel-grid coverage kernel
(gf_footprint_cover) against the oracle's, bit-exact, for every region kind
(rect in Galactic and equatorial coordinates, wrapping rects, beams with
holes, zones of avoidance, unions); and catalogs whose footprint is the
union of several footprint files (load_footprints, footprint.cpp:746-759)
against the oracle's replay of the same job, row for row, in both draw
modes, with the expectation of a union of disjoint beams equal to the sum
of the single-beam jobs'."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import galfast_amd as G
import gfx as X
import oracle as O
import replay as R
from galfast_amd import configs
from galfast_amd.catalog import Catalog
from galfast_amd.table import catalog_spec
from devutil import stream_ptr
from test_gpu_catalog import collect, check_reference_draw

pytestmark = pytest.mark.gpu

D2R = math.pi / 180

CASES = {
    "whole_sky": lambda: X.footprint(),
    "gal_rect_bmin": lambda: X.rect_footprint(0, 360, -90, 90, bmin=10.0),
    "wrapping_rect": lambda: X.rect_footprint(-10.0, 10.0, 20.0, 40.0),
    "equ_rect": lambda: X.rect_footprint(150.0, 210.0, -10.0, 20.0, equ=1),
    "beam_hole": lambda: X.footprint(X.beam_region(200.0, -50.0, 20.0, hole=8.0)),
    "beam_bmin": lambda: X.footprint(X.beam_region(30.0, 5.0, 25.0, bmin=10.0)),
    "union": lambda: X.footprint(X.beam_region(60.0, 40.0, 10.0), X.rect_footprint(150.0, 210.0, -10.0, 20.0,
                                                                                 equ=1).r[0],
                                 X.beam_region(240.0, -40.0, 12.0, hole=3.0, bmin=45.0)),
}


@pytest.mark.parametrize("dx_deg", [1.0, 0.5])
@pytest.mark.parametrize("case", sorted(CASES))
def test_cover_matches_oracle(case, dx_deg):
    import torch
    f = CASES[case]()
    dx = dx_deg * D2R
    n = int(math.ceil(2 * math.sqrt(2.0) / dx))
    print("cover", case, dx_deg, n, flush=True)
    d = torch.zeros(2 * n * n, dtype=torch.uint8, device="cuda")
    G.check(G.lib().gf_footprint_cover(C.byref(f), dx, n, d.data_ptr(), stream_ptr()), "gf_footprint_cover")
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    ref = np.zeros(2 * n * n, np.uint8)
    O.lib().gfo_footprint_cover(C.byref(f), dx, n, ref.ctypes.data)
    assert got.sum() > 0
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, "%d cells differ, first %s" % (bad.size, bad[:5])


def test_cover_rejects_bad_arguments():
    f = X.footprint()
    f.nregion = 9
    assert G.lib().gf_footprint_cover(C.byref(f), 0.01, 10, None, None) != 0


UNION = ["type = rect\ncoordsys = equ\nrect = 200 215 30 40",
         "type = beam\ncoordsys = gal\nfootprint_beam = 90 60 4 1.5",
         "type = rect\ncoordsys = gal\nrect = 80 100 50 54\nbmin = 52"]


@pytest.fixture(scope="module")
def c0_union(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c0u"))
    return configs.set_footprints(configs.write_c0(d), UNION)


def test_union_catalog_matches_oracle_replay(c0_union):
    cat = Catalog(c0_union, seed=7, chunk_rows=40000)
    expected = cat.count()
    view = R.host_view(G.lib(), cat._h)
    assert view.foot.nregion == 3
    spec = catalog_spec(nbands=view.nbands, nabsmag=view.nabsmag)
    got = collect(cat, spec)
    ref = []
    R.replay(view, chunk_rows=40000, on_chunk=lambda t, n, comp: ref.append((t, n, comp)))
    assert [(n, c) for _, n, c in got] == [(n, c) for _, n, c in ref]
    nvis = 0
    for (a, n, _), (b, _, _) in zip(got, ref):
        assert np.array_equal(a.arr("hidden"), b.arr("hidden"))
        vis = b.arr("hidden")[0] == 0
        nvis += vis.sum()
        for c in ("lb", "projXY", "DM", "absmag", "XYZ", "comp", "FeH", "mags"):
            assert np.array_equal(a.arr(c)[:, vis], b.arr(c)[:, vis]), c
        # every visible star lies in one of the three regions
        lb = a.arr("lb")[:, vis]
        out = [O.lib().gfo_outside_footprint(C.byref(view.foot), l * D2R, bb * D2R) for l, bb in lb.T[::97]]
        assert not any(out)
    assert nvis > 0.5 * expected
    cat.close()


def test_union_reference_draw_matches_oracle(c0_union):
    out, drawn, clipped = check_reference_draw(c0_union, 11)
    assert 0 < out <= drawn and clipped > 0


def test_union_of_disjoint_beams_expectation_adds(tmp_path):
    beams = ["type = beam\nfootprint_beam = 90 60 3", "type = beam\ncoordsys = equ\nfootprint_beam = 10 -30 2.5"]
    exp = []
    for k, feet in enumerate([beams, beams[:1], beams[1:]]):
        cmd = configs.set_footprints(configs.write_c0(str(tmp_path / ("j%d" % k))), feet)
        cat = Catalog(cmd, seed=3)
        exp.append((cat.count(), cat.info().npixels))
        cat.close()
    (eu, nu), (ea, na), (eb, nb) = exp
    assert nu == na + nb and na > 0 and nb > 0
    assert abs(eu - (ea + eb)) <= 1e-9 * eu


@pytest.mark.parametrize("body,msg", [
    ("type = rect\ncoordsys = ecl\nrect = 0 10 0 10", "coordinate system 'ecl'"),
    ("coordsys = gal\nrect = 0 10 0 10", "'type' keyword"),
    ("type = polygon", "footprint type polygon is unknown"),
    ("type = beam\ncoordsys = gal", "footprint_beam must be set"),
])
def test_footprint_configuration_errors(tmp_path, body, msg):
    """The reference's footprint.cpp errors (:493-494, :680-689, :729-743)."""
    cmd = configs.set_footprints(configs.write_c0(str(tmp_path)), [body])
    with pytest.raises(G.GalfastError) as e:
        Catalog(cmd, seed=1)
    assert msg in str(e.value)


def test_equatorial_beam_is_the_galactic_beam(tmp_path):
    """A beam centred on the north Galactic pole given in equatorial
    coordinates (ra 192.859508, dec 27.128336, J2000) covers the pixels of
    the beam at b = 90 given in Galactic ones; the ring tolerance turns a
    rect 1e-6 short of a full turn into the full ring."""
    res = []
    for k, body in enumerate(["type = beam\ncoordsys = equ\nfootprint_beam = 192.859508 27.128336 6",
                              "type = beam\ncoordsys = gal\nfootprint_beam = 0 90 6",
                              "type = rect\ncoordsys = gal\nrect = 0 359.999999 60 80",
                              "type = rect\ncoordsys = gal\nrect = 0 360 60 80"]):
        cmd = configs.set_footprints(configs.write_c0(str(tmp_path / ("j%d" % k))), [body])
        cat = Catalog(cmd, seed=2)
        res.append((cat.count(), cat.info().npixels, cat.info().footprint_covered_sr))
        cat.close()
    (e0, n0, a0), (e1, n1, a1), (e2, n2, a2), (e3, n3, a3) = res
    assert n0 == n1 and abs(a0 - a1) <= 2e-4 * a1 and abs(e0 - e1) <= 2e-4 * e1
    assert abs(a1 / (2 * math.pi * (1 - math.cos(6 * D2R))) - 1) < 5e-3
    assert (e2, n2, a2) == (e3, n3, a3)
