"""Pin the oracle's countsMap binning (os_countsMap::threaded_process,
pipeline.cpp:418-494) by what the mathematics fixes: bin centres and edges
land in known bins, the polar Lambert projection matches the textbook
radius 2 sin(theta / 2) and azimuth, coadd counts are the number of rows
whose depth bin is at or below Z, and the whole map equals a numpy
histogram of the rows.  Each check also rejects a planted mistake."""
import numpy as np

from countsmap_util import D2R, host_cols, oracle_bin, params, shape


def run(p, x, y, zs, hidden=None):
    n = len(x)
    xy = np.ascontiguousarray(np.stack([x, y]).astype(np.float64))
    z = np.ascontiguousarray(np.atleast_2d(zs).astype(np.float32))
    h = np.ascontiguousarray(np.zeros(n, np.int32) if hidden is None else hidden.astype(np.int32))
    counts, beams, tot = oracle_bin(p, *host_cols(xy, z, h), n)
    nb, nz, zw = shape(p)
    return counts.reshape(nb, nz, zw), beams, tot


def test_bin_centres_edges_and_clamping():
    # x bins of width 0.5 from 10: centres 10 + 0.5 k, edges half-way
    p = params(nx=7, ny=1, nz=3, x0=10.0, dx=0.5, y0=0.0, dy=1.0, z0=0.0, dz=1.0, zmag_width=1)
    k = np.arange(7)
    x = np.r_[10.0 + 0.5 * k, 10.0 + 0.5 * k + 0.2499, 10.0 + 0.5 * k - 0.2499, 9.0, 20.0, np.nan]
    want_x = np.r_[k, k, k, 0, 6, 0]          # below and above the range clamp; NaN -> bin 0
    z = np.r_[np.zeros(21), 0.0, 0.0, 0.0]
    for i, (xi, wx) in enumerate(zip(x, want_x)):
        c, beams, tot = run(p, [xi], [0.0], [[z[i]]])
        assert tot == 1
        assert np.flatnonzero(beams).tolist() == [wx], (xi, wx)
        assert c[wx, 0, 0] == 1
    # a planted mistake: rounding down instead of to the nearest centre
    wrong = np.floor((x[:21] - 10.0) / 0.5).astype(int)
    assert not np.array_equal(wrong, want_x[:21])


def test_polar_lambert_closed_form():
    # fine 1-D maps: the x (or y) bin gives the projected coordinate to 1e-5
    rng = np.random.default_rng(5)
    n = 1500
    lon = rng.uniform(0, 360, n)
    lat = np.where(rng.random(n) < 0.5, 1, -1) * rng.uniform(0.5, 89.5, n)
    d = 1e-5
    north = lat > 0
    # textbook polar Lambert azimuthal equal-area map about each pole:
    # radius 2 sin(theta / 2), theta the angle from the pole
    theta = np.where(north, 90.0 - lat, 90.0 + lat) * D2R
    rho = 2 * np.sin(theta / 2)
    az = np.where(north, (lon - 90.0) * D2R, (lon + 90.0) * D2R)
    ex = rho * np.sin(az)
    ey = np.where(north, -rho * np.cos(az), rho * np.cos(az))
    for axis, e in ((0, ex), (1, ey)):
        nbin = int(round(4 / d)) + 1
        if axis == 0:
            p = params(nx=nbin, ny=1, nz=1, x0=-2.0, dx=d, y0=0.0, dy=1.0, zmag_width=1, equalarea=True)
        else:
            p = params(nx=1, ny=nbin, nz=1, x0=0.0, dx=1.0, y0=-2.0, dy=d, zmag_width=1, equalarea=True)
        got = np.empty(n, np.int64)
        maps = np.empty(n, np.int64)
        nbm, nz, zw = shape(p)
        counts = np.zeros(nbm * nz * zw, np.uint32)
        beams = np.zeros(nbm, np.uint8)
        z = np.zeros((1, 1), np.float32)
        h = np.zeros(1, np.int32)
        for i in range(n):
            xy = np.array([[lon[i]], [lat[i]]], np.float64)
            oracle_bin(p, *host_cols(xy, z, h), 1, counts, beams)
            b = int(np.flatnonzero(beams)[0])
            beams[b] = 0
            got[i] = b // 2
            maps[i] = b % 2
        assert np.array_equal(maps, np.where(north, 0, 1))
        f32 = np.float32
        pos = (e - np.float64(f32(-2.0))) / np.float64(f32(d)) + 0.5
        want = np.floor(pos).astype(np.int64)
        edge = np.abs(pos - np.round(pos)) < 1e-6
        assert np.array_equal(got[~edge], want[~edge]), axis
        # a planted mistake: the south map with the north map's sign of y
        if axis == 1:
            bad = np.floor((rho * -np.cos(az) + 2.0) / d + 0.5).astype(np.int64)
            assert not np.array_equal(got[~north & ~edge], bad[~north & ~edge])


def test_coadd_counts_and_map_equal_histogram():
    rng = np.random.default_rng(9)
    n = 20000
    zm = 3
    offs = np.array([[0.5, 0.0, -0.3], [1.0, 0.2, 0.0]], np.float32)
    p = params(nx=36, ny=18, nz=12, x0=5.0, dx=10.0, y0=-85.0, dy=10.0, z0=14.0, dz=1.0, zmag_width=zm,
               coadd=offs)
    x = rng.uniform(0, 360, n)
    y = rng.uniform(-90, 90, n)
    z = rng.uniform(12, 28, (zm, n)).astype(np.float32)
    hidden = (rng.random(n) < 0.2).astype(np.int32)
    c, beams, tot = run(p, x, y, z, hidden)
    vis = hidden == 0
    assert tot == zm * vis.sum()

    def nb(v, v0, dv, nn):
        return np.clip(np.floor((v - np.float64(np.float32(v0))) / np.float64(np.float32(dv)) + 0.5), 0, nn - 1) \
            .astype(np.int64)
    X, Y = nb(x, 5.0, 10.0, 36), nb(y, -85.0, 10.0, 18)
    beam = X * 18 + Y
    want = np.zeros((36 * 18, 12, zm + 2), np.int64)
    for i in range(zm):
        np.add.at(want, (beam[vis], nb(z[i], 14.0, 1.0, 12)[vis], i), 1)
    for k in range(2):
        cut = np.min(z - offs[k][:, None], axis=0)
        db = nb(cut, 14.0, 1.0, 12)
        # the row counts in every bin from its depth bin up
        for Z in range(12):
            np.add.at(want, (beam[vis & (db <= Z)], Z, zm + k), 1)
    assert np.array_equal(c, want)
    assert np.array_equal(beams.astype(bool), want[:, :, :zm].sum(axis=(1, 2)) > 0)
    # coadd columns are non-decreasing in Z and end at the beam's row count
    assert np.all(np.diff(c[:, :, zm:], axis=1) >= 0)
    rows = np.bincount(beam[vis], minlength=36 * 18)
    assert np.array_equal(c[:, -1, zm], rows) and np.array_equal(c[:, -1, zm + 1], rows)
    # a planted mistake: a coadd counted only at its depth bin
    once = np.zeros_like(want[:, :, zm])
    np.add.at(once, (beam[vis], nb(np.min(z - offs[0][:, None], axis=0), 14.0, 1.0, 12)[vis]), 1)
    assert not np.array_equal(once, c[:, :, zm])


def test_nan_magnitude_skipped_in_coadd_depth():
    # std::min(cut, m) keeps cut when m is NaN: a row with one NaN band still
    # reaches its depth from the other bands
    p = params(nx=1, ny=1, nz=10, x0=0.0, dx=1.0, y0=0.0, dy=1.0, z0=10.0, dz=1.0, zmag_width=2,
               coadd=np.zeros((1, 2)))
    c, beams, tot = run(p, [0.0], [0.0], np.array([[np.nan], [13.2]]))
    assert c[0, :, 2].tolist() == [0, 0, 0, 1, 1, 1, 1, 1, 1, 1]
    assert c[0, 0, 0] == 1            # the NaN magnitude itself lands in bin 0
