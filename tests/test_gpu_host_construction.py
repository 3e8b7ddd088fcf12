"""The catalog's host-built parameter blocks against a construction made
here from the same configuration files, so that a mistake in the library's
host code (a frame sign, a normalisation, a default, a table wiring) cannot
hide behind the oracle replay, which reads these blocks from the library.

  models       kind, component, shape parameters, cuts (model_expDisk.cpp:
               40-90, model_powerLawEllipsoid.cpp:42-102); (T, M) from the
               frame geometry and scipy rotations; f = f0 / rho(LF site)
               with rho from its definition (model_lib.h:68-95); the LF
               texture as the clipped linear resampling of the file
               (common/cux.cpp:336-384)
  FeH          the defaults and keys of os_FeH::construct (FeH_host.cpp:85-123)
  Bond2010     the defaults and keys of os_Bond2010::construct
               (Bond2010_host.cpp:136-227)
  vel2pm       os_vel2pm::construct (vel2pm_host.cpp:76-95)
  photometry   bands, bootstrap index, reddening, colour tables from the
               system file (photometry_host.cpp:175-349)
  multiples    the fraction and LF tables (unresolvedMultiples_host.cpp:140-180)
  photoerr     the per-band sigma(m) knots in band order
  extinction   the per-pixel minima (os_skygen.cpp:285-341): for every
               pixel and distance modulus, a lower bound of the map's
               trilinear value anywhere inside the pixel, and no lower than
               the smallest texel near it
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator
from scipy.spatial.transform import Rotation

import galfast_amd as G
from galfast_amd import _abi as A
from galfast_amd import configs
from galfast_amd.catalog import Catalog
import gfx
import replay as R
from test_oracle_pins import galplane_geometric, unit

pytestmark = pytest.mark.gpu


def read_conf(path):
    kv = {}
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if "=" in line:
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    return kv


@pytest.fixture(scope="module")
def job(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("hc"))
    cmd = configs.write_c4_small(d)
    with open(os.path.join(d, "model.thin.conf"), "a") as f:
        f.write("rcut = 15000\n")
    with open(os.path.join(d, "model.halo.conf"), "a") as f:
        f.write("orientation = galplane axisangle 0 0 1 20\nrmin = 1000\nrmax = 40000\n"
                "lf_l = 20\nlf_b = 40\nlf_D = 1500\n")
    with open(os.path.join(d, "Bond2010.conf"), "w") as f:
        f.write("module = Bond2010\ncomp_thin = 0\ncomp_thick = 1\ncomp_halo = 2\n"
                "HsigmaR = 140 0 0 0 0\nHsigmaPhi = 105 0 0 0 0\nHsigmaTh = 75 0 0 0 0\nfk = 2.5\n")
    with open(os.path.join(d, "vel2pm.conf"), "w") as f:
        f.write("module = vel2pm\ncoordsys = equ\nvLSR = -230\n")
    txt = open(cmd).read().replace("modules = feh.conf multiples.conf photometry.conf photoerr.conf",
                                   "modules = feh.conf multiples.conf Bond2010.conf photometry.conf photoerr.conf "
                                   "vel2pm.conf")
    open(cmd, "w").write(txt)
    cat = Catalog(cmd, seed=7)
    view = R.host_view(G.lib(), cat._h)
    yield d, cat, view
    cat.close()


def rho_unit(kind, T, M, site, mc):
    x, y, z = M @ (site + T)
    if kind == A.GF_MODEL_EXPDISK:
        return math.exp(-math.hypot(x, y) / float(mc["l"]) - abs(z) / float(mc["h"]))
    ba, ca, n = float(mc.get("b_a", 1)), float(mc.get("c_a", 1)), float(mc.get("n", -3))
    return (x * x + (y / ba) ** 2 + (z / ca) ** 2) ** (0.5 * n)


def lf_texture_np(path, nsamp=1024):
    X, Y = np.loadtxt(path, unpack=True)
    rng = X[-1] - X[0]
    X = np.r_[X[0] - 1e-5 * rng, X, X[-1] + 1e-5 * rng]
    Y = np.r_[0.0, Y, 0.0]
    dX = (X[-1] - X[0]) / (nsamp - 1)
    return np.interp(X[0] + np.arange(nsamp) * dX, X, Y), X[0], 1.0 / dX


def test_models(job):
    d, cat, view = job
    defs = read_conf(os.path.join(d, "definitions.conf"))
    Rg, z0 = float(defs["Rg"]), float(defs["z0"])
    names = ["model.thin.conf", "model.thick.conf", "model.halo.conf"]
    assert view.nmodels == 3
    Tg, Mg = galplane_geometric(Rg, z0)
    for i, name in enumerate(names):
        mc = read_conf(os.path.join(d, name))
        m = view.models[i]
        kind = A.GF_MODEL_EXPDISK if mc["model"] == "exponential disk" else A.GF_MODEL_POWERLAW
        assert m.kind == kind and m.comp == int(mc["comp"])
        T = np.array(m.T[:], dtype=np.float64)
        M = np.array([m.M[r][:] for r in range(3)], dtype=np.float64)
        Mw = Mg
        if "orientation" in mc:
            w = mc["orientation"].split()
            assert w[:2] == ["galplane", "axisangle"]
            ax = np.array([float(v) for v in w[2:5]])
            Mw = Rotation.from_rotvec(math.radians(float(w[5])) * ax / np.linalg.norm(ax)).as_matrix().T @ Mg
        assert np.abs(T - Tg).max() < 1e-3 and np.abs(M - Mw).max() < 1e-6, name
        site = float(mc.get("lf_D", 0)) * unit(math.radians(float(mc.get("lf_l", 0))),
                                               math.radians(float(mc.get("lf_b", 0))))
        f = float(mc.get("f", 1)) / rho_unit(kind, Tg, Mw, site, mc)
        assert abs(m.f / f - 1) < 2e-5, (name, m.f, f)
        assert m.norm == 1.0
        if kind == A.GF_MODEL_EXPDISK:
            assert m.l == np.float32(mc["l"]) and m.h == np.float32(mc["h"])
            assert m.r_cut2 == np.float32(float(mc.get("rcut", 0))) ** 2
        else:
            assert m.n == np.float32(mc["n"]) and m.ca == np.float32(mc.get("c_a", 1))
            assert m.ba == np.float32(mc.get("b_a", 1))
            assert m.rminSq == np.float32(float(mc.get("rmin", 0))) ** 2
            assert m.rmaxSq == np.float32(float(mc.get("rmax", 0))) ** 2
        tex, x0, idx = lf_texture_np(os.path.join(d, mc["lumfunc"]))
        got = np.ctypeslib.as_array(C.cast(m.lf.data, C.POINTER(C.c_float)), shape=(m.lf.n,))
        assert m.lf.n == 1024
        np.testing.assert_allclose(got, tex, rtol=2e-7, atol=1e-12)
        assert m.lf.tc[0] == np.float32(x0) and m.lf.tc[1] == np.float32(idx)
    # planted mistake: the halo's normalisation at the Sun instead of its LF site
    mc = read_conf(os.path.join(d, names[2]))
    wrong = float(mc["f"]) / rho_unit(A.GF_MODEL_POWERLAW, Tg, Mg, np.zeros(3), mc)
    assert abs(view.models[2].f / wrong - 1) > 1e-2


def test_module_blocks(job):
    d, cat, view = job
    defs = read_conf(os.path.join(d, "definitions.conf"))
    Rg, z0 = float(defs["Rg"]), float(defs["z0"])
    Tg, Mg = galplane_geometric(Rg, z0)
    ch = view.chain
    # FeH: the feh.conf keys (every one is given there)
    fc = read_conf(os.path.join(d, "feh.conf"))
    p = ch.feh.contents
    A0, A1 = np.float32(fc["A0"]), np.float32(fc["A1"])
    s = np.float64(A0 + A1)
    assert p.A[0] == np.float32(A0 / s) and p.A[1] == np.float32(A1 / s)
    assert list(p.sigma) == [np.float32(fc[k]) for k in ("sigma0", "sigma1", "sigmaH")]
    assert list(p.offs) == [np.float32(fc[k]) for k in ("offs0", "offs1", "offsH")]
    assert (p.Hmu, p.muInf, p.DeltaMu) == tuple(np.float32(fc[k]) for k in ("Hmu", "muInf", "deltaMu"))
    assert p.comp_thin.bits[0] == 1 and p.comp_thick.bits[0] == 2 and p.comp_halo.bits[0] == 4
    assert np.abs(np.array(p.T[:]) - Tg).max() < 1e-3
    assert np.abs(np.array([p.M[r][:] for r in range(3)]) - Mg).max() < 1e-7
    # Bond2010: the reference defaults (gfx.bond_params) with the halo and fk keys of Bond2010.conf
    b = ch.bond.contents
    want = gfx.bond_params(halo_sigma=(140, 75, 105))
    assert abs(b.fk / (2.5 / 3.5) - 1) < 1e-7
    for part in ("disk_1", "disk_2", "halo"):
        gv, wv = getattr(b, part), getattr(want, part)
        assert [list(r) for r in gv.m] == [list(r) for r in wv.m], part
        assert [list(r) for r in gv.ss] == [list(r) for r in wv.ss], part
    assert np.abs(np.array(b.T[:]) - Tg).max() < 1e-3
    assert np.abs(np.array([b.M[r][:] for r in range(3)]) - Mg).max() < 1e-7
    # vel2pm: coordsys equ (1), vLSR from the file, reference defaults for the rest
    v = ch.v2pm.contents
    assert v.coordsys == 1 and v.vLSR == np.float32(-230.0) and v.Rg == np.float32(Rg)
    assert (v.u0, v.v0, v.w0) == (np.float32(-10.0), np.float32(-5.3), np.float32(7.2))
    # photometry: 6 bands, bootstrap SYNr, reddening as listed, tables of photosys.txt
    pc = read_conf(os.path.join(d, "photometry.conf"))
    bands = pc["bands"].split()
    ph = ch.phot.contents
    assert ph.ncolors == len(bands) - 1 and ph.bidx == bands.index(pc["bootstrap_band"])
    red = [float(x) for x in pc["reddening"].split()]
    assert list(ph.reddening[:len(red)]) == [np.float32(x) for x in red]
    assert ch.phot_comps.bits[0] == 7
    rows = np.loadtxt(os.path.join(d, pc["file"]))
    ntex, nF, nM, tc, cols, flags = gfx.photometry_tables(rows, ph.ncolors, float(pc["dFeH"]), float(pc["dM"]))
    assert (ph.nFeH, ph.nMr) == (nF, nM)
    assert list(ph.tcFeH) + list(ph.tcMr) == list(tc)
    sz = nF * nM * 4
    for i in range(ntex):
        got_c = np.ctypeslib.as_array(C.cast(ph.color[i], C.POINTER(C.c_float)), shape=(sz,))
        got_f = np.ctypeslib.as_array(C.cast(ph.cflags[i], C.POINTER(C.c_float)), shape=(sz,))
        assert np.array_equal(got_c, cols[i * sz:(i + 1) * sz]) and np.array_equal(got_f, flags[i * sz:(i + 1) * sz])
    # multiples: LF_M2_gt_M1 with fraction 0.5 and the LF of lf.txt
    mu = ch.mult.contents
    assert mu.algo == 1 and ch.mult_comps.bits[0] == 7
    lfM, lfphi = np.loadtxt(os.path.join(d, "lf.txt"), unpack=True)
    (sec, sec_tc), (cum, cum_tc), (inv, inv_tc) = gfx.multiples_tables(lfM, lfphi, fraction=0.5)
    for t, (arr, tcw) in ((mu.secProb, (sec, sec_tc)), (mu.cumLF, (cum, cum_tc)), (mu.invCumLF, (inv, inv_tc))):
        got = np.ctypeslib.as_array(C.cast(t.data, C.POINTER(C.c_float)), shape=(t.n,))
        assert t.n == len(arr) and np.allclose(got, arr, rtol=1e-6, atol=1e-7)
        assert abs(t.tc[0] - tcw[0]) <= 1e-6 * max(1.0, abs(tcw[0])) and abs(t.tc[1] / tcw[1] - 1) < 1e-6
    # photometric errors: sigma(m) knots of each band's file, in band order
    pe = ch.perr.contents
    assert pe.nbands == len(bands)
    kx = np.ctypeslib.as_array(C.cast(pe.kx, C.POINTER(C.c_double)), shape=(pe.koff[pe.nbands],))
    ky = np.ctypeslib.as_array(C.cast(pe.ky, C.POINTER(C.c_double)), shape=(pe.koff[pe.nbands],))
    for i, bname in enumerate(bands):
        mm, ss = np.loadtxt(os.path.join(d, "SYN.%s.photoerr.txt" % bname), unpack=True)
        a, e = pe.koff[i], pe.koff[i + 1]
        assert np.array_equal(kx[a:e], mm) and np.array_equal(ky[a:e], ss), bname


def _cube(tex):
    data = np.ctypeslib.as_array(C.cast(tex.data, C.POINTER(C.c_float)), shape=(tex.nz, tex.ny, tex.nx))
    grids = [tex.tc[a][0] + np.arange((tex.nx, tex.ny, tex.nz)[a]) / np.float64(tex.tc[a][1]) for a in range(3)]
    return data.astype(np.float64), grids


def test_extinction_minima_bound_the_map_inside_each_pixel(job):
    d, cat, view = job
    ext = view.ext
    assert ext.north.data and ext.south.data and ext.beam.data
    sp = view.sp
    pix = (A.gf_pencil_beam * sp.npixels).from_address(view.pixels)
    maps = [_cube(ext.north), _cube(ext.south)]
    beam, bgrid = _cube(ext.beam)
    interp = [RegularGridInterpolator((g[2], g[1], g[0]), c) for c, g in maps]
    rng = np.random.default_rng(1)
    DMs = np.linspace(0.0, 28.0, 57)
    nlow = 0
    for i in range(sp.npixels):
        p = pix[i]
        assert p.extIdx == i
        h = p.projIdx
        c, g = maps[h]
        # 64 random points of the pixel square in its own projection
        x = p.X + p.dx * (rng.uniform(size=64) - 0.5)
        y = p.Y + p.dx * (rng.uniform(size=64) - 0.5)
        X, D = np.meshgrid(x, DMs)
        Y, _ = np.meshgrid(y, DMs)
        pts = np.stack([np.clip(D.ravel(), g[2][0], g[2][-1]), np.clip(Y.ravel(), g[1][0], g[1][-1]),
                        np.clip(X.ravel(), g[0][0], g[0][-1])], axis=1)
        am = interp[h](pts).reshape(D.shape)
        # the minimum table, linear in DM between its slices
        col = beam[:, i % 2048, i // 2048]           # texel (x, y) = (i div 2048, i mod 2048)
        amin = np.interp(np.clip(DMs, bgrid[2][0], bgrid[2][-1]), bgrid[2], col)
        assert np.all(amin[:, None] <= am + 1e-6), i
        # not below the smallest texel within two pixel widths (a superset of the stencils)
        xs, ys = g[0], g[1]
        ix = np.flatnonzero(np.abs(xs - p.X) <= 2 * p.dx + 2 * (xs[1] - xs[0]))
        iy = np.flatnonzero(np.abs(ys - p.Y) <= 2 * p.dx + 2 * (ys[1] - ys[0]))
        for z in range(1, beam.shape[0]):
            lo = c[z - 1][np.ix_(iy, ix)].min()
            assert col[z] >= lo - 1e-7, (i, z)
        nlow += int((col[1:] > 0).sum())
    assert nlow > 0
    # planted mistake: minima taken at the same slice instead of one closer
    # would lie above the library's at the pixel centre's own texels
    p = pix[0]
    c, g = maps[p.projIdx]
    i0 = int((p.X - g[0][0]) / (g[0][1] - g[0][0]))
    j0 = int((p.Y - g[1][0]) / (g[1][1] - g[1][0]))
    ix, iy = [i0, i0 + 1], [j0, j0 + 1]            # the bilinear stencil of the pixel centre
    same = np.array([c[z][np.ix_(iy, ix)].min() for z in range(c.shape[0])])
    assert np.any(same[1:] > beam[1:, 0, 0] + 1e-6)


def test_no_star_is_below_its_pixel_minimum(job):
    """With minima that bound the map, the draw never marks a star's
    extinction as below its pixel's minimum (AmInf < 0)."""
    d, cat, view = job
    neg = [0]
    total = [0]

    def on_chunk(t, n, comp):
        vis = t.arr("hidden")[0][:n] == 0
        neg[0] += int((t.arr("AmInf")[0][:n][vis] < 0).sum())
        total[0] += int(vis.sum())

    R.replay(view, pixel_index=list(range(0, view.sp.npixels, 3)), on_chunk=on_chunk)
    assert total[0] > 1000 and neg[0] == 0
