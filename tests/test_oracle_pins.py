"""Pins of the oracle's sky-generation functions against what the mathematics
fixes, each computed here in FP64 numpy from its definition (not from the
oracle's code), and each with a planted mistake that the same comparison
must reject, so the check has the power to catch one.

  density      f exp(-r/l - |z|/h) and f (x^2 + (y/ba)^2 + (z/ca)^2)^(n/2)
               in the Galactic-plane frame built from its geometry
               (skygen/model_expDisk.h:71-121, model_powerLawEllipsoid.h:60-105)
  LF norm      rho(LF site) = f0, cuts ignored (skygen/model_lib.h:68-95,
               model_expDisk.cpp:81-89)
  count        a constant-density model with a flat LF: per (pixel, DM bin)
               f phi D^3 dOmega (0.2 ln 10) dm dM times the number of valid
               M bins -- the Euclidean N(m) ~ 10^(0.6 m) -- for the bins
               form and the reference's crawl (skygen/skygen.cu.h:469-618)
  emit / draw  per star: |XYZ| = 10^(0.2 DM + 1), XYZ along (l, b), (l, b)
               the Lambert deprojection of projXY inside its pixel, M and m
               uniform within their bins, M bins drawn with LF weights,
               hidden = (m + Am > m1) or outside the footprint, Am the
               trilinear map value, AmInf the far value or -AmMin
               (skygen/skygen.cu.h:366-457)
  clipper      hidden |= outside the footprint at the deprojected position
               (os_clipper, skygen/os_skygen.cpp)
  tex3d        scipy's trilinear interpolation with clamped coordinates
               (common/cux.h:849-874)
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats
from scipy.interpolate import RegularGridInterpolator

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O
import gfx

P_OK = 1e-4
P_BAD = 1e-8
LN10 = math.log(10.0)


def o():
    lib = O.lib()
    lib.gfo_tex3d.restype = C.c_float
    lib.gfo_tex3d.argtypes = [C.POINTER(A.gf_tex3d), C.c_float, C.c_float, C.c_float]
    return lib


# ---------------------------------------------------------------------------
# geometry from definitions
# ---------------------------------------------------------------------------
def galplane_geometric(Rg, z0):
    """The Galactic-plane frame from its definition: origin at the Galactic
    centre, which is (Rg, 0, 0) of the Sun-centred Galactic frame (x toward
    l = 0, z toward the north Galactic pole); the Sun at distance Rg from the
    centre and height z0 above the plane, on the frame's +x side; the
    frame's y axis opposite the Galactic y axis.  Returns (T, M) with
    v' = M (v + T); the columns of M are the images of the Galactic axes."""
    d = math.sqrt(Rg * Rg - z0 * z0)
    sun = np.array([d, 0.0, z0])
    ex = -sun / Rg                    # Sun -> centre
    ey = np.array([0.0, -1.0, 0.0])
    ez = np.cross(ex, ey)             # right-handed
    return np.array([-Rg, 0.0, 0.0]), np.column_stack([ex, ey, ez])


def lambert_project(l0, b0, l, b):
    """Lambert azimuthal equal-area projection about (l0, b0) (Snyder 1987,
    eqs. 24-2 to 24-4), radians."""
    cosd = np.cos(l - l0)
    k = np.sqrt(2.0 / (1.0 + math.sin(b0) * np.sin(b) + math.cos(b0) * np.cos(b) * cosd))
    return k * np.cos(b) * np.sin(l - l0), k * (math.cos(b0) * np.sin(b) - math.sin(b0) * np.cos(b) * cosd)


def lambert_deproject(l0, b0, x, y):
    """Inverse of lambert_project (Snyder 1987, eqs. 20-14, 24-16, 24-18)."""
    rho = np.hypot(x, y)
    c = 2.0 * np.arcsin(np.minimum(rho / 2.0, 1.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        b = np.arcsin(np.cos(c) * math.sin(b0) + np.where(rho > 0, y * np.sin(c) * math.cos(b0) / rho, 0.0))
        l = l0 + np.arctan2(x * np.sin(c), rho * math.cos(b0) * np.cos(c) - y * math.sin(b0) * np.sin(c))
    return np.mod(l, 2 * math.pi), b


def unit(l, b):
    return np.stack([np.cos(b) * np.cos(l), np.cos(b) * np.sin(l), np.sin(b)])


def worst(got, want, rel=True):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    d = np.abs(got - want)
    if rel:
        d = d / np.maximum(np.abs(want), 1e-300)
    return float(d.max())


# ---------------------------------------------------------------------------
# density and normalisation
# ---------------------------------------------------------------------------
def _model(kind, T, M, **kw):
    m = A.gf_model()
    m.kind = kind
    gfx.set_TM(m, T, M)
    for k, v in kw.items():
        setattr(m, k, v)
    m.norm = 1.0
    return m


def _rho_oracle(m, pts):
    lib = o()
    return np.array([lib.gfo_model_rho(C.byref(m), *map(float, p)) for p in pts.astype(np.float32)])


def _frame(pts, T, M):
    return (np.asarray(M, dtype=np.float64) @ (pts.astype(np.float32).astype(np.float64) + T).T)


def _sample_points(n, seed):
    rng = np.random.default_rng(seed)
    # Galactic cartesian points out to 20 kpc around the Sun, both sides of the plane
    d = 10 ** rng.uniform(1.0, 4.3, n)
    l = rng.uniform(0, 2 * math.pi, n)
    b = np.arcsin(rng.uniform(-1, 1, n))
    return (d * unit(l, b)).T


def test_expdisk_density_against_fp64():
    Rg, z0 = 8000.0, 25.0
    T, M = galplane_geometric(Rg, z0)
    f, L, H, rcut = 0.037, 2600.0, 300.0, 12000.0
    pts = _sample_points(4000, 1)
    m = _model(A.GF_MODEL_EXPDISK, T, M, f=f, l=L, h=H, r_cut2=np.float32(rcut) * np.float32(rcut))
    got = _rho_oracle(m, pts)

    def rho(pts, T, M, absz=True):
        x, y, z = _frame(pts, T, M)
        zz = np.abs(z) if absz else z
        v = f * np.exp(-np.hypot(x, y) / L - zz / H)
        return np.where(x * x + y * y + z * z > rcut * rcut, 0.0, v)

    x, y, z = _frame(pts, T, M)
    away = np.abs(np.sqrt(x * x + y * y + z * z) - rcut) > 1.0      # clear of the cut
    assert (~away).sum() < 10
    want = rho(pts, T, M)
    assert (want[away] == 0).sum() > 100 and (want[away] > 0).sum() > 1000
    assert np.array_equal(got[away] == 0, want[away] == 0)
    pos = away & (want > 0)
    assert worst(got[pos], want[pos]) < 2e-5
    # planted mistakes: z for |z|, the plane frame with the Sun below the plane
    assert worst(got[pos], rho(pts, T, M, absz=False)[pos]) > 1e-2
    Tb, Mb = galplane_geometric(Rg, -z0)
    assert worst(got[pos], rho(pts, Tb, Mb)[pos]) > 1e-3


def test_powerlaw_density_against_fp64():
    Rg, z0 = 8000.0, 25.0
    T, M = galplane_geometric(Rg, z0)
    f, n, ba, ca, rmin, rmax = 0.0051, -2.77, 0.8, 0.64, 4000.0, 14000.0
    pts = _sample_points(4000, 2)
    m = _model(A.GF_MODEL_POWERLAW, T, M, f=f, n=n, ba=ba, ca=ca,
               rminSq=np.float32(rmin) * np.float32(rmin), rmaxSq=np.float32(rmax) * np.float32(rmax))
    got = _rho_oracle(m, pts)

    def rho(pts, ba=ba, ca=ca, half=0.5):
        x, y, z = _frame(pts, T, M)
        r2 = x * x + (y / ba) ** 2 + (z / ca) ** 2
        D2 = x * x + y * y + z * z
        return np.where((D2 < rmin * rmin) | (D2 >= rmax * rmax), 0.0, f * r2 ** (half * n))

    x, y, z = _frame(pts, T, M)
    D = np.sqrt(x * x + y * y + z * z)
    away = (np.abs(D - rmin) > 0.5) & (np.abs(D - rmax) > 0.5)
    want = rho(pts)
    assert (want[away] == 0).sum() > 10 and (want[away] > 0).sum() > 1000
    assert np.array_equal(got[away] == 0, want[away] == 0)
    pos = away & (want > 0)
    assert worst(got[pos], want[pos]) < 2e-5
    # planted mistakes: the axis ratios swapped, n for n/2
    assert worst(got[pos], rho(pts, ba=ca, ca=ba)[pos]) > 1e-2
    assert worst(got[pos], rho(pts, half=1.0)[pos]) > 1e-2


@pytest.mark.parametrize("kind", [A.GF_MODEL_EXPDISK, A.GF_MODEL_POWERLAW])
def test_lf_normalisation_sets_rho_at_the_site(kind):
    T, M = galplane_geometric(8000.0, 25.0)
    lib = o()
    for site_lbd, f0 in [((0.0, 0.0, 0.0), 1.0), ((30.0, 60.0, 2500.0), 0.12), ((250.0, -40.0, 900.0), 0.005)]:
        l, b, d = site_lbd
        site = (d * unit(math.radians(l), math.radians(b))).astype(np.float32)
        kw = dict(l=2600.0, h=300.0, r_cut2=1.0) if kind == A.GF_MODEL_EXPDISK else \
            dict(n=-2.77, ba=1.0, ca=0.64, rminSq=1.0, rmaxSq=4.0)    # cuts that exclude the site
        m = _model(kind, T, M, f=1.0, **kw)
        f = lib.gfo_lf_norm_f(m, C.c_float(f0), *map(float, site))
        # the density with f = 1 at the site, from its definition, cuts ignored
        x, y, z = _frame(site[None, :], T, M)[:, 0]
        if kind == A.GF_MODEL_EXPDISK:
            unit_rho = math.exp(-math.hypot(x, y) / 2600.0 - abs(z) / 300.0)
        else:
            unit_rho = (x * x + y * y + (z / 0.64) ** 2) ** (-2.77 / 2)
        assert abs(f / (f0 / unit_rho) - 1) < 2e-5
        m2 = _model(kind, T, M, f=f, **{k: 0.0 if k.startswith("r") else v for k, v in kw.items()})
        assert abs(lib.gfo_model_rho(C.byref(m2), *map(float, site)) / f0 - 1) < 2e-5   # model_lib.h:91


# ---------------------------------------------------------------------------
# absolute count: constant density, flat LF
# ---------------------------------------------------------------------------
def _flat_lf(phi=0.003, M0=3.0, M1=15.0):
    Ms = np.round(np.arange(M0, M1 + 0.05, 0.1), 8)
    return gfx.lf_texture(Ms, np.full(len(Ms), phi))


def _count_setup(f0=0.02, m0=14.0, m1=22.0, M0=4.0, M1=8.0):
    pix = gfx.pixelize_rect(20.0, 30.0, 40.0, 50.0, 1.0)
    sp = gfx.skygen_params(m0, m1, 0.1, M0, M1, len(pix))
    tex, tc = _flat_lf()
    m = gfx.make_model(A.GF_MODEL_POWERLAW, 0, f0, tex, tc, n=0.0, ca=1.0, ba=1.0)
    return sp, pix, m, (tex, tc)


def _closed_form_counts(sp, pix, f0, phi, D_power=3, m1=None):
    """W[p, k]: f0 phi D_k^D_power dOmega_p (0.2 ln 10) dm dM n_k, with D_k the
    distance of DM = m0 - M1 + k dm and n_k the M bins of the diagonal inside
    the m grid and under the flux limit; NaN where no bin is valid."""
    K = sp.nm + sp.nM - 1
    m0, M1, dm, dM = (np.float64(np.float32(v)) for v in (sp.m0, sp.M1, sp.dm, sp.dM))
    k = np.arange(K)
    DM = m0 - M1 + k * dm
    D = 10 ** (0.2 * DM + 1)
    lo = np.maximum(0, k - (sp.nm - 1))
    hi = np.minimum(sp.nM - 1, k)
    nk = np.zeros(K)
    for kk in range(K):
        im = kk - np.arange(lo[kk], hi[kk] + 1)
        ok = m0 + dm * im <= (sp.m1 if m1 is None else m1)
        nk[kk] = ok.sum()
    dA = np.array([p.dA for p in pix], dtype=np.float64)
    W = f0 * phi * (D ** D_power)[None, :] * dA[:, None] * 0.2 * LN10 * dm * dM * nk[None, :]
    return np.where(nk[None, :] > 0, W, np.nan)


def _bins(sp, m, pix):
    K = sp.nm + sp.nM - 1
    wpk = np.zeros(len(pix) * K)
    res = A.gf_count_result()
    O.lib().gfo_count_bins(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, C.byref(res), None,
                           O.ptr(wpk))
    return wpk.reshape(len(pix), K), res


@pytest.mark.parametrize("flux_cut", [None, 19.93])
def test_count_of_a_constant_density_model_is_euclidean(flux_cut):
    f0, phi = 0.02, 0.003
    sp, pix, m, keep = _count_setup(f0=f0)
    if flux_cut is not None:
        sp.m1 = flux_cut           # below the top of the m grid: a diagonal loses its faint cells
    W, res = _bins(sp, m, pix)
    want = _closed_form_counts(sp, pix, f0, phi)
    valid = ~np.isnan(want)
    assert np.array_equal(W > 0, valid) and np.all(W[~valid] == -1.0)
    assert worst(W[valid], want[valid]) < 3e-5
    # the Euclidean slope: where the M window is full, W[k+1] / W[k] = 10^(0.6 dm)
    nk = np.round(want[0] / want[0].max() * 1e9)
    k = np.arange(W.shape[1] - 1)
    full = valid[0, :-1] & valid[0, 1:] & (sp.nM - 1 <= k) & (k + 1 <= sp.nm - 1)
    if flux_cut is not None:
        full &= np.arange(W.shape[1] - 1) + 1 < int((flux_cut - sp.m0) / sp.dm)
    assert full.sum() > 10
    ratio = W[:, 1:][:, full] / W[:, :-1][:, full]
    assert np.allclose(ratio, 10 ** (0.6 * sp.dm), rtol=3e-5)
    # planted mistakes: D^2 (area without depth) and a volume element of 0.4 ln 10
    bad = _closed_form_counts(sp, pix, f0, phi, D_power=2)
    assert worst(W[valid], bad[valid]) > 0.5
    assert worst(W[valid], 2 * want[valid]) > 0.4
    # the reference's crawl totals the same expectation, covered or not
    crawl = A.gf_count_result()
    O.lib().gfo_count_crawl(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, C.byref(crawl), None)
    cov = np.array([p.coveredFraction for p in pix])
    assert abs(crawl.nstarsExpectedToGenerate / np.nansum(want) - 1) < 3e-5
    assert abs(crawl.nstarsExpected / np.nansum(want * cov[:, None]) - 1) < 3e-5
    assert abs(res.nstarsExpected / np.nansum(want * cov[:, None]) - 1) < 3e-5


def test_count_distance_cut():
    f0, phi = 0.02, 0.003
    sp, pix, m, keep = _count_setup(f0=f0)
    sp.dmin, sp.dmax = 3000.0, 20000.0
    W, _ = _bins(sp, m, pix)
    want = _closed_form_counts(sp, pix, f0, phi)
    K = W.shape[1]
    DM = np.float64(np.float32(sp.m0)) - np.float64(np.float32(sp.M1)) + np.arange(K) * np.float64(np.float32(sp.dm))
    D = 10 ** (0.2 * DM + 1)
    inside = (D >= 3000.0 * (1 + 1e-5)) & (D < 20000.0 * (1 - 1e-5))
    edge = ~inside & (D > 3000.0 * (1 - 1e-5)) & (D < 20000.0 * (1 + 1e-5))
    k_ok = inside & ~np.isnan(want[0])
    assert k_ok.sum() > 5 and (~inside & ~edge).sum() > 5
    assert worst(W[:, k_ok], want[:, k_ok]) < 3e-5
    assert np.all(W[:, ~inside & ~edge] == -1.0)


# ---------------------------------------------------------------------------
# emission: per-star formulas
# ---------------------------------------------------------------------------
def _linear_ext_cube(hemi, n=24, nz=11):
    """A map linear in (x, y, DM): trilinear sampling reproduces it inside
    the grid.  Returns (tex3d, data, f) with f(x, y, DM) the FP64 value."""
    xs = np.linspace(-1.5, 1.5, n)
    zs = np.linspace(0.0, 20.0, nz)
    a = (0.05, 0.02, -0.03, 0.01) if hemi == 0 else (0.08, -0.01, 0.02, 0.012)
    f = lambda x, y, z: a[0] + a[1] * x + a[2] * y + a[3] * z
    Z, Y, X = np.meshgrid(zs, xs, xs, indexing="ij")
    data = np.ascontiguousarray(f(X, Y, Z), dtype=np.float32)
    t = A.gf_tex3d()
    t.data = data.ctypes.data
    t.nx, t.ny, t.nz = n, n, nz
    for ax, g in enumerate((xs, xs, zs)):
        t.tc[ax][0] = g[0]
        t.tc[ax][1] = 1.0 / (g[1] - g[0])
    clampz = lambda z: np.clip(z, zs[0], zs[-1])
    return t, data, (lambda x, y, z: f(x, y, clampz(z)))


class Emitted:
    pass


def _emit(with_ext=False, with_foot=False, target=150_000, seed=11):
    lf_M, lf_phi = gfx.seeded_lf()
    tex, tc = gfx.lf_texture(lf_M, lf_phi)
    pix = gfx.pixelize_rect(30.0, 40.0, 50.0, 60.0, 1.0)
    npix = len(pix)
    sp = gfx.skygen_params(15.0, 22.0, 0.1, 4.0, 10.0, npix)
    m = gfx.make_model(A.GF_MODEL_POWERLAW, 2, 0.005, tex, tc, n=-2.77, ca=0.64)
    ext = None
    keep = []
    fext = None
    amin = np.zeros(npix)
    if with_ext:
        ext = A.gf_extinction()
        tn, dn, fn = _linear_ext_cube(0)
        ts, ds, fs = _linear_ext_cube(1)
        ext.north, ext.south = tn, ts
        # per-pixel minimum table: AmMin = 0.2 on every third pixel, 0 elsewhere
        amin = np.where(np.arange(npix) % 3 == 0, 0.2, 0.0)
        beam = np.ascontiguousarray(np.repeat(amin.astype(np.float32)[None, :], 11, axis=0).reshape(11, 1, npix))
        tb = A.gf_tex3d()
        tb.data = beam.ctypes.data
        tb.nx, tb.ny, tb.nz = 1, npix, 11
        tb.tc[0][0], tb.tc[0][1] = 0.0, 1.0
        tb.tc[1][0], tb.tc[1][1] = 0.0, 1.0
        tb.tc[2][0], tb.tc[2][1] = tn.tc[2][0], tn.tc[2][1]
        ext.beam = tb
        for i in range(npix):
            pix[i].extIdx = i          # x index i // 2048 = 0, y index i
        keep += [dn, ds, beam]
        fext = fn
    foot = A.gf_footprint()
    if with_foot:
        foot = gfx.footprint(gfx.beam_region(35.0, 55.0, 4.0))
    K = sp.nm + sp.nM - 1
    U = npix * K
    lib = O.lib()
    extp = C.byref(ext) if ext is not None else None
    wpk = np.zeros(U)
    res = A.gf_count_result()
    lib.gfo_count_bins(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), extp, C.byref(res), None, O.ptr(wpk))
    m.norm = np.float32(target / wpk[wpk > 0].sum())
    lib.gfo_count_bins(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), extp, C.byref(res), None, O.ptr(wpk))
    st = gfx.mwc_state(seed)
    counts = np.zeros(U, dtype=np.uint32)
    lib.gfo_plan(O.ptr(wpk), U, O.ptr(st), 65536, O.ptr(counts))
    offs = np.zeros(U + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(counts, dtype=np.uint64)
    n = int(offs[-1])
    t = HostTable(n, catalog_spec())
    cols = t.skygen_cols()
    lib.gfo_emit(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), extp, C.byref(foot), O.ptr(offs), U, 0, n,
                 O.ptr(st), 65536, C.byref(cols))
    e = Emitted()
    e.sp, e.pix, e.m, e.t, e.n, e.wpk, e.counts, e.K = sp, pix, m, t, n, wpk, counts, K
    e.unit = np.repeat(np.arange(U), counts.astype(np.int64))
    e.lf = (tex, tc)
    e.fext, e.amin, e.foot, e.keep = fext, amin, foot, keep
    return e


def _check_star_geometry(t, n, sp, pix, pidx):
    """|XYZ|, direction, projection and pixel of every drawn star."""
    lb = t.arr("lb")
    l, b = np.radians(lb[0][:n]), np.radians(lb[1][:n])
    xyz = t.arr("XYZ")[:, :n].astype(np.float64)
    DM = t.arr("DM")[0][:n].astype(np.float64)
    D = np.linalg.norm(xyz, axis=0)
    assert worst(D, 10 ** (0.2 * DM + 1)) < 4e-6
    assert worst(D, 10 ** (0.2 * DM)) > 1.0           # planted: 10 pc lost
    assert np.abs(xyz / D - unit(l, b)).max() < 2e-6
    assert np.abs(xyz / D - unit(-l, b)).max() > 0.1    # planted: l mirrored
    proj = t.arr("projIdx")[0][:n]
    xy = t.arr("projXY")[:, :n].astype(np.float64)
    for h in (0, 1):
        s = proj == h
        if not s.any():
            continue
        pr = sp.proj[h]
        x, y = lambert_project(pr.l0, pr.b0, l[s], b[s])
        assert np.abs(x - xy[0][s]).max() < 2e-6 and np.abs(y - xy[1][s]).max() < 2e-6
    X = np.array([p.X for p in pix])[pidx]
    Y = np.array([p.Y for p in pix])[pidx]
    dx = np.array([p.dx for p in pix])[pidx]
    ux = (xy[0] - X) / dx + 0.5
    uy = (xy[1] - Y) / dx + 0.5
    assert ux.min() >= -1e-5 and ux.max() <= 1 + 1e-5 and uy.min() >= -1e-5 and uy.max() <= 1 + 1e-5
    return ux, uy


def _bin_positions(M, m, sp):
    """Fractional positions of M and m inside their bins, and the bin indices."""
    M1, dM, m0, dm = (np.float64(np.float32(v)) for v in (sp.M1, sp.dM, sp.m0, sp.dm))
    iM = np.round((M1 - M) / dM).astype(np.int64)
    uM = (M - (M1 - iM * dM)) / dM + 0.5
    im = np.round((m - m0) / dm).astype(np.int64)
    um = (m - (m0 + im * dm)) / dm + 0.5
    return iM, uM, im, um


def test_emit_per_star_formulas():
    e = _emit()
    t, n, sp, K = e.t, e.n, e.sp, e.K
    assert n > 100_000
    # Poisson plan: the total against the expectation
    W = e.wpk[e.wpk > 0].sum()
    assert abs(n - W) < 5 * math.sqrt(W)
    p_of_row = e.unit // K
    k_of_row = e.unit % K
    ux, uy = _check_star_geometry(t, n, sp, e.pix, p_of_row)
    assert stats.kstest(ux, "uniform").pvalue > P_OK and stats.kstest(uy, "uniform").pvalue > P_OK
    assert stats.kstest(np.clip(ux, 0, 1) ** 0.9, "uniform").pvalue < P_BAD
    M = t.arr("absmag")[0][:n].astype(np.float64)
    DM = t.arr("DM")[0][:n].astype(np.float64)
    m = DM + M
    iM, uM, im, um = _bin_positions(M, m, sp)
    # the cell of every star lies on its unit's DM diagonal, inside the grid
    # (m = DM + M is rebuilt here to ~2e-6 mag, so a star within 1e-4 of a
    # bin edge may round into the neighbouring bin)
    edge = np.minimum(um, 1 - um) < 1e-4
    assert np.all((im + iM == k_of_row) | edge) and edge.sum() < 4e-4 * n
    im = k_of_row - iM
    um = (m - (np.float64(np.float32(sp.m0)) + im * np.float64(np.float32(sp.dm)))) / np.float64(np.float32(sp.dm)) + 0.5
    assert iM.min() >= 0 and iM.max() < sp.nM and im.min() >= 0 and im.max() < sp.nm
    assert uM.min() > -1e-4 and uM.max() < 1 + 1e-4 and um.min() > -1e-3 and um.max() < 1 + 1e-3
    assert stats.kstest(uM, "uniform").pvalue > P_OK and stats.kstest(um, "uniform").pvalue > P_OK
    assert stats.kstest(np.clip(uM, 0, 1) ** 1.1, "uniform").pvalue < P_BAD
    # M bins: LF-weighted within the unit's window (chi-square of the bin histogram)
    tex, tc = e.lf
    xs = tc[0] + np.arange(len(tex)) / np.float64(tc[1])
    Mc = np.float64(np.float32(sp.M1)) - np.arange(sp.nM) * np.float64(np.float32(sp.dM))
    lf = np.interp(Mc, xs, tex.astype(np.float64))
    expect = np.zeros(sp.nM)
    flat = np.zeros(sp.nM)
    ku, nu = np.unique(k_of_row, return_counts=True)
    for k, cnt in zip(ku, nu):
        lo, hi = max(0, k - (sp.nm - 1)), min(sp.nM - 1, k)
        w = lf[lo:hi + 1]
        expect[lo:hi + 1] += cnt * w / w.sum()
        flat[lo:hi + 1] += cnt / (hi - lo + 1)
    obs = np.bincount(iM, minlength=sp.nM)
    assert stats.chisquare(obs, expect).pvalue > P_OK
    assert stats.chisquare(obs, flat * obs.sum() / flat.sum()).pvalue < P_BAD
    # no footprint, no extinction: hidden only past the flux limit; comp the model's
    hid = t.arr("hidden")[0][:n]
    mt = np.float32(t.arr("DM")[0][:n]) + np.float32(t.arr("absmag")[0][:n])
    clear = np.abs(mt.astype(np.float64) - sp.m1) > 1e-5
    assert np.array_equal(hid[clear] == 1, (mt > np.float32(sp.m1))[clear])
    assert np.all(t.arr("comp")[0][:n] == 2)
    assert np.all(t.arr("Am")[0][:n] == 0) and np.all(t.arr("AmInf")[0][:n] == 0)


def test_emit_extinction_and_footprint():
    e = _emit(with_ext=True, with_foot=True, target=60_000, seed=12)
    t, n, sp, K = e.t, e.n, e.sp, e.K
    p_of_row = e.unit // K
    _check_star_geometry(t, n, sp, e.pix, p_of_row)
    xy = t.arr("projXY")[:, :n].astype(np.float64)
    DM = t.arr("DM")[0][:n].astype(np.float64)
    Am = t.arr("Am")[0][:n].astype(np.float64)
    AmInf = t.arr("AmInf")[0][:n].astype(np.float64)
    assert np.all(t.arr("projIdx")[0][:n] == 0)
    want = e.fext(xy[0], xy[1], DM)
    assert np.abs(Am - want).max() < 2e-6
    assert np.abs(Am - e.fext(xy[1], xy[0], DM)).max() > 1e-3       # planted: x, y swapped
    far = e.fext(xy[0], xy[1], 100.0)
    amin = e.amin[p_of_row]
    inf = np.where(amin > Am, -amin, np.maximum(far, Am))
    assert np.abs(AmInf - inf).max() < 2e-6
    assert (amin > Am).sum() > 100 and (amin <= Am).sum() > 100
    # hidden: over the flux limit after extinction, or outside the 4 deg beam
    lb = t.arr("lb")
    v = unit(np.radians(lb[0][:n]), np.radians(lb[1][:n]))
    c = unit(math.radians(35.0), math.radians(55.0))
    cosd = c @ v
    outside = cosd < math.cos(math.radians(4.0))
    mt = np.float32(t.arr("DM")[0][:n]) + np.float32(t.arr("absmag")[0][:n])
    faint = (mt + np.float32(t.arr("Am")[0][:n])) > np.float32(sp.m1)
    clear = (np.abs(cosd - math.cos(math.radians(4.0))) > 1e-7) & \
            (np.abs(mt.astype(np.float64) + Am - sp.m1) > 1e-5)
    hid = t.arr("hidden")[0][:n] == 1
    assert np.array_equal(hid[clear], (outside | faint)[clear])
    assert outside.sum() > 1000 and faint.sum() > 1000 and (~outside & ~faint).sum() > 1000


def test_reference_draw_per_star_formulas():
    """The reference-order draw (one star per Poisson draw of every cell)
    obeys the same per-star formulas; its total is Poisson about the crawl's
    expectation."""
    lf_M, lf_phi = gfx.seeded_lf()
    tex, tc = gfx.lf_texture(lf_M, lf_phi)
    pix = gfx.pixelize_rect(30.0, 36.0, 50.0, 56.0, 1.0)
    sp = gfx.skygen_params(15.0, 21.0, 0.1, 4.0, 9.0, len(pix), nthreads=1024)
    m = gfx.make_model(A.GF_MODEL_EXPDISK, 0, 1.0, tex, tc, l=2600.0, h=300.0)
    lib = O.lib()
    crawl = A.gf_count_result()
    lib.gfo_count_crawl(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, C.byref(crawl), None)
    m.norm = np.float32(40_000 / crawl.nstarsExpectedToGenerate)
    lib.gfo_count_crawl(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, C.byref(crawl), None)
    st = gfx.mwc_state(5)
    n = lib.gfo_draw_compat(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(st.copy()),
                            65536, None, 0)
    W = crawl.nstarsExpectedToGenerate
    assert abs(n - W) < 5 * math.sqrt(W)
    t = HostTable(n, catalog_spec())
    cols = t.skygen_cols()
    assert lib.gfo_draw_compat(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), None, None, O.ptr(st), 65536,
                               C.byref(cols), n) == n
    # the pixel of each star: the grid cell holding its projected position
    xy = t.arr("projXY")[:, :n].astype(np.float64)
    PX = np.array([p.X for p in pix])
    PY = np.array([p.Y for p in pix])
    pidx = np.argmin((xy[0][:, None] - PX[None, :]) ** 2 + (xy[1][:, None] - PY[None, :]) ** 2, axis=1)
    ux, uy = _check_star_geometry(t, n, sp, pix, pidx)
    assert stats.kstest(ux, "uniform").pvalue > P_OK
    M = t.arr("absmag")[0][:n].astype(np.float64)
    m_ = M + t.arr("DM")[0][:n].astype(np.float64)
    iM, uM, im, um = _bin_positions(M, m_, sp)
    assert iM.min() >= 0 and iM.max() < sp.nM and im.min() >= 0 and im.max() < sp.nm
    assert stats.kstest(uM, "uniform").pvalue > P_OK and stats.kstest(um, "uniform").pvalue > P_OK


# ---------------------------------------------------------------------------
# clipper
# ---------------------------------------------------------------------------
def test_clip_hides_rows_outside_the_footprint():
    rng = np.random.default_rng(4)
    n = 40_000
    sp = gfx.skygen_params(15.0, 22.0, 0.1, 4.0, 10.0, 1)
    t = HostTable(n, catalog_spec())
    proj = rng.integers(0, 2, n).astype(np.int32)
    r = np.sqrt(rng.uniform(0, 1.99, n))
    a = rng.uniform(0, 2 * math.pi, n)
    x, y = (r * np.cos(a)).astype(np.float32), (r * np.sin(a)).astype(np.float32)
    t.arr("projIdx")[0] = proj
    t.arr("projXY")[0], t.arr("projXY")[1] = x, y
    hid0 = (rng.uniform(size=n) < 0.1).astype(np.int32)
    t.arr("hidden")[0] = hid0
    foot = gfx.footprint(gfx.beam_region(60.0, 30.0, 25.0, hole=5.0),
                         gfx.rect_footprint(200.0, 260.0, -50.0, -10.0).r[0])
    O.lib().gfo_clip(C.byref(sp), C.byref(foot), C.byref(t.skygen_cols()), n)
    hid = t.arr("hidden")[0]
    l = np.zeros(n)
    b = np.zeros(n)
    for h in (0, 1):
        s = proj == h
        l[s], b[s] = lambert_deproject(sp.proj[h].l0, sp.proj[h].b0, x[s].astype(np.float64), y[s].astype(np.float64))
    # outside, from the regions' definitions
    cosd = unit(math.radians(60.0), math.radians(30.0)) @ unit(l, b)
    in_beam = (cosd >= math.cos(math.radians(25.0))) & (cosd < math.cos(math.radians(5.0)))
    ld, bd = np.degrees(l), np.degrees(b)
    in_rect = (ld >= 200.0) & (ld <= 260.0) & (bd >= -50.0) & (bd <= -10.0)
    outside = ~(in_beam | in_rect)
    edges = np.minimum.reduce([np.abs(cosd - math.cos(math.radians(25.0))), np.abs(cosd - math.cos(math.radians(5.0))),
                               np.abs(ld - 200.0) * 1e-2, np.abs(ld - 260.0) * 1e-2, np.abs(bd + 50.0) * 1e-2,
                               np.abs(bd + 10.0) * 1e-2])
    clear = edges > 1e-6
    assert clear.sum() > n - 50
    assert np.array_equal(hid[clear] == 1, (hid0.astype(bool) | outside)[clear])
    assert (outside & ~hid0.astype(bool)).sum() > 1000 and (~outside).sum() > 1000
    # and the oracle's own footprint test at the same positions
    lib = O.lib()
    of = np.array([lib.gfo_outside_footprint(C.byref(foot), float(l[i]), float(b[i])) for i in range(0, n, 7)])
    assert np.array_equal((of == 1)[clear[::7]], outside[::7][clear[::7]])
    # planted mistake: the hole ignored
    in_beam_bad = cosd >= math.cos(math.radians(25.0))
    assert not np.array_equal(hid[clear] == 1, (hid0.astype(bool) | ~(in_beam_bad | in_rect))[clear])


# ---------------------------------------------------------------------------
# trilinear texture sampling
# ---------------------------------------------------------------------------
def test_tex3d_is_clamped_trilinear_interpolation():
    rng = np.random.default_rng(9)
    nx, ny, nz = 7, 5, 4
    data = rng.uniform(-1, 1, (nz, ny, nx)).astype(np.float32)     # x fastest
    t = A.gf_tex3d()
    t.data = data.ctypes.data
    t.nx, t.ny, t.nz = nx, ny, nz
    origin = (-1.2, 0.5, 3.0)
    step = (0.4, 0.25, 2.5)
    grids = []
    for ax in range(3):
        t.tc[ax][0] = origin[ax]
        t.tc[ax][1] = 1.0 / step[ax]
        grids.append(np.float64(np.float32(origin[ax])) + np.arange((nx, ny, nz)[ax]) / np.float64(np.float32(1.0 / step[ax])))
    f = RegularGridInterpolator((grids[2], grids[1], grids[0]), data.astype(np.float64))
    lib = o()
    pts = np.stack([rng.uniform(g[0] - 0.5, g[-1] + 0.5, 3000) for g in grids], axis=1).astype(np.float32)
    got = np.array([lib.gfo_tex3d(C.byref(t), *map(float, p)) for p in pts])
    clamped = np.stack([np.clip(pts[:, a], grids[a][0], grids[a][-1]) for a in range(3)], axis=1)
    want = f(clamped[:, ::-1])
    assert np.abs(got - want).max() < 2e-6
    # planted mistake: y and z axes exchanged in the memory layout
    g = RegularGridInterpolator((grids[2], grids[1], grids[0]),
                                data.reshape(ny, nz, nx).transpose(1, 0, 2).astype(np.float64))
    assert np.abs(got - g(clamped[:, ::-1])).max() > 0.1
