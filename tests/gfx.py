"""Test fixtures: parameter blocks with the reference's defaults, synthetic
star tables, a small sky configuration and seeded stream tables.  Test
infrastructure only."""
import ctypes as C
import math
import os

import numpy as np

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRIMES = np.fromfile(os.path.join(ROOT, "galfast_amd", "data", "safeprimes32.u32"), dtype="<u4")

Rg, z0 = 8000.0, 25.0


def galplane_transform(Rg=Rg, z0=z0):
    """load_transform with center = orientation = galplane (common/transform.cpp:122-135, 199-217, 345-388)."""
    sina = -z0 / Rg
    cosa = math.sqrt(1 - sina * sina)
    M = [[-cosa, 0.0, sina], [0.0, -1.0, 0.0], [sina, 0.0, cosa]]
    # centre of the galplane system in Galactic cartesian: (Rg - x, -y, z) of origin
    ctr = [Rg - (cosa * 0 - sina * 0), -0.0, sina * 0 + cosa * 0]
    T = [-ctr[0], -ctr[1], -ctr[2]]
    return T, M


def set_TM(obj, T, M):
    for i in range(3):
        obj.T[i] = T[i]
        for j in range(3):
            obj.M[i][j] = M[i][j]


def mwc_state(seed, nstreams=65536):
    st = np.zeros(3 * nstreams, dtype=np.uint32)
    n = O.lib().gfo_mwc_seed(seed, O.ptr(PRIMES), len(PRIMES), nstreams, O.ptr(st))
    assert n == nstreams
    return st


def feh_params():
    """os_FeH::construct defaults (modules/FeH_host.cpp:85-123)."""
    p = A.gf_FeH_data()
    A0, A1 = 0.63, 0.37
    s = A0 + A1
    p.A[0], p.A[1] = np.float32(np.float32(A0) / np.float64(np.float32(A0) + np.float32(A1))), \
        np.float32(np.float32(A1) / np.float64(np.float32(A0) + np.float32(A1)))
    p.sigma[0], p.sigma[1], p.sigma[2] = 0.2, 0.2, 0.3
    p.offs[0], p.offs[1], p.offs[2] = 0.0, 0.14, -1.46
    p.Hmu, p.muInf, p.DeltaMu = 500.0, -0.82, 0.55
    p.comp_thin = A.bitmap(0)
    p.comp_thick = A.bitmap(1)
    p.comp_halo = A.bitmap(2)
    set_TM(p, *galplane_transform())
    return p


def _farr(dst, vals):
    for i in range(5):
        dst[i] = vals[i] if i < len(vals) else 0.0


def bond_params(halo_sigma=(141, 75, 85)):
    """os_Bond2010::construct defaults (modules/Bond2010_host.cpp:136-227)."""
    p = A.gf_Bond2010_data()
    p.comp_thin, p.comp_thick, p.comp_halo = A.bitmap(0), A.bitmap(1), A.bitmap(2)
    d = p.disk_1
    _farr(d.m[0], [0, 0, 0, 0, 0])
    _farr(d.m[1], [-194, 19.2, 1.25, 0, 0])
    _farr(d.m[2], [0, 0, 0, 0, 0])
    _farr(d.ss[0], [40, 5, 1.5, 0, 0])
    _farr(d.ss[1], [0] * 5)
    _farr(d.ss[2], [0] * 5)
    _farr(d.ss[3], [12, 1.8, 2, 0, 0])
    _farr(d.ss[4], [0] * 5)
    _farr(d.ss[5], [25, 4, 1.5, 0, 0])
    C.memmove(C.addressof(p.disk_2), C.addressof(p.disk_1), C.sizeof(A.gf_vmoments))
    _farr(p.disk_2.ss[3], [34, 1.2, 2, 0, 0])
    p.disk_2.m[1][0] = np.float32(p.disk_2.m[1][0]) + np.float32(34.0)
    fk = np.float32(3.0)
    p.fk = np.float32(np.float64(fk) / (1.0 + np.float64(fk)))
    h = p.halo
    for i in range(3):
        _farr(h.m[i], [0] * 5)
    _farr(h.ss[0], [halo_sigma[0], 0, 0, 0, 0])
    _farr(h.ss[1], [0] * 5)
    _farr(h.ss[2], [0] * 5)
    _farr(h.ss[3], [halo_sigma[1], 0, 0, 0, 0])
    _farr(h.ss[4], [0] * 5)
    _farr(h.ss[5], [halo_sigma[2], 0, 0, 0, 0])
    set_TM(p, *galplane_transform())
    return p


def vel2pm_params(coordsys=1):
    p = A.gf_vel2pm_data()
    p.coordsys = coordsys
    p.vLSR, p.u0, p.v0, p.w0, p.Rg = -220.0, -10.0, -5.3, 7.2, Rg
    return p


def synthetic_stars(n, seed=1, nbands=5, hidden_frac=0.05, ncomp=3):
    """A host table with plausible skygen columns."""
    rng = np.random.default_rng(seed)
    t = HostTable(n, catalog_spec(nbands=nbands, nabsmag=2))
    l = rng.uniform(0, 360, n)
    b = np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
    D = 10 ** rng.uniform(1.5, 4.5, n)
    lr, br = np.radians(l), np.radians(b)
    t.arr("lb")[0] = l
    t.arr("lb")[1] = b
    xyz = t.arr("XYZ")
    xyz[0] = (D * np.cos(lr) * np.cos(br)).astype(np.float32)
    xyz[1] = (D * np.sin(lr) * np.cos(br)).astype(np.float32)
    xyz[2] = (D * np.sin(br)).astype(np.float32)
    t.arr("DM")[0] = (5 * np.log10(D / 10)).astype(np.float32)
    t.arr("absmag")[0] = rng.uniform(3, 15, n).astype(np.float32)
    t.arr("comp")[0] = rng.integers(0, ncomp, n).astype(np.int32)
    t.arr("hidden")[0] = (rng.uniform(size=n) < hidden_frac).astype(np.int32)
    t.arr("Am")[0] = rng.uniform(0, 0.3, n).astype(np.float32)
    t.arr("FeH")[0] = rng.uniform(-2.0, 0.0, n).astype(np.float32)
    v = t.arr("vcyl")
    v[0] = rng.normal(0, 40, n)
    v[1] = rng.normal(-200, 30, n)
    v[2] = rng.normal(0, 25, n)
    return t


def seeded_lf(seed=42, M0=3.0, M1=15.0, dM=0.1):
    """Smooth seeded luminosity function tabulated on [M0, M1] step dM."""
    rng = np.random.default_rng(seed)
    M = np.round(np.arange(M0, M1 + dM / 2, dM), 10)
    c = rng.uniform(-0.3, 0.3, 4)
    x = (M - 9.0) / 6.0
    phi = 10 ** (-3.46 + 0.25 * (M - 3.0) - 0.02 * (M - 3.0) ** 2 + 0.1 * (c[0] * x + c[1] * x * x + c[2] * x ** 3))
    return M, phi


def lf_texture(M, phi, nsamp=1024):
    tex = np.zeros(nsamp, dtype=np.float32)
    tc = np.zeros(2, dtype=np.float32)
    M = np.ascontiguousarray(M, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    O.lib().gfo_resample_clip(O.ptr(M), O.ptr(phi), len(M), C.c_float(0.0), nsamp, O.ptr(tex), O.ptr(tc))
    return tex, tc


def make_model(kind, comp, f0, lf_tex, lf_tc, **kw):
    m = A.gf_model()
    m.kind = kind
    m.comp = comp
    set_TM(m, *galplane_transform())
    m.l, m.h = kw.get("l", 2600.0), kw.get("h", 300.0)
    m.n, m.ba, m.ca = kw.get("n", -2.77), kw.get("ba", 1.0), kw.get("ca", 0.64)
    m.norm = 1.0
    m.lf.data = lf_tex.ctypes.data
    m.lf.n = len(lf_tex)
    m.lf.tc[0], m.lf.tc[1] = lf_tc[0], lf_tc[1]
    m.f = O.lib().gfo_lf_norm_f(m, C.c_float(f0), 0.0, 0.0, 0.0)
    rmax = kw.get("rmax", 0.0)
    m.rmaxSq = np.float32(rmax) * np.float32(rmax)
    return m


def c0_models(lf_tex, lf_tc):
    return [
        make_model(A.GF_MODEL_EXPDISK, 0, 1.0, lf_tex, lf_tc, l=2600.0, h=300.0),
        make_model(A.GF_MODEL_EXPDISK, 1, 0.12, lf_tex, lf_tc, l=3600.0, h=900.0),
        make_model(A.GF_MODEL_POWERLAW, 2, 0.005, lf_tex, lf_tc, n=-2.77, ca=0.64),
    ]


def lambert(l0, b0):
    p = A.gf_lambert()
    p.l0, p.b0 = l0, b0
    p.cl, p.cb, p.sl, p.sb = math.cos(l0), math.cos(b0), math.sin(l0), math.sin(b0)
    return p


def skygen_params(m0, m1, dm, M0, M1, npix, nthreads=7680):
    """load_skyPixelizationConfig + reset_absmag (skygen/os_skygen.cpp:401-430)."""
    sp = A.gf_skygen_params()
    f = np.float32
    sp.dm = dm
    sp.dM = dm
    nM = int(round((f(M1) - f(M0)) / f(dm)))
    sp.M0 = M0
    sp.M1 = f(f(M1) - 0.5 * np.float64(f(dm)))
    sp.nM = nM
    sp.nm = int(round((f(m1) - f(m0)) / f(dm)))
    sp.m0 = f(f(m0) + 0.5 * np.float64(f(dm)))
    sp.m1 = m1
    sp.npixels = npix
    sp.nthreads = nthreads
    sp.proj[0] = lambert(math.radians(90), math.radians(90))
    sp.proj[1] = lambert(math.radians(270), math.radians(-90))
    return sp


def pixelize_rect(l0, l1, b0, b1, dx_deg, sub=8):
    """Square Lambert pixels (north projection) covering a Galactic rectangle,
    covered fraction by sub x sub point sampling."""
    proj = lambert(math.radians(90), math.radians(90))
    dx = math.radians(dx_deg)
    # bounding box of the rectangle in projected coordinates
    ls = np.radians(np.linspace(l0, l1, 200))
    bs = np.radians(np.linspace(b0, b1, 200))
    L, B = np.meshgrid(ls, bs)
    kp = np.sqrt(2.0 / (1.0 + np.sin(B)))
    X = kp * np.cos(B) * np.sin(L - proj.l0)
    Y = -kp * np.cos(B) * np.cos(L - proj.l0)
    x0, y0 = X.min(), Y.min()
    nx = int(math.ceil((X.max() - x0) / dx))
    ny = int(math.ceil((Y.max() - y0) / dx))
    pix = []
    for i in range(nx):
        for j in range(ny):
            xs = x0 + dx * (i + (np.arange(sub) + 0.5) / sub)
            ys = y0 + dx * (j + (np.arange(sub) + 0.5) / sub)
            XX, YY = np.meshgrid(xs, ys)
            r2 = XX ** 2 + YY ** 2
            bb = np.arcsin(1 - 0.5 * r2)
            ll = (proj.l0 + np.arctan2(XX, -YY)) % (2 * np.pi)
            inside = (np.degrees(bb) >= b0) & (np.degrees(bb) <= b1) & \
                     (np.degrees(ll) >= l0) & (np.degrees(ll) <= l1)
            frac = inside.mean()
            if frac == 0:
                continue
            xc, yc = x0 + dx * (i + 0.5), y0 + dx * (j + 0.5)
            r2c = xc * xc + yc * yc
            bc = math.asin(1 - 0.5 * r2c)
            lc = proj.l0 + math.atan2(xc, -yc)
            pb = A.gf_pencil_beam()
            pb.cl, pb.cb, pb.sl, pb.sb = math.cos(lc), math.cos(bc), math.sin(lc), math.sin(bc)
            pb.coveredFraction = frac
            pb.dx = dx
            pb.dA = np.float32(dx) * np.float32(dx)
            pb.projIdx = 0
            pb.X, pb.Y = xc, yc
            pb.extIdx = -1
            pix.append(pb)
    arr = (A.gf_pencil_beam * len(pix))(*pix)
    return arr


def rect_footprint(l0, l1, b0, b1, bmin=0.0, equ=0):
    f = A.gf_footprint()
    f.nregion = 1
    r = f.r[0]
    r.kind, r.equ = 1, equ
    r.l0, r.l1, r.b0, r.b1, r.bmin = l0, l1, b0, b1, bmin
    return f


def beam_region(l, b, radius, hole=0.0, bmin=0.0):
    """A beam region (Galactic centre and radii in degrees)."""
    r = A.gf_foot_region()
    lr, br = np.radians(l), np.radians(b)
    r.kind = 2
    r.cl, r.cb, r.sl, r.sb = np.cos(lr), np.cos(br), np.sin(lr), np.sin(br)
    r.cosr = np.cos(np.radians(radius))
    r.hole = 1.0 - np.cos(np.radians(hole)) if hole > 0 else 0.0
    r.bmin = bmin
    return r


def footprint(*regions):
    f = A.gf_footprint()
    f.nregion = len(regions)
    for k, r in enumerate(regions):
        f.r[k] = r
    return f


def synthetic_photosys(seed=42, ncolors=4, Mr=(3.0, 15.0, 0.1), FeH=(-2.5, 0.0, 0.25)):
    """Rows (Mr, FeH, colours...) with colours as seeded cubic polynomials of
    an r-i like variable, shifted linearly with FeH."""
    rng = np.random.default_rng(seed)
    Ms = np.round(np.arange(Mr[0], Mr[1] + Mr[2] / 2, Mr[2]), 8)
    Fs = np.round(np.arange(FeH[0], FeH[1] + FeH[2] / 2, FeH[2]), 8)
    coef = rng.uniform(-0.05, 0.05, (ncolors, 4))
    coef[:, 1] += rng.uniform(0.5, 1.5, ncolors)
    slope = rng.uniform(0.02, 0.2, ncolors)
    rows = []
    for f in Fs:
        for m in Ms:
            ri = 0.1 + 0.12 * (m - 3.0)
            cols = [coef[c, 0] + coef[c, 1] * ri + coef[c, 2] * ri ** 2 + coef[c, 3] * ri ** 3 + slope[c] * f
                    for c in range(ncolors)]
            rows.append([m, f] + cols)
    return np.array(rows, dtype=np.float64)


def photometry_tables(rows, ncolors, dFeH=0.01, dMr=0.01):
    cap = 4 * 4 * 2000 * 3000
    colors = np.zeros(cap, dtype=np.float32)
    flags = np.zeros(cap, dtype=np.float32)
    nF, nM = C.c_int(), C.c_int()
    tc = np.zeros(4, dtype=np.float32)
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    ntex = O.lib().gfo_photometry_tables(O.ptr(rows), rows.shape[0], ncolors, dFeH, dMr, O.ptr(colors),
                                         O.ptr(flags), cap, C.byref(nF), C.byref(nM), O.ptr(tc))
    assert ntex > 0
    sz = nF.value * nM.value * 4
    return ntex, nF.value, nM.value, tc, colors[: ntex * sz].copy(), flags[: ntex * sz].copy()


def photometry_params(ntex, nF, nM, tc, color_ptrs, flag_ptrs, ncolors, bidx, reddening):
    p = A.gf_photometry_data()
    p.ncolors, p.bidx = ncolors, bidx
    for i in range(17):
        p.reddening[i] = reddening[i] if i < len(reddening) else 1.0
    p.nFeH, p.nMr = nF, nM
    p.tcFeH[0], p.tcFeH[1], p.tcMr[0], p.tcMr[1] = tc[0], tc[1], tc[2], tc[3]
    for i in range(4):
        p.color[i] = color_ptrs[i] if i < ntex else None
        p.cflags[i] = flag_ptrs[i] if i < ntex else None
    return p


def multiples_tables(lf_M, lf_phi, fraction=0.5, npix=256):
    """secProb, cumLF, invCumLF as built by os_unresolvedMultiples::construct
    (modules/unresolvedMultiples_host.cpp:140-180)."""
    x = np.ascontiguousarray(lf_M, dtype=np.float64)
    y = np.ascontiguousarray(lf_phi, dtype=np.float64)
    sec = np.full(2, fraction, dtype=np.float32)
    sec_tc = np.array([-100.0, 1.0 / 200.0], dtype=np.float32)
    dx = (x[-1] - x[0]) / (npix - 1)
    xc = np.zeros(npix)
    yc = np.zeros(npix)
    xc[0] = x[0]
    yprev = O.lib().gfo_spline(O.ptr(x), O.ptr(y), len(x), x[0])
    for i in range(1, npix):
        xx = x[0] + i * dx
        yy = O.lib().gfo_spline(O.ptr(x), O.ptr(y), len(x), xx)
        dy = yy - yprev
        dA = (yprev + 0.5 * dy) * dx
        xc[i] = xx
        yc[i] = yc[i - 1] + dA
        yprev = yy
    yc /= yc[-1]
    cum = np.zeros(npix, dtype=np.float32)
    cum_tc = np.zeros(2, dtype=np.float32)
    inv = np.zeros(npix, dtype=np.float32)
    inv_tc = np.zeros(2, dtype=np.float32)
    O.lib().gfo_resample(O.ptr(xc), O.ptr(yc), npix, npix, O.ptr(cum), O.ptr(cum_tc))
    O.lib().gfo_resample(O.ptr(yc), O.ptr(xc), npix, npix, O.ptr(inv), O.ptr(inv_tc))
    return (sec, sec_tc), (cum, cum_tc), (inv, inv_tc)


def tex1d(data_ptr, n, tc):
    t = A.gf_tex1d()
    t.data = data_ptr
    t.n = n
    t.tc[0], t.tc[1] = tc[0], tc[1]
    return t


def photoerr_knots(nbands, m0=10.0, m1=30.0, dm=0.01):
    """sigma(m) = 0.02 + exp(0.8 (m - 24)) tabulated for every band."""
    m = np.round(np.arange(m0, m1 + dm / 2, dm), 8)
    s = 0.02 + np.exp(0.8 * (m - 24.0))
    kx = np.tile(m, nbands)
    ky = np.tile(s, nbands)
    koff = [i * len(m) for i in range(nbands + 1)]
    return kx, ky, koff
