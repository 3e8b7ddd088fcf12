"""Writers for the benchmark workloads as galfast configuration files.

Each workload is written as the file set a reference user would pass to
``galfast catalog cmd.conf``: definitions, skygen input, density models,
footprint, module configurations, plus the synthetic tables they name
(a seeded smooth luminosity function, a seeded polynomial photometric
system and tabulated photometric-error curves).

* ``c0``  -- thin disk (2600/300 pc) + thick disk (3600/900 pc, 0.12) +
  power-law halo (n=-2.77, c/a=0.64, 0.005); Sun at Rg=8000, z0=25;
  10x10 deg patch at (l,b)=(90,60) in 1 deg pixels; r in [15,21], dm=0.05;
  Mr in [3,15]; 5 synthetic bands; Ivezic [Fe/H] (the photometry module
  needs FeH).  About 1e5 stars.
* ``c1``  -- the same model over |b| > 10 deg in 0.5 deg pixels, r in
  [14,24]; Ivezic [Fe/H]; Bond2010 kinematics with halo dispersions
  (sigma_r, sigma_phi, sigma_theta) = (140, 105, 75) km/s; equatorial proper
  motions and coordinates; photometric errors sigma(m)=0.02+exp(0.8(m-24)).
"""
import os

import numpy as np

BANDS = ["SYNu", "SYNg", "SYNr", "SYNi", "SYNz"]


def seeded_lf(seed=42, M0=3.0, M1=15.0, dM=0.1):
    """Smooth luminosity function on [M0, M1] step dM (stars / pc^3 / mag)."""
    rng = np.random.default_rng(seed)
    M = np.round(np.arange(M0, M1 + dM / 2, dM), 10)
    c = rng.uniform(-0.3, 0.3, 4)
    x = (M - 9.0) / 6.0
    logphi = -3.46 + 0.25 * (M - 3.0) - 0.02 * (M - 3.0) ** 2 + 0.1 * (c[0] * x + c[1] * x * x + c[2] * x ** 3)
    return M, 10 ** logphi


def seeded_photosys(seed=42, ncolors=4, Mr=(-1.0, 18.0, 0.1), FeH=(-3.0, 0.5, 0.25)):
    """Rows (Mr, FeH, u-g, g-r, r-i, i-z): colours are seeded cubic polynomials
    of an r-i proxy that grows with Mr, plus a linear metallicity term."""
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
            rows.append([m, f] + [coef[c, 0] + coef[c, 1] * ri + coef[c, 2] * ri ** 2 + coef[c, 3] * ri ** 3
                                  + slope[c] * f for c in range(ncolors)])
    return np.array(rows)


def _w(path, text):
    with open(path, "w") as f:
        f.write(text.strip() + "\n")


def write_common(d, seed=42):
    os.makedirs(d, exist_ok=True)
    M, phi = seeded_lf(seed)
    np.savetxt(os.path.join(d, "lf.txt"), np.c_[M, phi], fmt="%.10g")
    np.savetxt(os.path.join(d, "photosys.txt"), seeded_photosys(seed), fmt="%.8f")
    m = np.round(np.arange(10.0, 30.0 + 0.005, 0.01), 6)
    err = 0.02 + np.exp(0.8 * (m - 24.0))
    for b in BANDS:
        np.savetxt(os.path.join(d, "SYN.%s.photoerr.txt" % b), np.c_[m, err], fmt="%.8g")
    _w(os.path.join(d, "definitions.conf"), "module = definitions\nRg = 8000\nz0 = 25")
    _w(os.path.join(d, "model.thin.conf"), """
module = model
model = exponential disk
comp = 0
l = 2600
h = 300
lumfunc = lf.txt
f = 1""")
    _w(os.path.join(d, "model.thick.conf"), """
module = model
model = exponential disk
comp = 1
l = 3600
h = 900
lumfunc = lf.txt
f = 0.12""")
    _w(os.path.join(d, "model.halo.conf"), """
module = model
model = power law ellipsoid
comp = 2
n = -2.77
c_a = 0.64
lumfunc = lf.txt
f = 0.005""")
    _w(os.path.join(d, "feh.conf"), """
module = FeH
comp_thin = 0
comp_thick = 1
comp_halo = 2
A0 = 0.63
sigma0 = 0.2
offs0 = 0.0
A1 = 0.37
sigma1 = 0.2
offs1 = 0.14
muInf = -0.82
deltaMu = 0.55
Hmu = 500
offsH = -1.46
sigmaH = 0.3""")
    _w(os.path.join(d, "photometry.conf"), """
module = photometry
comp = 0-2
bandset = SYN
file = photosys.txt
bands = %s
bootstrap_band = SYNr
absband = absSYNr
dFeH = 0.01
dM = 0.01
reddening = 1.8739 1.3788 1 0.7583 0.5376""" % " ".join(BANDS))


def write_c0(d, seed=42):
    write_common(d, seed)
    _w(os.path.join(d, "skygen.conf"), """
module = skygen
band = SYNr
dm = 0.05
m0 = 15
m1 = 21
M0 = 3
M1 = 15
dx = 1""")
    _w(os.path.join(d, "foot.conf"), "module = footprint\ntype = rect\ncoordsys = gal\nrect = 85 95 55 65")
    _w(os.path.join(d, "cmd.conf"), """
module = config
seed = %d
definitions = definitions.conf
input = skygen.conf
models = model.thin.conf model.thick.conf model.halo.conf
footprints = foot.conf
modules = feh.conf photometry.conf""" % seed)
    return os.path.join(d, "cmd.conf")


def write_c1(d, seed=42):
    write_common(d, seed)
    _w(os.path.join(d, "skygen.conf"), """
module = skygen
band = SYNr
dm = 0.05
m0 = 14
m1 = 24
M0 = 3
M1 = 15
dx = 0.5
maxstars = 0""")
    _w(os.path.join(d, "foot.conf"), "module = footprint\ntype = rect\ncoordsys = gal\nrect = 0 360 -90 90\nbmin = 10")
    _w(os.path.join(d, "Bond2010.conf"), """
module = Bond2010
comp_thin = 0
comp_thick = 1
comp_halo = 2
HsigmaR = 140 0 0 0 0
HsigmaPhi = 105 0 0 0 0
HsigmaTh = 75 0 0 0 0""")
    _w(os.path.join(d, "vel2pm.conf"), "module = vel2pm\ncoordsys = equ")
    _w(os.path.join(d, "gal2other.conf"), "module = gal2other\ncoordsys = equ")
    _w(os.path.join(d, "photoerr.conf"), "module = photometricErrors\n" +
       "\n".join("SYN.%s.file = SYN.%s.photoerr.txt" % (b, b) for b in BANDS))
    _w(os.path.join(d, "cmd.conf"), """
module = config
seed = %d
definitions = definitions.conf
input = skygen.conf
models = model.thin.conf model.thick.conf model.halo.conf
footprints = foot.conf
modules = feh.conf Bond2010.conf photometry.conf photoerr.conf vel2pm.conf gal2other.conf""" % seed)
    return os.path.join(d, "cmd.conf")


def set_footprints(cmd, bodies):
    """Replace the footprints of the job in `cmd` (a cmd.conf path) by one
    footprint file per entry of `bodies` (the file text after
    "module = footprint"); the job's footprint is their union."""
    d = os.path.dirname(cmd)
    names = []
    for k, body in enumerate(bodies):
        names.append("foot%d.conf" % k)
        _w(os.path.join(d, names[-1]), "module = footprint\n" + body)
    with open(cmd) as f:
        lines = [ln for ln in f.read().splitlines() if not ln.strip().startswith("footprints")]
    lines.append("footprints = " + " ".join(names))
    _w(cmd, "\n".join(lines))
    return cmd


def add_countsmap(cmd, keys):
    """Give the job in `cmd` (a cmd.conf path) a countsMap output stage with
    the given {key: value} (xy, z, x0, dx, equalarea, "coadd.<name>",
    "output all bins", ...)."""
    d = os.path.dirname(cmd)
    _w(os.path.join(d, "countsMap.conf"), "module = counts map\n" +
       "\n".join("%s = %s" % (k, v) for k, v in keys.items()))
    with open(cmd, "a") as f:
        f.write("output = countsMap.conf\n")
    return cmd


def write_c3(d, seed=42, m1=25.0, beam=(90.0, 60.0, 1.0), dx=0.25, table_seed=42):
    """Serving base configuration: the config-1 models, tables (seeded with
    table_seed) and modules, a beam footprint, r in [14, m1], dx-degree
    pixels and pipeline seed `seed`.  The server takes each request's own
    beam, faint limit and seed; a catalog opened on this configuration with
    a request's values is that request alone."""
    path = write_c1(d, table_seed)
    with open(path) as f:
        text = f.read()
    _w(path, text.replace("seed = %d\n" % table_seed, "seed = %d\n" % seed))
    _w(os.path.join(d, "skygen.conf"), """
module = skygen
band = SYNr
dm = 0.05
m0 = 14
m1 = %r
M0 = 3
M1 = 15
dx = %r""" % (float(m1), float(dx)))
    _w(os.path.join(d, "foot.conf"), "module = footprint\ntype = beam\ncoordsys = gal\nfootprint_beam = %r %r %r" %
       tuple(float(v) for v in beam))
    return os.path.join(d, "cmd.conf")


def request_seed(rid):
    """32-bit seed of request rid (SplitMix64 finaliser of the id)."""
    z = (rid + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return int((z ^ (z >> 31)) & 0xFFFFFFFF)


def c3_requests(n, seed=0, area=(0.01, 100.0), bmin=5.0, m1=(20.0, 25.0), id0=1):
    """n serving requests (id, l, b, radius, m1, seed): circular beams with
    area log-uniform in `area` (deg^2), centres uniform on the sphere with
    |b| > bmin, faint limit uniform in m1, seed = request_seed(id)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        z = rng.uniform(-1.0, 1.0)
        lon = rng.uniform(0.0, 360.0)
        b = np.degrees(np.arcsin(z))
        a = np.exp(rng.uniform(np.log(area[0]), np.log(area[1])))
        mag = rng.uniform(*m1)
        if abs(b) <= bmin:
            continue
        # cap of solid angle a: 2 pi (1 - cos r)
        r = np.degrees(np.arccos(1.0 - a * (np.pi / 180.0) ** 2 / (2.0 * np.pi)))
        rid = id0 + len(out)
        out.append((rid, float(lon), float(b), float(r), float(mag), request_seed(rid)))
    return out


BANDS6 = BANDS + ["SYNy"]


def lambert_deproject_grid(x, y, hemi):
    """(l, b) in radians of Lambert-plane points of the north (hemi 0, pole
    l=90, b=90) or south (hemi 1, pole l=270, b=-90) projection."""
    r2 = x * x + y * y
    sb = np.clip(1.0 - 0.5 * r2, -1.0, 1.0)
    if hemi == 0:
        return np.pi / 2 + np.arctan2(x, -y), np.arcsin(sb)
    return 1.5 * np.pi + np.arctan2(x, y), -np.arcsin(sb)


def dust_disk_cube(hemi, n=256, ndm=61, dm_max=30.0, a0=0.5e-3, h=100.0):
    """r-band extinction of an exponential dust disk (a0 mag/pc in the
    plane, scale height h pc, the Sun in the plane) integrated along each
    line of sight: A(D, b) = a0 h / |sin b| (1 - exp(-D |sin b| / h)).
    Axes X, Y in [-2, 2] (Lambert plane of hemisphere `hemi`) and DM in
    [0, dm_max]; zero outside the hemisphere (r^2 > 2).  Returns the cube as
    [DM][Y][X] and its (CRVAL, CDELT) per axis."""
    xs = np.linspace(-2.0, 2.0, n)
    X, Y = np.meshgrid(xs, xs)                       # [y][x]
    _, b = lambert_deproject_grid(X, Y, hemi)
    sb = np.abs(np.sin(b))
    out = np.zeros((ndm, n, n), dtype=np.float64)
    for k in range(ndm):
        D = 10 ** (0.2 * (dm_max * k / (ndm - 1)) + 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            A = np.where(sb > 1e-9, a0 * h / sb * (1.0 - np.exp(-D * sb / h)), a0 * D)
        out[k] = A
    out[:, (X * X + Y * Y) > 2.0] = 0.0
    axes = [(-2.0, 4.0 / (n - 1)), (-2.0, 4.0 / (n - 1)), (0.0, dm_max / (ndm - 1))]
    return out.astype(np.float32), axes


def write_fits_cube(path, cube, axes, extra=None):
    """A FITS primary image of 32-bit floats (cube indexed [z][y][x]) with
    CRPIXn = 1, CRVALn, CDELTn coordinates."""
    nz, ny, nx = cube.shape
    cards = [("SIMPLE", "T"), ("BITPIX", "-32"), ("NAXIS", "3"), ("NAXIS1", str(nx)), ("NAXIS2", str(ny)),
             ("NAXIS3", str(nz))]
    for i, (crval, cdelt) in enumerate(axes):
        cards += [("CRPIX%d" % (i + 1), "1"), ("CRVAL%d" % (i + 1), repr(float(crval))),
                  ("CDELT%d" % (i + 1), repr(float(cdelt)))]
    for k, v in (extra or {}).items():
        cards.append((k, repr(float(v))))
    hdr = "".join(("%-8s= %20s" % (k, v)).ljust(80) for k, v in cards) + "END".ljust(80)
    hdr = hdr.ljust((len(hdr) + 2879) // 2880 * 2880)
    data = np.ascontiguousarray(cube, dtype=">f4").tobytes()
    data += b"\0" * ((-len(data)) % 2880)
    with open(path, "wb") as f:
        f.write(hdr.encode("ascii"))
        f.write(data)


def write_c4(d, seed=42, small=False, foot=None, m1=None):
    """Config 4: 6 synthetic bands, unresolved binaries (fraction 0.5,
    companion magnitudes from the LF, fainter than the primary) and an
    analytic exponential dust disk (h = 100 pc, 0.5 mag/kpc in r) as north
    and south (X, Y, DM) extinction cubes, r in [14, 27.5] over
    l in [60, 100], b in [20, 50] (~970 deg^2) in 0.25 deg pixels.  small:
    the config-0 sky (10x10 deg, 1 deg pixels, r 15-21) moved to b in
    [10, 20] with 64x64x31 cubes, for tests.  foot: the footprint file's
    text after "module = footprint" instead of the rect; m1: the faint
    limit."""
    write_common(d, seed)
    np.savetxt(os.path.join(d, "photosys.txt"), seeded_photosys(seed, ncolors=5), fmt="%.8f")
    m = np.round(np.arange(10.0, 30.0 + 0.005, 0.01), 6)
    err = 0.02 + np.exp(0.8 * (m - 24.0))
    np.savetxt(os.path.join(d, "SYN.SYNy.photoerr.txt"), np.c_[m, err], fmt="%.8g")
    _w(os.path.join(d, "photometry.conf"), """
module = photometry
comp = 0-2
bandset = SYN
file = photosys.txt
bands = %s
bootstrap_band = SYNr
absband = absSYNr
dFeH = 0.01
dM = 0.01
reddening = 1.8739 1.3788 1 0.7583 0.5376 0.4""" % " ".join(BANDS6))
    _w(os.path.join(d, "multiples.conf"), """
module = unresolvedMultiples
comp = 0-2
algorithm = LF_M2_gt_M1
fraction = 0.5
lumfunc = lf.txt""")
    _w(os.path.join(d, "photoerr.conf"), "module = photometricErrors\n" +
       "\n".join("SYN.%s.file = SYN.%s.photoerr.txt" % (b, b) for b in BANDS6))
    n, ndm = (64, 31) if small else (256, 61)
    for hemi, name, pole in ((0, "north", (90.0, 90.0)), (1, "south", (-90.0, -90.0))):
        cube, axes = dust_disk_cube(hemi, n=n, ndm=ndm)
        write_fits_cube(os.path.join(d, "ext.%s.fits" % name), cube, axes, {"LAMBDA0": pole[0], "PHI1": pole[1]})
    _w(os.path.join(d, "ext.conf"), "north = ext.north.fits\nsouth = ext.south.fits")
    if small:
        sky, body = ("m0 = 15\nm1 = %r\ndx = 1" % (21.0 if m1 is None else float(m1)),
                     "type = rect\ncoordsys = gal\nrect = 85 95 10 20")
    else:
        sky, body = ("m0 = 14\nm1 = %r\ndx = 0.25" % (27.5 if m1 is None else float(m1)),
                     "type = rect\ncoordsys = gal\nrect = 60 100 20 50")
    _w(os.path.join(d, "skygen.conf"), "module = skygen\nband = SYNr\ndm = 0.05\nM0 = 3\nM1 = 15\n%s\n"
       "extmaps = ext.conf\nmaxstars = 0" % sky)
    _w(os.path.join(d, "foot.conf"), "module = footprint\n%s" % (body if foot is None else foot))
    _w(os.path.join(d, "cmd.conf"), """
module = config
seed = %d
definitions = definitions.conf
input = skygen.conf
models = model.thin.conf model.thick.conf model.halo.conf
footprints = foot.conf
modules = feh.conf multiples.conf photometry.conf photoerr.conf""" % seed)
    return os.path.join(d, "cmd.conf")


def write_c4_small(d, seed=42):
    return write_c4(d, seed, small=True)


def foot_body(l, b, radius, l1=0.0, b1=0.0):
    """Footprint file text of a server request: a beam, or with radius 0 the
    Galactic rect [l, l1] x [b, b1]."""
    if radius > 0:
        return "type = beam\ncoordsys = gal\nfootprint_beam = %r %r %r" % (float(l), float(b), float(radius))
    return "type = rect\ncoordsys = gal\nrect = %r %r %r %r" % (float(l), float(l1), float(b), float(b1))


def c4_tiles(nl=20, nb=10, l=(60.0, 100.0), b=(20.0, 50.0), m1=27.5, id0=1):
    """The config-4 footprint as nl x nb Galactic rect sub-requests
    (id, l0, b0, 0, m1, seed, l1, b1) that tile it exactly, seed per id."""
    out = []
    dl, db = (l[1] - l[0]) / nl, (b[1] - b[0]) / nb
    for i in range(nl):
        for j in range(nb):
            rid = id0 + len(out)
            out.append((rid, l[0] + i * dl, b[0] + j * db, 0.0, float(m1), request_seed(rid),
                        l[0] + (i + 1) * dl, b[0] + (j + 1) * db))
    return out


WRITERS = {"c0": write_c0, "c1": write_c1, "c3": write_c3, "c4": write_c4, "c4s": write_c4_small}
