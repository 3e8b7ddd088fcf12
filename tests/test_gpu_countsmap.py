"""countsMap on the device: gf_countsmap_bin / gf_countsmap_finish against the
oracle's binning (bit-exact counts, beams and totals over two chunks), and
the countsMap writer on config-0 and config-1-shaped catalogs: the text table
parsed back and compared with the oracle's map of the same catalog's visible
rows, every line's coordinates, z, area or (lon, lat) as the reference's
destructor computes and prints them (os_countsMap::~os_countsMap,
pipeline.cpp:684-762; printed with glibc's "%g")."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import _abi as A, configs
from galfast_amd.catalog import Catalog
from galfast_amd.table import HostTable, DeviceTable, catalog_spec
from countsmap_util import D2R, host_cols, oracle_bin, params, shape
from devutil import stream_ptr
from test_gpu_catalog import collect

pytestmark = pytest.mark.gpu

LIBC = C.CDLL("libc.so.6")
R2D = 57.29577951308232


def g(v):
    buf = C.create_string_buffer(48)
    LIBC.snprintf(buf, 48, b"%g", C.c_double(float(v)))
    return buf.value.decode()


def device_bin(p, tabs):
    """Bin device tables [(DeviceTable, n), ...] into one map; returns the
    host copies of (counts, beams, total) after gf_countsmap_finish."""
    import torch
    nb, nz, zw = shape(p)
    dp = A.gf_countsmap_params.from_buffer_copy(p)
    co = torch.from_numpy(p.coadd.copy()).cuda() if p.coadd.size else None
    dp.coadd_offsets = co.data_ptr() if co is not None else None
    counts = torch.zeros(nb * nz * zw, dtype=torch.int32, device="cuda")
    beams = torch.zeros(nb, dtype=torch.uint8, device="cuda")
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    L = G.lib()
    for d, n in tabs:
        print("gf_countsmap_bin n=%d equalarea=%d ncoadd=%d" % (n, p.equalarea, p.ncoadd), flush=True)
        G.check(L.gf_countsmap_bin(C.byref(dp), d.col("lb"), d.col("mags"), d.col("hidden"), n, counts.data_ptr(),
                                   beams.data_ptr(), total.data_ptr(), stream_ptr()), "gf_countsmap_bin")
    G.check(L.gf_countsmap_finish(C.byref(dp), counts.data_ptr(), stream_ptr()), "gf_countsmap_finish")
    torch.cuda.synchronize()
    return counts.cpu().numpy().view(np.uint32), beams.cpu().numpy(), int(total.item())


@pytest.mark.parametrize("equalarea", [0, 1])
def test_binning_matches_oracle(equalarea):
    rng = np.random.default_rng(21 + equalarea)
    spec = catalog_spec()
    offs = np.array([[0.3, 0.0, -0.2, 0.5, 1.0], [0.0, 0.0, 0.0, 0.0, 0.0]], np.float32)
    if equalarea:
        p = params(nx=231, ny=231, nz=41, x0=-2.0, dx=np.float32(D2R), y0=-2.0, dy=np.float32(D2R), z0=0.0,
                   dz=1.0, zmag_width=5, coadd=offs, equalarea=True)
    else:
        p = params(nx=360, ny=180, nz=41, x0=0.5, dx=1.0, y0=-89.5, dy=1.0, z0=0.0, dz=1.0, zmag_width=5,
                   coadd=offs)
    tabs, hosts = [], []
    for n in (200_000, 77_777):
        h = HostTable(n, spec)
        h.arr("lb")[0] = rng.uniform(0, 360, n)
        h.arr("lb")[1] = np.degrees(np.arcsin(rng.uniform(-1, 1, n)))
        h.arr("lb")[1][:50] = [90.0, -90.0, 0.0, -0.0, 45.0] * 10       # poles, the equator, both signs of zero
        h.arr("lb")[0][:50:7] = [0.0, 90.0, 180.0, 270.0, 360.0, -90.0, 450.0, 0.0][:len(range(0, 50, 7))]
        h.arr("mags")[:] = rng.uniform(10, 45, (5, n)).astype(np.float32)
        h.arr("mags")[2, 60:70] = np.nan
        h.arr("mags")[3, 70:80] = -1e30
        h.arr("hidden")[0] = rng.random(n) < 0.15
        d = DeviceTable(n, spec)
        d.upload(h)
        tabs.append((d, n))
        hosts.append(h)
    counts, beams, total = device_bin(p, tabs)
    oc = ob = None
    ot = 0
    for h in hosts:
        xy = np.ascontiguousarray(h.arr("lb"))
        z = np.ascontiguousarray(h.arr("mags"))
        hid = np.ascontiguousarray(h.arr("hidden")[0])
        oc, ob, t = oracle_bin(p, *host_cols(xy, z, hid), h.n, oc, ob)
        ot += t
    assert total == ot == 5 * sum(int((h.arr("hidden")[0] == 0).sum()) for h in hosts)
    assert np.array_equal(beams, ob)
    assert np.array_equal(counts, oc)


def parse_map(path):
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[-1] == ""
    head = [l for l in lines[:-1] if l.startswith("#")]
    body = [l.split("\t") for l in lines[:-1] if not l.startswith("#")]
    return head, body


def visible_rows(cmd, kw, nstars=None):
    cat = Catalog(cmd, seed=42, **kw)
    if nstars:
        cat.set("nstars", nstars)
    cat.count()
    info = cat.info()
    spec = catalog_spec(nbands=info.nbands, nabsmag=max(1, info.nabsmag))
    got = collect(cat, spec)
    cat.close()
    return got


@pytest.mark.parametrize("cfg,keys", [
    ("c0", {"xy": "lb", "z": "SYN", "x0": 84.5, "x1": 95.5, "dx": 0.5, "y0": 54.5, "y1": 65.5, "dy": 0.5,
            "z0": 12, "z1": 24, "dz": 0.5, "coadd.deep": "0.5 0 0 0 0.5", "coadd.r": "9 0 9 9 9"}),
    ("c0", {"xy": "lb", "z": "SYN", "equalarea": 1, "dx": 0.005, "dy": 0.005, "x0": -0.1, "x1": 0.1,
            "y0": -0.65, "y1": -0.35, "z0": 12, "z1": 24, "dz": 0.5, "output all bins": 1}),
    ("c1", {"xy": "radec", "z": "obsSYN", "x0": 1, "x1": 359, "dx": 2, "y0": -89, "y1": 89, "dy": 2,
            "coadd.any": "0 0 0 0 0"}),
    ("c1", {"xy": "lb", "z": "obsSYN", "equalarea": 1, "z0": 10, "z1": 30, "dz": 2}),
])
def test_countsmap_writer(tmp_path, cfg, keys):
    d = str(tmp_path)
    if cfg == "c0":
        cmd, kw, nstars = configs.write_c0(d), dict(chunk_rows=40000), None
        band_col = "mags"
    else:
        # the full-sky job scaled to 3e5 stars (the reference's nstars key)
        cmd, kw, nstars = configs.write_c1(d), dict(chunk_rows=30_000_000), 300_000
        band_col = "obsMags"
    configs.add_countsmap(cmd, keys)
    got = visible_rows(cmd, kw, nstars)
    cat = Catalog(cmd, seed=42, **kw)
    if nstars:
        cat.set("nstars", nstars)
    cat.count()
    path = os.path.join(d, "counts.txt")
    print("writing countsMap %s %s" % (cfg, sorted(keys)), flush=True)
    st, rows = cat.write(path, "countsmap")
    cat.close()
    assert rows == st.visible > 1000

    # the map as the host builds its parameters (os_countsMap::construct);
    # a key missing from the stage's file is looked up in the job's
    # definitions, as every configuration lookup is (definitions.conf
    # defines z0 = 25, the Sun's height)
    defs = {}
    for line in open(os.path.join(d, "definitions.conf")):
        k, _, v = line.partition("=")
        if v and k.strip() != "module":
            defs[k.strip()] = v.strip()
    keys = dict(defs, **keys)
    ea = bool(keys.get("equalarea", 0))
    f32 = np.float32
    dflt = (-2.0, 2.0, f32(D2R)) if ea else None
    x0 = f32(keys.get("x0", dflt[0] if ea else 0.0))
    x1 = float(keys.get("x1", dflt[1] if ea else 360.0))
    dx = f32(keys.get("dx", dflt[2] if ea else 1.0))
    y0 = f32(keys.get("y0", dflt[0] if ea else -90.0))
    y1 = float(keys.get("y1", dflt[1] if ea else 90.0))
    dy = f32(keys.get("dy", dflt[2] if ea else 1.0))
    z0, z1, dz = f32(keys.get("z0", 0.0)), float(keys.get("z1", 40.0)), f32(keys.get("dz", 1.0))

    def nbins(a0, a1, da):
        t = f32((a1 - float(a0)) / float(da))
        r = round(float(t))
        return int(r if abs(float(t) - r) < 1e-4 else math.ceil(float(t))) + 1
    nx, ny, nz = nbins(x0, x1, dx), nbins(y0, y1, dy), nbins(z0, z1, dz)
    names = sorted(k for k in keys if k.startswith("coadd."))
    offs = np.array([[float(v) for v in keys[k].split()] for k in names], np.float32).reshape(len(names), 5)
    p = params(nx=nx, ny=ny, nz=nz, x0=float(x0), dx=float(dx), y0=float(y0), dy=float(dy), z0=float(z0),
               dz=float(dz), zmag_width=5, coadd=offs, equalarea=ea)
    xy_col = {"lb": "lb", "radec": "radec"}[keys["xy"]]
    oc = ob = None
    for t, n, _ in got:
        xy = np.ascontiguousarray(t.arr(xy_col)[:, :n])
        z = np.ascontiguousarray(t.arr(band_col)[:, :n])
        hid = np.ascontiguousarray(t.arr("hidden")[0, :n])
        oc, ob, _ = oracle_bin(p, *host_cols(xy, z, hid), n, oc, ob)
    nmaps = 2 if ea else 1
    nb, _, zw = shape(p)
    oc = oc.reshape(nb, nz, zw)

    head, body = parse_map(path)
    if ea:
        assert head[0] == "# Binned in lambert equal area projection, using the following poles:"
        assert head[1] == "#\tmap = 0   :  l0 = 90,  phi1 = 90"
        assert head[2] == "#\tmap = 1   :  l0 = -90,  phi1 = -90"
        cols = head[4]
        assert cols.startswith("# map\tlam_X\tlam_Y\t%s\t" % keys["z"])
        assert cols.endswith("\t%s[0]\t%s[1]" % (keys["xy"], keys["xy"]))
    else:
        cols = head[0]
        assert cols.startswith("# map\t%s[0]\t%s[1]\t%s\t" % (keys["xy"], keys["xy"], keys["z"]))
        assert cols.endswith("\tdA")
    assert head[-1] == "#"
    assert "\t".join("N(%s[%d])" % (keys["z"], k) for k in range(5)) in cols
    for k in names:
        assert "\t" + k + "\t" in cols

    # beams written: every beam that received a row; with "output all bins"
    # also every map-0 beam
    dense = bool(keys.get("output all bins", 0))
    want_beams = [b for b in range(nb) if ob[b] or (dense and b % nmaps == 0)]
    assert len(body) == len(want_beams) * nz
    gx = [g(float(f32(x0 + f32(dx * f32(X))))) for X in range(nx)]
    gy = [g(float(f32(y0 + f32(dy * f32(Y))))) for Y in range(ny)]
    gz = [g(f32(z0 + f32(dz * f32(f32(Z * zw) / f32(zw))))) for Z in range(nz)]

    def asin(v):
        return math.asin(v) if -1.0 <= v <= 1.0 else float("nan")

    total = 0
    for j, b in enumerate(want_beams):
        X, rem = divmod(b, ny * nmaps)
        Y, m = divmod(rem, nmaps)
        x = float(f32(x0 + f32(dx * f32(X))))
        y = float(f32(y0 + f32(dy * f32(Y))))
        if ea:
            # peyton::math::lambert::deproject (common/projections.h:58-66)
            l0, sinphi1, cosphi1 = p.proj[m][0], p.proj[m][1], p.proj[m][2]
            phi1 = (90.0 if m == 0 else -90.0) * D2R
            r = math.sqrt(x * x + y * y)
            cc = 2 * asin(0.5 * r)
            lat = asin(math.cos(cc) * sinphi1 + y * math.sin(cc) * cosphi1 / r) if r != 0 else phi1
            lon = l0 + (math.atan2(x * math.sin(cc), r * cosphi1 * math.cos(cc) - y * sinphi1 * math.sin(cc))
                        if r != 0 else 0.0)
            tail = [g(lon * R2D), g(lat * R2D)]
        else:
            area = float(dx) * ((math.sin(min(y + 0.5 * float(dy), 90.0) * D2R) -
                                 math.sin(max(y - 0.5 * float(dy), -90.0) * D2R)) * R2D)
            tail = [g(area)]
        head4 = [str(m), gx[X], gy[Y]]
        for Z in range(nz):
            f = body[j * nz + Z]
            assert f[:3] == head4 and f[3] == gz[Z], (b, Z, f[:4])
            vals = [int(v) for v in f[4:4 + zw]]
            assert vals == oc[b, Z].tolist(), (b, Z)
            total += sum(vals[:5])
            for got_s, want_s in zip(f[4 + zw:], tail):
                # a NaN prints as "-nan" or "nan" by its sign bit
                assert got_s == want_s or ("nan" in got_s and "nan" in want_s), (b, Z, f)
            assert len(f) == 4 + zw + len(tail)
    assert total == 5 * rows
