"""Output stages: the device text formatter against glibc's snprintf, and
config-0 / config-1 catalogs written as FITS and text, read back and compared
with the catalog's own device columns (visible rows, reference column order
and names, component ids)."""
import ctypes as C
import os

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import _abi as A, configs
from galfast_amd.catalog import Catalog
from galfast_amd.table import DeviceTable, HostTable, catalog_spec
from devutil import stream_ptr
from test_gpu_catalog import collect

pytestmark = pytest.mark.gpu

LIBC = C.CDLL("libc.so.6")


def glibc(fmt, v):
    buf = C.create_string_buffer(64)
    if fmt.endswith("d"):
        LIBC.snprintf(buf, 64, fmt.encode(), C.c_int(int(v)))
    else:
        LIBC.snprintf(buf, 64, fmt.encode(), C.c_double(float(v)))
    return buf.value.decode()


class gf_outcol(C.Structure):
    _fields_ = [("col", A.gf_col), ("map", C.c_void_p), ("nmap", C.c_uint32), ("type", C.c_char),
                ("flags", C.c_char), ("width", C.c_int16), ("prec", C.c_int32)]


def outcol(col, fmt, type_):
    o = gf_outcol()
    o.col = col
    o.type = type_.encode()
    body = fmt[1:-1]
    o.flags = b"\x01" if body.startswith(" ") else b"\x00"
    body = body.lstrip(" ")
    w, _, p = body.partition(".")
    o.width = int(w or 0)
    o.prec = int(p) if fmt.endswith("f") else -1
    return o


FORMATS = [("% 13.8f", "D", "lb"), ("% 10.2f", "E", "XYZ"), ("% 7.3f", "E", "DM"), ("%6.3f", "E", "Am"),
           ("% 6.3f", "E", "FeH"), ("% 7.1f", "E", "vcyl"), ("% 8.2f", "E", "pm"), ("%3d", "J", "comp"),
           ("% 4d", "J", "photoFlags"), ("%1d", "J", "absmagSysNcomp")]


def test_text_formatter_matches_glibc():
    rng = np.random.default_rng(11)
    n = 4000
    spec = catalog_spec()
    h = HostTable(n, spec)
    # magnitudes of every scale, exact ties at the printed precision, signed zeros
    scale = 10.0 ** rng.integers(-9, 9, n)
    vals = rng.normal(size=n) * scale
    ties = (rng.integers(-10 ** 6, 10 ** 6, n) + 0.5) / 10 ** rng.integers(0, 3, n)
    vals[::3] = ties[::3]
    vals[:6] = [0.0, -0.0, 0.125, -0.125, 2.5e-9, -4.9999999e-4]
    h.arr("lb")[0] = vals
    h.arr("lb")[1] = -vals[::-1]
    for c in ("XYZ", "DM", "Am", "FeH", "vcyl", "pm"):
        a = h.arr(c)
        a[:] = np.resize(vals, a.shape).astype(np.float32)
    for c in ("comp", "photoFlags", "absmagSysNcomp"):
        h.arr(c)[0] = rng.integers(-5000, 5000, n)
    h.arr("hidden")[0] = rng.random(n) < 0.1
    d = DeviceTable(n, spec)
    d.upload(h)
    cols = (gf_outcol * len(FORMATS))(*[outcol(d.col(c), f, t) for f, t, c in FORMATS])
    cap = n * 400
    import torch
    out = torch.zeros(cap, dtype=torch.uint8, device="cuda")
    nb, nv = C.c_uint64(), C.c_uint64()
    L = G.lib()
    L.gf_text_format_rows.argtypes = [C.c_void_p, C.c_int, A.gf_col, C.c_uint64, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_void_p]
    print("gf_text_format_rows n=%d" % n, flush=True)
    G.check(L.gf_text_format_rows(cols, len(FORMATS), d.col("hidden"), n, out.data_ptr(), cap, C.byref(nb),
                                  C.byref(nv), stream_ptr()), "gf_text_format_rows")
    text = bytes(out[:nb.value].cpu().numpy()).decode()
    lines = text.split("\n")
    assert lines[-1] == "" and len(lines) - 1 == nv.value
    vis = np.flatnonzero(h.arr("hidden")[0] == 0)
    assert nv.value == vis.size
    for li, r in zip(lines[:-1], vis):
        want = []
        for f, t, c in FORMATS:
            a = h.arr(c)
            for e in range(a.shape[0]):
                want.append(glibc(f, a[e, r]))
        assert li == " ".join(want), (r, li, " ".join(want))


def read_fits_tables(path):
    """{EXTNAME: (header dict, ordered (TTYPE, TFORM) list, raw data bytes)}
    of the binary tables of a FITS file."""
    raw = open(path, "rb").read()
    out = {}
    at = 0
    first = True
    while at < len(raw):
        hdr = {}
        order = []
        while True:
            blk = raw[at:at + 2880]
            at += 2880
            done = False
            for i in range(36):
                card = blk[80 * i:80 * i + 80].decode()
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if card[8:10] == "= ":
                    v = card[10:].split(" / ")[0].strip()
                    v = v.strip("'").strip() if v.startswith("'") else v
                    hdr[key] = v
                    order.append(key)
            if done:
                break
        if first:
            first = False
            continue
        n1, n2 = int(hdr["NAXIS1"]), int(hdr["NAXIS2"])
        nbytes = n1 * n2
        data = raw[at:at + nbytes]
        at += (nbytes + 2879) // 2880 * 2880
        cols = [(hdr["TTYPE%d" % i], hdr["TFORM%d" % i]) for i in range(1, int(hdr["TFIELDS"]) + 1)]
        out[hdr["EXTNAME"]] = (hdr, cols, data)
    return out


def table_column(name, band="SYNr", bandset="SYN"):
    return {"lb": "lb", "radec": "radec", "XYZ": "XYZ", "DM": "DM", "abs" + band: "absmag", "comp": "comp",
            "FeH": "FeH", "FeH_comp": "FeH_comp", "vcyl": "vcyl", "pmradec": "pm", "pmlb": "pm", "Am": "Am",
            "AmInf": "AmInf", bandset: "mags", bandset + "PhotoFlags": "photoFlags", "obs" + bandset: "obsMags",
            "abs" + band + "Sys": "absmagSys", "abs" + band + "SysNcomp": "absmagSysNcomp"}[name]


@pytest.mark.parametrize("cfg,order", [
    ("c0", ["lb", "XYZ", "DM", "absSYNr", "comp", "FeH", "Am", "AmInf", "SYN", "SYNPhotoFlags"]),
    ("c1", ["lb", "radec", "XYZ", "DM", "absSYNr", "comp", "FeH", "vcyl", "pmradec", "Am", "AmInf", "obsSYN", "SYN",
            "SYNPhotoFlags"])])
def test_fits_and_text_output(tmp_path, cfg, order):
    d = str(tmp_path)
    if cfg == "c0":
        cmd = configs.write_c0(d)
        kw = dict(chunk_rows=40000)
    else:
        cmd = configs.write_c1(d)
        kw = dict(chunk_rows=30_000_000)
    # the expected rows: the same catalog (same seed) run into host copies
    cat = Catalog(cmd, seed=42, **kw)
    cat.count()
    ids = np.array(cat.component_ids(), dtype=np.int64)
    info = cat.info()
    spec = catalog_spec(nbands=info.nbands, nabsmag=max(1, info.nabsmag))
    if cfg == "c1":
        # the full-sky job scaled to 3e5 stars (the reference's nstars key)
        cat.set("nstars", 300_000)
    cat.count()
    got = collect(cat, spec)
    cat.close()
    vis_rows = {}
    for name in spec:
        parts = [t.arr(name)[:, :n][:, t.arr("hidden")[0, :n] == 0] for t, n, _ in got]
        vis_rows[name] = np.concatenate(parts, axis=1)
    nvis = vis_rows["hidden"].shape[1]
    assert nvis > 1000
    for kind in ("fits", "text"):
        cat = Catalog(cmd, seed=42, **kw)
        if cfg == "c1":
            cat.set("nstars", 300_000)
        cat.count()
        path = os.path.join(d, "sky." + kind)
        print("writing %s %s" % (cfg, kind), flush=True)
        st, rows = cat.write(path, kind)
        cat.close()
        assert rows == nvis == st.visible
        if kind == "fits":
            tabs = read_fits_tables(path)
            hdr, cols, data = tabs["CATALOG"]
            assert [c for c, _ in cols] == order
            assert int(hdr["NAXIS2"]) == nvis
            dt = []
            for name, form in cols:
                w, t = int(form[:-1]), form[-1]
                dt.append((name, {"E": ">f4", "D": ">f8", "J": ">i4"}[t], (w,)))
            arr = np.frombuffer(data, dtype=np.dtype(dt), count=nvis)
            for name, form in cols:
                want = vis_rows[table_column(name)].T
                if name == "comp":
                    want = ids[want]
                assert np.array_equal(arr[name].astype(want.dtype), want), name
            meta = tabs["METADATA"]
            assert meta[2].decode().startswith("lb[2] ")
        else:
            with open(path) as f:
                head = f.readline()
                assert head.startswith("# lb[2] ")
                assert head.split()[1:len(order) + 1][0] == "lb[2]"
                lines = f.read().split("\n")
            assert lines[-1] == "" and len(lines) - 1 == nvis
            fmts = {"lb": "% 13.8f", "radec": "% 13.8f", "XYZ": "% 10.2f", "DM": "% 7.3f", "absSYNr": "% 7.3f",
                    "comp": "%3d", "FeH": "% 6.3f", "vcyl": "% 7.1f", "pmradec": "% 8.2f", "Am": "%6.3f",
                    "AmInf": "%6.3f", "obsSYN": "% 7.3f", "SYN": "% 7.3f", "SYNPhotoFlags": "% 4d"}
            step = max(1, nvis // 2000)
            for r in range(0, nvis, step):
                want = []
                for name in order:
                    a = vis_rows[table_column(name)]
                    for e in range(a.shape[0]):
                        v = a[e, r]
                        if name == "comp":
                            v = ids[v]
                        want.append(fmts[name] % (float(v) if fmts[name].endswith("f") else int(v)))
                assert lines[r] == " ".join(want), (r, lines[r])


def test_fits_pieces_across_chunks(tmp_path):
    """A config-1-shaped job of ~1.2e7 stars in 4.5e6-row chunks: the writer packs
    each chunk in pieces of 2^22 rows, copies one piece to the host while
    packing the next, and carries the last piece of a chunk into the next
    chunk's callback; the file must hold every visible row once, in order."""
    d = str(tmp_path)
    cmd = configs.write_c1(d)
    kw = dict(chunk_rows=4_500_000)
    cat = Catalog(cmd, seed=42, **kw)
    cat.set("nstars", 12_000_000)
    cat.count()
    ids = np.array(cat.component_ids(), dtype=np.int64)
    info = cat.info()
    spec = catalog_spec(nbands=info.nbands, nabsmag=max(1, info.nabsmag))
    got = collect(cat, spec)
    cat.close()
    assert len(got) >= 2 and max(n for _, n, _ in got) > (1 << 22)
    cat = Catalog(cmd, seed=42, **kw)
    cat.set("nstars", 12_000_000)
    cat.count()
    path = os.path.join(d, "sky.fits")
    print("writing pieces", flush=True)
    st, rows = cat.write(path, "fits")
    cat.close()
    hdr, cols, data = read_fits_tables(path)["CATALOG"]
    nvis = int(hdr["NAXIS2"])
    assert rows == nvis == st.visible
    dt = [(name, {"E": ">f4", "D": ">f8", "J": ">i4"}[form[-1]], (int(form[:-1]),)) for name, form in cols]
    arr = np.frombuffer(data, dtype=np.dtype(dt), count=nvis)
    for name, form in cols:
        parts = [t.arr(table_column(name))[:, :n][:, t.arr("hidden")[0, :n] == 0] for t, n, _ in got]
        want = np.concatenate(parts, axis=1).T
        if name == "comp":
            want = ids[want]
        assert np.array_equal(arr[name].astype(want.dtype), want), name
