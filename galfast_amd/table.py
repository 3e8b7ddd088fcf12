"""Structure-of-arrays catalog columns in host (numpy) or device (torch) memory.

Layout follows the reference's otable columns (common/otable.h:275-288,
common/cux.h:567-571): a column of width w holds w planes of ``n`` elements,
plane j starting ``j * pitch`` bytes after the first, pitch rounded up to a
multiple of 128 bytes.
"""
import numpy as np

from ._abi import gf_col, gf_table, gf_skygen_cols, TABLE_COLUMNS, SKYGEN_COLUMNS

PITCH_ALIGN = 128


def catalog_spec(nbands=5, nabsmag=1):
    """Column name -> (numpy dtype, width) of the catalog table."""
    f4, i4, f8 = np.float32, np.int32, np.float64
    return {
        "lb": (f8, 2), "projIdx": (i4, 1), "projXY": (f4, 2), "DM": (f4, 1), "absmag": (f4, 1),
        "XYZ": (f4, 3), "Am": (f4, 1), "AmInf": (f4, 1), "comp": (i4, 1), "hidden": (i4, 1),
        "FeH": (f4, 1), "FeH_comp": (i4, 1), "absmagSys": (f4, max(nabsmag, 1)),
        "absmagSysNcomp": (i4, 1), "vcyl": (f4, 3), "mags": (f4, nbands), "photoFlags": (i4, 1),
        "obsMags": (f4, nbands), "pm": (f4, 3), "radec": (f8, 2),
    }


def pitch_of(n, itemsize):
    b = n * itemsize
    return max(PITCH_ALIGN, (b + PITCH_ALIGN - 1) // PITCH_ALIGN * PITCH_ALIGN)


class HostTable:
    """Columns in host memory (numpy); views are (width, n) arrays."""

    def __init__(self, n, spec=None, fill=0):
        self.n = n
        self.spec = dict(spec or catalog_spec())
        self.buf = {}
        for name, (dt, w) in self.spec.items():
            it = np.dtype(dt).itemsize
            p = pitch_of(n, it)
            raw = np.zeros(w * p // it, dtype=dt)
            if fill:
                raw[:] = fill
            self.buf[name] = (raw, p)

    def arr(self, name):
        raw, p = self.buf[name]
        dt, w = self.spec[name]
        it = np.dtype(dt).itemsize
        return raw.reshape(w, p // it)[:, : self.n]

    def col(self, name):
        if name not in self.buf:
            return gf_col()
        raw, p = self.buf[name]
        dt, w = self.spec[name]
        return gf_col(raw.ctypes.data, p, w, np.dtype(dt).itemsize, 0)

    def table(self):
        return gf_table(*[self.col(n) for n in TABLE_COLUMNS])

    def skygen_cols(self):
        m = {"M": "absmag"}
        return gf_skygen_cols(*[self.col(m.get(n, n)) for n in SKYGEN_COLUMNS])


class DeviceTable:
    """The same layout in device memory (torch tensors on cuda:0)."""

    def __init__(self, n, spec=None, device="cuda"):
        import torch
        self.n = n
        self.spec = dict(spec or catalog_spec())
        self.buf = {}
        tmap = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32}
        for name, (dt, w) in self.spec.items():
            it = np.dtype(dt).itemsize
            p = pitch_of(n, it)
            t = torch.zeros(w * p // it, dtype=tmap[dt], device=device)
            self.buf[name] = (t, p)

    def col(self, name):
        if name not in self.buf:
            return gf_col()
        t, p = self.buf[name]
        dt, w = self.spec[name]
        return gf_col(t.data_ptr(), p, w, np.dtype(dt).itemsize, 0)

    def table(self):
        return gf_table(*[self.col(n) for n in TABLE_COLUMNS])

    def skygen_cols(self):
        m = {"M": "absmag"}
        return gf_skygen_cols(*[self.col(m.get(n, n)) for n in SKYGEN_COLUMNS])

    def upload(self, host):
        import torch
        for name, (t, p) in self.buf.items():
            if name in host.buf:
                raw, hp = host.buf[name]
                assert hp == p
                t.copy_(torch.from_numpy(raw))

    def to_host(self):
        h = HostTable(self.n, self.spec)
        for name, (t, p) in self.buf.items():
            raw, _ = h.buf[name]
            raw[:] = t.cpu().numpy()
        return h
