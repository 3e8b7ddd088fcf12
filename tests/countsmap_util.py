"""Helpers of the countsMap tests: parameter blocks as the host builds them
and the oracle's binning over host columns.  Test infrastructure."""
import ctypes as C
import math

import numpy as np

from galfast_amd import _abi as A
import oracle as O

D2R = 0.017453292519943295


def params(nx, ny, nz, x0=0.0, dx=1.0, y0=-90.0, dy=1.0, z0=0.0, dz=1.0, zmag_width=5, coadd=None,
           equalarea=False):
    """gf_countsmap_params; `coadd` a (ncoadd, zmag_width) array (host
    memory, kept alive on the returned block as .coadd)."""
    p = A.gf_countsmap_params()
    p.x0, p.dx, p.y0, p.dy, p.z0, p.dz = x0, dx, y0, dy, z0, dz
    p.nx, p.ny, p.nz = nx, ny, nz
    p.zmag_width = zmag_width
    p.equalarea = int(equalarea)
    co = np.ascontiguousarray(np.zeros((0, zmag_width)) if coadd is None else coadd, dtype=np.float32)
    p.ncoadd = co.shape[0]
    p.coadd_offsets = co.ctypes.data if co.size else None
    for i, a in enumerate((90.0, -90.0)):
        r = a * D2R
        p.proj[i][0] = r
        p.proj[i][1] = math.sin(r)
        p.proj[i][2] = math.cos(r)
    p.coadd = co
    return p


def shape(p):
    nmaps = 2 if p.equalarea else 1
    return p.nx * p.ny * nmaps, p.nz, p.zmag_width + p.ncoadd


def oracle_bin(p, xy, z, hidden, n, counts=None, beams=None):
    """The oracle's binning of n rows of host columns, added to counts /
    beams (fresh zero arrays when None).  Returns (counts, beams, total)."""
    nb, nz, zw = shape(p)
    if counts is None:
        counts = np.zeros(nb * nz * zw, dtype=np.uint32)
        beams = np.zeros(nb, dtype=np.uint8)
    tot = O.lib().gfo_countsmap_bin(C.byref(p), xy, z, hidden, n, counts.ctypes.data, beams.ctypes.data)
    return counts, beams, tot


def host_cols(xy, z, hidden):
    """gf_col views of numpy arrays: xy (2, n) float64, z (w, n) float32,
    hidden (n,) int32; the arrays are kept alive by the caller."""
    assert xy.flags.c_contiguous and z.flags.c_contiguous and hidden.flags.c_contiguous
    n = xy.shape[1]
    return (A.gf_col(xy.ctypes.data, n * 8, 2, 8, 0), A.gf_col(z.ctypes.data, n * 4, z.shape[0], 4, 0),
            A.gf_col(hidden.ctypes.data, n * 4, 1, 4, 0))
