"""The catalog run with the module chain's row passes on a second HIP stream
and two output chunks in flight, against the same run on one stream.

Workload: the full-sky catalog (|b| > 10 deg, 0.5 deg pixels, r 14-24) in
30,000,000-row chunks, so every chunk takes two chain calls (65,536 streams x
256 rows, then the rest) and the position buffers alternate within a chunk
and across chunks.  Every chunk's row count, component, visible count and
the rows of windows at its start, its end and both sides of the chain-call
boundary must be bit-identical between the two runs (the one-stream run is
the ordering the module kernels define, pinned against the oracle by the
config-0 and config-4 replay tests)."""
import ctypes as C

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import configs
from galfast_amd.catalog import Catalog
from galfast_amd.table import catalog_spec
import replay as R

pytestmark = pytest.mark.gpu

WIN = 8192
SUB = 65536 * 256


def windows(n):
    starts = {0, max(0, n - WIN)}
    for b in range(SUB, n, SUB):
        starts |= {b - WIN, b}
    return sorted(starts)


def sample(table, n, spec):
    """Rows of every window of every column, on the host."""
    L = G.lib()
    L.gf_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    out = {}
    for name, (dt, _) in spec.items():
        col = getattr(table, name)
        if not col.ptr:
            continue
        parts = []
        for a in windows(n):
            w = min(WIN, n - a)
            buf = np.empty((col.width, w), dtype=dt)
            assert buf.itemsize == col.elem
            for j in range(col.width):
                G.check(L.gf_memcpy_d2h(buf[j].ctypes.data, col.ptr + j * col.pitch + a * col.elem, w * col.elem,
                                        None))
            parts.append(buf)
        out[name] = np.concatenate(parts, axis=1)
    return out


def run(cmd, overlap):
    cat = Catalog(cmd, seed=42, chunk_rows=30_000_000)
    cat.set("overlap", overlap)
    cat.count()
    info = cat.info()
    spec = catalog_spec(nbands=info.nbands, nabsmag=max(1, info.nabsmag))
    for c in R.unwritten_columns(R.host_view(G.lib(), cat._h)):
        spec.pop(c, None)
    chunks = []
    st = cat.run(lambda t, n, comp: chunks.append((n, comp, sample(t, n, spec))))
    cat.close()
    return st, chunks


def test_overlapped_run_matches_one_stream_c1(tmp_path):
    cmd = configs.write_c1(str(tmp_path))
    s1, a = run(cmd, 1)
    s0, b = run(cmd, 0)
    assert s1.generated == s0.generated and s1.visible == s0.visible and s1.rows_chain == s0.rows_chain
    assert s1.n_chain == s0.n_chain and s1.n_chain > s1.n_emit   # chunks with two chain calls
    assert len(a) == len(b) >= 4
    for (n1, c1, t1), (n0, c0, t0) in zip(a, b):
        assert (n1, c1) == (n0, c0)
        assert t1.keys() == t0.keys()
        np.testing.assert_array_equal(t1["hidden"], t0["hidden"])
        vis = t0["hidden"][0] == 0
        assert vis.any()
        for name in t0:
            x, y = np.ascontiguousarray(t1[name][:, vis]), np.ascontiguousarray(t0[name][:, vis])
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name
