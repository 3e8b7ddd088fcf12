"""gf_skygen_order (drawSources' batch order) on the device against the
oracle's restatement, on random columns with planted ties in every key,
negative values and signed zeros, at a size that spans many sort tiles."""
import ctypes as C

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd.table import DeviceTable, HostTable, catalog_spec
import oracle as O
from devutil import stream_ptr, sync

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 1000, 300000])
def test_order_matches_oracle(n):
    rng = np.random.default_rng(n)
    spec = catalog_spec()
    h = HostTable(n, spec)
    h.arr("hidden")[0] = rng.random(n) < 0.15
    h.arr("lb")[0] = rng.integers(0, 200, n) * 1.8
    h.arr("lb")[1] = rng.integers(-50, 50, n) * 1.8
    h.arr("lb")[1][rng.random(n) < 0.01] = -0.0
    h.arr("DM")[0] = rng.integers(0, 5, n).astype(np.float32) - 2.0
    h.arr("absmag")[0] = rng.integers(-3, 3, n).astype(np.float32)
    h.arr("Am")[0] = rng.integers(0, 2, n).astype(np.float32)
    for c in ("projIdx", "comp"):
        h.arr(c)[0] = rng.integers(0, 2 ** 20, n)
    for c in ("projXY", "XYZ", "AmInf"):
        h.arr(c)[:] = rng.normal(size=h.arr(c).shape)
    h.arr("comp")[0] = np.arange(n)   # identifies the row
    d = DeviceTable(n, spec)
    d.upload(h)
    out = DeviceTable(n, spec)
    nv = C.c_uint64()
    print("gf_skygen_order n=%d" % n, flush=True)
    G.check(G.lib().gf_skygen_order(C.byref(d.skygen_cols()), n, C.byref(out.skygen_cols()), C.byref(nv),
                                    stream_ptr()), "gf_skygen_order")
    sync()
    ref = HostTable(n, spec)
    nref = O.lib().gfo_order(C.byref(h.skygen_cols()), n, C.byref(ref.skygen_cols()))
    assert nv.value == nref == int((h.arr("hidden")[0] == 0).sum())
    got = out.to_host()
    for c in ("lb", "projIdx", "projXY", "DM", "absmag", "XYZ", "Am", "AmInf", "comp", "hidden"):
        a = np.ascontiguousarray(got.arr(c)[:, :nref])
        b = np.ascontiguousarray(ref.arr(c)[:, :nref])
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), c
