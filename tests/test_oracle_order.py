"""Pin the oracle's batch order (drawSources' std::sort with star_comp and
the truncation at the first hidden row) against numpy's stable lexicographic
sort, on random columns with planted ties in every key."""
import ctypes as C

import numpy as np

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O


def test_order_is_lexsort_of_visible_rows():
    rng = np.random.default_rng(3)
    n = 5000
    spec = catalog_spec()
    t = HostTable(n, spec)
    # few distinct values so that every key ties often
    t.arr("hidden")[0] = rng.random(n) < 0.2
    t.arr("lb")[0] = rng.integers(0, 40, n) * 0.5
    t.arr("lb")[1] = rng.integers(-5, 5, n) * 0.25
    t.arr("DM")[0] = rng.integers(0, 4, n).astype(np.float32)
    t.arr("absmag")[0] = rng.integers(0, 3, n).astype(np.float32)
    t.arr("Am")[0] = rng.integers(0, 2, n).astype(np.float32)
    t.arr("comp")[0] = np.arange(n)
    out = HostTable(n, spec)
    nv = O.lib().gfo_order(C.byref(t.skygen_cols()), n, C.byref(out.skygen_cols()))
    vis = np.flatnonzero(t.arr("hidden")[0] == 0)
    assert nv == vis.size
    keys = [t.arr("Am")[0][vis], t.arr("absmag")[0][vis], t.arr("DM")[0][vis], t.arr("lb")[1][vis],
            t.arr("lb")[0][vis]]
    expect = vis[np.lexsort(keys)]   # stable: ties keep the input order
    assert np.array_equal(out.arr("comp")[0][:nv], expect)
    assert np.all(out.arr("hidden")[0][:nv] == 0)
    # a planted mistake (b before l) orders differently
    wrong = vis[np.lexsort([keys[0], keys[1], keys[2], keys[4], keys[3]])]
    assert not np.array_equal(wrong, expect)
