"""End-to-end catalog job (config 0) on the GPU against the oracle replay."""
import ctypes as C
import os

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import configs
from galfast_amd.catalog import Catalog
from galfast_amd.table import HostTable, catalog_spec
from galfast_amd._abi import TABLE_COLUMNS
import replay as R

pytestmark = pytest.mark.gpu


def fetch(table, n, spec):
    """Copy n rows of every device column into a host table."""
    L = G.lib()
    L.gf_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    h = HostTable(n, spec)
    for name in spec:
        col = getattr(table, name)
        if not col.ptr:
            continue
        raw, p = h.buf[name]
        assert col.elem == raw.dtype.itemsize
        for j in range(col.width):
            dst = raw.ctypes.data + j * p
            G.check(L.gf_memcpy_d2h(dst, col.ptr + j * col.pitch, n * col.elem, None))
    return h


def collect(cat, spec):
    out = []
    cat.run(lambda t, n, comp: out.append((fetch(t, n, spec), n, comp)))
    return out


@pytest.fixture(scope="module")
def c0(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c0"))
    return configs.write_c0(d)


def test_c0_catalog_matches_oracle_replay(c0):
    cat = Catalog(c0, seed=42, chunk_rows=40000)
    info = cat.info()
    assert info.nmodels == 3 and info.nbands == 5
    expected = cat.count()
    assert 5e4 < expected < 2e5
    view = R.host_view(G.lib(), cat._h)
    spec = catalog_spec(nbands=view.nbands, nabsmag=view.nabsmag)
    got = []
    st = cat.run(lambda t, n, comp: got.append((fetch(t, n, spec), n, comp)))
    ref = []
    rs = R.replay(view, chunk_rows=40000, on_chunk=lambda t, n, comp: ref.append((t, n, comp)))
    # the run's visible count (added up by the draw passes) is the replay's
    assert st.visible == rs["visible"] and st.generated == rs["generated"]
    assert [(n, c) for _, n, c in got] == [(n, c) for _, n, c in ref]
    # every column of every visible row is bit-identical
    exact = ["lb", "projIdx", "projXY", "DM", "absmag", "XYZ", "comp", "Am", "AmInf", "hidden", "FeH", "mags",
             "photoFlags"]
    nvis = 0
    for (a, n, _), (b, _, _) in zip(got, ref):
        assert np.array_equal(a.arr("hidden"), b.arr("hidden"))
        vis = b.arr("hidden")[0] == 0
        nvis += vis.sum()
        for c in exact:
            bad = a.arr(c)[:, vis] != b.arr(c)[:, vis]
            assert not bad.any(), "%s: %d of %d differ" % (c, bad.sum(), bad.size)
    assert nvis > 0.5 * expected


def check_reference_draw(cmd, seed, capacity=None, keep=None):
    """The catalog's reference draw mode: the crawl in batches of at most
    `capacity` stars (one batch per model by default), per batch the order
    of drawSources (hidden rows dropped, the rest sorted by l, b, DM, M,
    Am), the clipper and the module chain over the visible rows; every
    column bit-identical to the oracle's restatement.  keep: a list that
    receives the (table, rows) of every batch.
    Returns (rows handed out, rows drawn, rows hidden by the clipper)."""
    cat = Catalog(cmd, seed=seed)
    cat.set("draw", 1)
    if capacity:
        cat.set("chunk_rows", capacity)
    cat.count()
    view = R.host_view(G.lib(), cat._h)
    spec = catalog_spec(nbands=view.nbands, nabsmag=view.nabsmag)
    got = collect(cat, spec)
    ref = []
    info = R.replay_reference(view, on_chunk=lambda t, n, comp: ref.append((t, n, comp)), capacity=capacity)
    cat.close()
    if keep is not None:
        keep.extend(got)
    if capacity is None:
        assert len(got) == view.nmodels
    assert [(n, c) for _, n, c in got] == [(n, c) for _, n, c in ref]
    unwritten = R.unwritten_columns(view)
    clipped = 0
    for (a, n, _), (b, _, _) in zip(got, ref):
        assert np.all(np.diff(b.arr("lb")[0, :n]) >= 0)
        assert np.array_equal(a.arr("hidden")[:, :n], b.arr("hidden")[:, :n])
        vis = b.arr("hidden")[0, :n] == 0
        clipped += int((~vis).sum())
        for c in spec:
            if c in unwritten:
                continue
            x = np.ascontiguousarray(a.arr(c)[:, :n][:, vis])
            y = np.ascontiguousarray(b.arr(c)[:, :n][:, vis])
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), c
        for c in ("lb", "projIdx", "projXY", "DM", "absmag", "XYZ", "comp", "Am", "AmInf"):
            assert np.array_equal(a.arr(c)[:, :n], b.arr(c)[:, :n]), c
    return sum(n for _, n, _ in got), info["generated"], clipped


def test_c0_reference_draw_matches_oracle(c0):
    out, drawn, clipped = check_reference_draw(c0, 42)
    assert 0 < out <= drawn
    assert clipped > 0   # the rect footprint's edge pixels


SKY_COLS = ("lb", "projIdx", "projXY", "DM", "absmag", "XYZ", "comp", "Am", "AmInf")


def _visible_rows(batches, comp):
    """The skygen columns of every handed-out row of model `comp` as one
    sorted byte array."""
    rows = []
    for t, n, c in batches:
        if c != comp:
            continue
        vis = t.arr("hidden")[0, :n] == 0
        parts = [np.ascontiguousarray(t.arr(c)[:, :n][:, vis].T).view(np.uint8).reshape(int(vis.sum()), -1)
                 for c in SKY_COLS]
        rows.append(np.concatenate(parts, axis=1))
    r = np.concatenate(rows)
    return r[np.lexsort(r.T[::-1])]


def test_c0_capacity_limited_reference_draw(c0):
    """The draw stopped every 4,999 stars and resumed where each batch
    stopped (skygenHost::drawSources with runtime_state): every batch
    bit-identical to the oracle's batch, and the first model's stars handed
    out over all its batches the same multiset as its one unlimited batch.
    (The module chain runs between batches on the same streams, so later
    models start from other stream states, as in the reference.)"""
    limited, unlimited = [], []
    check_reference_draw(c0, 42, capacity=4999, keep=limited)
    check_reference_draw(c0, 42, keep=unlimited)
    assert sum(1 for _, _, c in limited if c == 0) > 3
    assert max(n for _, n, _ in limited) <= 4999
    a, b = _visible_rows(limited, 0), _visible_rows(unlimited, 0)
    assert a.shape == b.shape and np.array_equal(a, b)


def test_chunk_rows_bounded_by_the_column_pitch(c0):
    """A chunk's 8-byte columns must keep a 32-bit plane pitch: chunk_rows
    above ((2^32 - 128) / 8) rows (and below one) is refused, the limit
    itself is accepted."""
    cat = Catalog(c0, seed=7)
    try:
        limit = ((1 << 32) - 128) // 8
        cat.set("chunk_rows", limit)
        for bad in (limit + 1, 1 << 32, 0):
            with pytest.raises(G.GalfastError):
                cat.set("chunk_rows", bad)
    finally:
        cat.close()


def test_c0_generated_matches_expected(c0):
    cat = Catalog(c0, seed=7)
    cat.count()
    st = cat.run()
    assert abs(st.generated - st.expected) < 5 * np.sqrt(st.expected)
    assert 0 < st.visible <= st.generated


@pytest.fixture(scope="module")
def c4s(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c4s"))
    return configs.write_c4_small(d)


def test_c4_reference_draw_matches_oracle(c4s):
    """With extinction, stars pushed past the faint limit are hidden by the
    draw and dropped by the batch order."""
    out, drawn, _ = check_reference_draw(c4s, 5)
    assert 0 < out < drawn


def test_c4_extinction_multiples_six_bands_match_oracle_replay(c4s):
    """Config 4's ingredients on a small sky at low latitude: extinction
    cubes (FITS, analytic dust disk) sampled in the count (per-pixel
    minimum) and the draw (Am, AmInf), unresolved binaries and 6 bands;
    every column of every visible row bit-identical to the oracle replay."""
    cat = Catalog(c4s, seed=5, chunk_rows=60000)
    info = cat.info()
    assert info.nbands == 6
    expected = cat.count()
    view = R.host_view(G.lib(), cat._h)
    assert view.ext.beam.data and view.ext.north.data and view.ext.south.data
    spec = catalog_spec(nbands=view.nbands, nabsmag=view.nabsmag)
    got = collect(cat, spec)
    ref = []
    R.replay(view, chunk_rows=60000, on_chunk=lambda t, n, comp: ref.append((t, n, comp)))
    assert [(n, c) for _, n, c in got] == [(n, c) for _, n, c in ref]
    unwritten = R.unwritten_columns(view)
    assert "FeH_comp" in unwritten and "mags" not in unwritten
    nvis = 0
    am_pos = 0
    for (a, n, _), (b, _, _) in zip(got, ref):
        assert np.array_equal(a.arr("hidden"), b.arr("hidden"))
        vis = b.arr("hidden")[0] == 0
        nvis += vis.sum()
        am_pos += (b.arr("Am")[0][vis] > 0).sum()
        for c in spec:
            if c in unwritten:
                continue
            x, y = a.arr(c)[:, vis], b.arr(c)[:, vis]
            bad = (x != y) & ~(np.isnan(x) & np.isnan(y)) if x.dtype.kind == "f" else (x != y)
            assert not bad.any(), "%s: %d of %d differ" % (c, bad.sum(), bad.size)
    assert nvis > 0.3 * expected
    assert am_pos > 0.9 * nvis          # the dust disk reddens nearly every star


@pytest.fixture(scope="module")
def c0narrow(tmp_path_factory):
    """Config 0 with the thin disk's luminosity function tabulated only over
    Mr in [5, 12] (the job's range is 3-15)."""
    d = str(tmp_path_factory.mktemp("c0n"))
    cmd = configs.write_c0(d)
    M, phi = configs.seeded_lf()
    keep = (M >= 5.0 - 1e-9) & (M <= 12.0 + 1e-9)
    np.savetxt(os.path.join(d, "lf_narrow.txt"), np.c_[M[keep], phi[keep]], fmt="%.10g")
    path = os.path.join(d, "model.thin.conf")
    with open(path) as f:
        text = f.read()
    with open(path, "w") as f:
        f.write(text.replace("lumfunc = lf.txt", "lumfunc = lf_narrow.txt"))
    return cmd


def test_model_absolute_magnitude_range_narrowed_to_its_lf(c0narrow):
    """hint_absmag + reset_absmag (skygen/skyconfig_impl.h:258-263,
    skygen/model_lib.h:97-112, skygen/os_skygen.cpp:401-407): the thin
    disk's bins cover only its LF's texture span, the other models keep
    the job's; the catalog matches the oracle replay row for row, and
    the thin disk's stars lie in its bins."""
    cat = Catalog(c0narrow, seed=11, chunk_rows=40000)
    cat.count()
    view = R.host_view(G.lib(), cat._h)
    m0 = view.models[0]
    X0 = np.float32(m0.lf.tc[0])
    X1 = np.float32(np.float64(m0.lf.n - 1) / np.float64(np.float32(m0.lf.tc[1])) + np.float64(X0))
    dM = np.float32(view.sp.dM)
    sp0 = R.model_sp(view, 0)
    assert np.float32(sp0.M0) == X0 and X0 > np.float32(view.sp.M0)
    assert sp0.nM == int(round(float((X1 - X0) / dM)))
    assert np.float32(sp0.M1) == np.float32(np.float64(X1) - 0.5 * np.float64(dM))
    assert sp0.nM < view.sp.nM
    for mi in (1, 2):
        sp = R.model_sp(view, mi)
        assert (sp.M0, sp.M1, sp.nM) == (view.sp.M0, view.sp.M1, view.sp.nM)
    spec = catalog_spec(nbands=view.nbands, nabsmag=view.nabsmag)
    got = collect(cat, spec)
    ref = []
    R.replay(view, chunk_rows=40000, on_chunk=lambda t, n, comp: ref.append((t, n, comp)))
    cat.close()
    assert [(n, c) for _, n, c in got] == [(n, c) for _, n, c in ref]
    thin = 0
    for (a, n, comp), (b, _, _) in zip(got, ref):
        vis = b.arr("hidden")[0] == 0
        for c in ("lb", "DM", "absmag", "XYZ", "hidden", "FeH", "mags"):
            assert np.array_equal(a.arr(c)[:, vis], b.arr(c)[:, vis]), c
        if comp == 0:
            M = a.arr("absmag")[0][:n]
            assert M.min() >= np.float32(sp0.M1) - (sp0.nM - 0.5) * dM - 1e-5
            assert M.max() <= np.float32(sp0.M1) + 0.5 * dM + 1e-5
            thin += n
    assert thin > 1000
