"""Request server on the GPU: every request's rows are the same alone and
in a batch, and match the oracle replay of that request."""
import ctypes as C

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import configs
from galfast_amd.catalog import Catalog
from galfast_amd.server import Server
from galfast_amd.table import HostTable, catalog_spec
import replay as R

pytestmark = pytest.mark.gpu

SPEC = catalog_spec(nbands=5, nabsmag=1)

# (id, l, b, radius deg, m1, seed): uneven sizes, both hemispheres, one beam
# across l = 0 and one near the pole
REQUESTS = [
    (1, 90.0, 60.0, 0.6, 21.0, 11),
    (2, 200.0, -35.0, 1.5, 20.5, 12),
    (3, 359.8, 25.0, 0.3, 22.0, 13),
    (4, 30.0, 85.0, 2.0, 20.0, 14),
    (5, 120.0, -70.0, 0.1, 23.0, 15),
    (6, 250.0, 12.0, 1.0, 20.2, 16),
    (7, 10.0, -8.0, 0.4, 21.5, 17),
    (8, 300.0, 45.0, 0.05, 24.0, 18),
]


def fetch_rows(table, row0, n, spec=SPEC):
    L = G.lib()
    L.gf_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    h = HostTable(n, spec)
    if n == 0:
        return h
    for name in spec:
        col = getattr(table, name)
        raw, p = h.buf[name]
        for j in range(col.width):
            G.check(L.gf_memcpy_d2h(raw.ctypes.data + j * p, col.ptr + j * col.pitch + row0 * col.elem,
                                    n * col.elem, None))
    return h


def serve(conf, reqs, **kw):
    got = {}

    def on_result(res, table, row0, nrows):
        assert res.status == 0, res.status
        got[res.id] = (res, fetch_rows(table, row0, nrows))

    srv = Server(conf, on_result, **kw)
    try:
        for r in reqs:
            srv.submit(*r)
        srv.wait()
        st = srv.stats()
    finally:
        srv.close()
    return got, st


SKYGEN = ("lb", "projIdx", "projXY", "DM", "absmag", "XYZ", "Am", "AmInf", "comp", "hidden")
# columns only GaussianFeH and unresolvedMultiples write: the serving job has
# neither, so the batch tables hold whatever their memory held before
UNWRITTEN = ("FeH_comp", "absmagSys", "absmagSysNcomp")


def same_rows(a, b, what, spec=SPEC, unwritten=UNWRITTEN):
    """Skygen columns on every row; module columns on visible rows (the
    modules skip hidden rows and leave those cells as they were)."""
    assert a.n == b.n, "%s: %d vs %d rows" % (what, a.n, b.n)
    vis = b.arr("hidden")[0] == 0
    for c in spec:
        if c in unwritten:
            continue
        x, y = a.arr(c), b.arr(c)
        if c not in SKYGEN:
            x, y = x[:, vis], y[:, vis]
        bad = (x != y) & ~(np.isnan(x) & np.isnan(y)) if x.dtype.kind == "f" else (x != y)
        assert not bad.any(), "%s: column %s, %d of %d differ" % (what, c, bad.sum(), bad.size)


@pytest.fixture(scope="module")
def base(tmp_path_factory):
    return configs.write_c3(str(tmp_path_factory.mktemp("c3")), m1=25.0)


@pytest.fixture(scope="module")
def alone(base):
    return serve(base, REQUESTS, plain=True)


def test_plain_server_serves_one_at_a_time(alone):
    got, st = alone
    assert st.submitted == st.completed == len(REQUESTS)
    assert st.batches == len(REQUESTS)
    assert all(res.batch_size == 1 for res, _ in got.values())
    assert st.rows == sum(res.generated for res, _ in got.values())
    assert st.visible == sum(res.visible for res, _ in got.values())


def test_request_rows_same_alone_and_in_mix(base, alone):
    ref, _ = alone
    mix, st = serve(base, REQUESTS, max_batch=16, workers=2, linger_us=300000)
    assert st.completed == len(REQUESTS)
    assert max(res.batch_size for res, _ in mix.values()) > 1, "no batching happened"
    for rid, (res, t) in mix.items():
        r0, t0 = ref[rid]
        assert (res.generated, res.visible, res.npixels) == (r0.generated, r0.visible, r0.npixels)
        assert res.expected == r0.expected
        same_rows(t, t0, "request %d" % rid)
    # a second mix in another order and batch shape gives the same rows again
    mix2, _ = serve(base, REQUESTS[::-1], max_batch=3, workers=2, linger_us=100000)
    for rid, (res, t) in mix2.items():
        same_rows(t, ref[rid][1], "request %d (reversed)" % rid)


def test_request_without_stars_beside_others(base, alone):
    """A beam of 3e-6 deg^2 whose faint limit leaves one apparent-magnitude
    bin draws no star; batched with ordinary requests it completes with no
    rows and leaves theirs unchanged."""
    ref, _ = alone
    empty = (50, 140.0, 30.0, 0.001, 14.05, 77)
    got, st = serve(base, REQUESTS[:3] + [empty], max_batch=8, workers=1, linger_us=300000)
    assert st.completed == 4
    res, t = got[50]
    assert res.expected < 1e-3 and res.generated == 0 and res.visible == 0 and t.n == 0
    assert res.batch_size > 1
    for rid in (1, 2, 3):
        same_rows(got[rid][1], ref[rid][1], "request %d beside an empty one" % rid)


def test_submit_rejects_malformed_requests(base):
    srv = Server(base, None, plain=True)
    try:
        for bad in [(1, 90.0, 60.0, 0.5, 14.0, 1),            # no magnitude bin between m0 = 14 and m1
                    (2, 90.0, 60.0, 0.5, float("nan"), 1),
                    (3, float("nan"), 60.0, 0.5, 22.0, 1),
                    (4, 90.0, 95.0, 0.5, 22.0, 1),            # |b| > 90
                    (5, 90.0, 60.0, 0.0, 22.0, 1)]:           # rect with l1 <= l
            with pytest.raises(G.GalfastError):
                srv.submit(*bad)
        srv.submit(6, 90.0, 60.0, 0.5, 14.05, 1)            # one bin: accepted
        srv.wait()
    finally:
        srv.close()


def test_captured_chain_matches_direct_launches(base, alone):
    """Every batch replays the worker's captured HIP graph of the module
    chain (plan, emission and the count run eagerly between replays); the
    same requests with the chain launched directly give the same rows."""
    ref, _ = alone
    eager, st = serve(base, REQUESTS, plain=True, eager_chain=True)
    assert st.completed == len(REQUESTS)
    for rid, (res, t) in eager.items():
        same_rows(t, ref[rid][1], "request %d (direct launches)" % rid)


@pytest.mark.parametrize("rid", [1, 3, 5])
def test_request_matches_oracle_replay(tmp_path, alone, rid):
    got, _ = alone
    _, l, b, rad, m1, seed = next(r for r in REQUESTS if r[0] == rid)
    m1 = float(np.float32(m1))
    conf = configs.write_c3(str(tmp_path), m1=m1, beam=(l, b, rad))
    cat = Catalog(conf, seed=seed)
    cat.count()
    view = R.host_view(G.lib(), cat._h)
    res, t = got[rid]
    assert res.npixels == view.sp.npixels
    ref, expected = R.replay_request(view)
    assert abs(res.expected - expected) <= 1e-12 * expected
    assert res.generated == ref.n
    same_rows(t, ref, "request %d vs oracle" % rid)
    assert res.visible == int((ref.arr("hidden")[0] == 0).sum())


GF_E_BUDGET = -7
BIG = (99, 60.0, 40.0, 6.0, 25.0, 19)    # ~113 deg^2 to r = 25


def serve_status(conf, reqs, spec=SPEC, **kw):
    """serve(), keeping failed requests (status != 0, no rows)."""
    got = {}

    def on_result(res, table, row0, nrows):
        got[res.id] = (res, fetch_rows(table, row0, nrows, spec) if res.status == 0 else None)

    srv = Server(conf, on_result, **kw)
    try:
        for r in reqs:
            srv.submit(*r)
        srv.wait()
    finally:
        srv.close()
    return got


def test_budget_admission_by_counted_rows(base, alone):
    """Requests are admitted by their counted expectation: a batch's requests
    fit the worker's share of the budget together, a request that cannot fit
    it alone fails with GF_E_BUDGET, and the others' rows are unchanged."""
    ref, _ = alone
    mix = REQUESTS[:4] + [BIG] + REQUESTS[4:]
    roomy = serve_status(base, mix, max_batch=16, workers=1, linger_us=300000)
    assert all(res.status == 0 for res, _ in roomy.values())
    need = {rid: res.device_bytes for rid, (res, _) in roomy.items()}
    small = sorted(need[r[0]] for r in REQUESTS)
    budget = small[-1] + small[-2] // 2          # the largest small request fits, with few others
    assert need[BIG[0]] > budget and sum(small) > budget
    got = serve_status(base, mix, max_batch=16, workers=1, linger_us=300000, hbm_budget=budget)
    assert set(got) == set(need)
    assert got[BIG[0]][0].status == GF_E_BUDGET
    per_batch = {}
    for rid, (res, t) in got.items():
        if rid == BIG[0]:
            continue
        assert res.status == 0, (rid, res.status, G.lib().gf_last_error(), {k: v[0].status for k, v in got.items()},
                                 {k: v for k, v in need.items()}, budget)
        r0, t0 = ref[rid]
        assert (res.generated, res.visible, res.expected) == (r0.generated, r0.visible, r0.expected)
        assert res.device_bytes == need[rid]
        same_rows(t, t0, "request %d under a small budget" % rid)
        per_batch[res.batch] = per_batch.get(res.batch, 0) + res.device_bytes
    assert len(per_batch) > 1 and max(per_batch.values()) <= budget
    # a request that fits alone also succeeds when it shares its first batch
    # with one that does not (whichever the preparation threads finish
    # first leads the batch, so the pair is served several times)
    fit = max(REQUESTS, key=lambda r: need[r[0]])
    for _ in range(4):
        pair = serve_status(base, [fit, BIG], max_batch=16, workers=1, linger_us=300000, hbm_budget=budget)
        assert pair[BIG[0]][0].status == GF_E_BUDGET
        assert pair[fit[0]][0].status == 0, G.lib().gf_last_error()
        same_rows(pair[fit[0]][1], ref[fit[0]][1], "request %d beside an over-budget one" % fit[0])
    # the same requests with two workers share the budget between them
    got2 = serve_status(base, REQUESTS, max_batch=16, workers=2, linger_us=100000, hbm_budget=2 * budget)
    for rid, (res, t) in got2.items():
        assert res.status == 0
        same_rows(t, ref[rid][1], "request %d, two workers" % rid)


# Requests on the small config-4 base (extinction maps, unresolved binaries,
# six bands): Galactic rect tiles and a beam, (id, l, b, radius, m1, seed, l1, b1)
C4_SPEC = catalog_spec(nbands=6, nabsmag=2)
C4_REQUESTS = [
    (201, 86.0, 11.0, 0.0, 21.0, 31, 89.0, 14.0),
    (202, 90.0, 15.0, 2.0, 20.5, 32, 0.0, 0.0),
    (203, 93.0, 16.0, 0.0, 21.0, 33, 95.0, 20.0),
    (204, 358.0, -30.0, 0.0, 20.0, 34, 362.0, -27.0),    # across l = 0, south
]


@pytest.fixture(scope="module")
def c4served(tmp_path_factory):
    conf = configs.write_c4(str(tmp_path_factory.mktemp("c4s")), small=True)
    alone = serve_status(conf, C4_REQUESTS, spec=C4_SPEC, plain=True)
    mix = serve_status(conf, C4_REQUESTS, spec=C4_SPEC, max_batch=8, workers=1, linger_us=300000)
    return alone, mix


@pytest.mark.parametrize("rid", [r[0] for r in C4_REQUESTS])
def test_rect_and_extinction_requests_match_oracle_replay(tmp_path, c4served, rid):
    """A rect or beam request on a base with extinction maps: the same rows
    alone and in a batch, and the oracle's replay of a catalog job on that
    footprint (its pixels, extinction minima and draws)."""
    alone, mix = c4served
    req = next(r for r in C4_REQUESTS if r[0] == rid)
    _, l, b, rad, m1, seed, l1, b1 = req
    res, t = alone[rid]
    assert res.status == 0 and res.generated > 0
    assert max(x.batch_size for x, _ in mix.values()) > 1
    same_rows(mix[rid][1], t, "request %d in a batch" % rid, C4_SPEC, ("FeH_comp",))
    conf = configs.write_c4(str(tmp_path), small=True, m1=float(np.float32(m1)),
                            foot=configs.foot_body(l, b, rad, l1, b1))
    cat = Catalog(conf, seed=seed)
    cat.count()
    view = R.host_view(G.lib(), cat._h)
    assert res.npixels == view.sp.npixels
    ref, expected = R.replay_request(view)
    assert abs(res.expected - expected) <= 1e-12 * expected
    assert res.generated == ref.n
    same_rows(t, ref, "request %d vs oracle" % rid, C4_SPEC, ("FeH_comp",))
    assert (t.arr("Am")[0] > 0).mean() > 0.9        # the dust is there
    cat.close()


def test_request_with_a_narrowed_model_matches_oracle_replay(tmp_path_factory):
    """A model whose LF covers less than the job's absolute-magnitude range
    is binned over its own range in a request too."""
    import os
    d = str(tmp_path_factory.mktemp("c0n"))
    cmd = configs.write_c0(d)
    M, phi = configs.seeded_lf()
    keep = (M >= 5.0 - 1e-9) & (M <= 12.0 + 1e-9)
    np.savetxt(os.path.join(d, "lf_narrow.txt"), np.c_[M[keep], phi[keep]], fmt="%.10g")
    path = os.path.join(d, "model.thin.conf")
    text = open(path).read()
    open(path, "w").write(text.replace("lumfunc = lf.txt", "lumfunc = lf_narrow.txt"))
    req = (301, 90.0, 60.0, 2.0, 21.0, 41, 0.0, 0.0)
    got = serve_status(cmd, [req], plain=True)
    res, t = got[301]
    assert res.status == 0 and res.generated > 0
    configs.set_footprints(cmd, [configs.foot_body(90.0, 60.0, 2.0)])
    cat = Catalog(cmd, seed=41)
    cat.count()
    view = R.host_view(G.lib(), cat._h)
    assert R.model_sp(view, 0).nM < view.sp.nM
    ref, expected = R.replay_request(view)
    assert abs(res.expected - expected) <= 1e-12 * expected
    assert res.generated == ref.n
    same_rows(t, ref, "narrowed request vs oracle")
    cat.close()
