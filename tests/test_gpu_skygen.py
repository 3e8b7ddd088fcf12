"""Sky generator on the GPU: count integral, reference-order draw and the
cumulative-table draw, against the oracle (config-0 sized sky)."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

import galfast_amd as G
from galfast_amd import _abi as A
from galfast_amd.table import DeviceTable, HostTable
import oracle as O
from devutil import DevRng, dev_array, stream_ptr, sync, torch
import gfx

pytestmark = pytest.mark.gpu


class Sky:
    """Config-0 geometry: l 85-95, b 55-65 at 1 degree, r 15-21, Mr 3-15, dm 0.05."""

    def __init__(self, dx=1.0, m=(15.0, 21.0), nthreads=7680):
        self.pix = gfx.pixelize_rect(85, 95, 55, 65, dx)
        self.np = len(self.pix)
        self.sp = gfx.skygen_params(m[0], m[1], 0.05, 3.0, 15.0, self.np, nthreads)
        M, phi = gfx.seeded_lf()
        self.lf, self.lf_tc = gfx.lf_texture(M, phi)
        self.models_h = gfx.c0_models(self.lf, self.lf_tc)
        self.lf_dev = dev_array(self.lf)
        self.models_d = []
        for m in self.models_h:
            md = A.gf_model.from_buffer_copy(m)
            md.lf.data = self.lf_dev.data_ptr()
            self.models_d.append(md)
        T = torch()
        self.pix_dev = T.from_numpy(np.frombuffer(bytes(self.pix), dtype=np.uint8).copy()).cuda()
        self.foot = gfx.rect_footprint(85, 95, 55, 65)
        self.K = self.sp.nm + self.sp.nM - 1


@pytest.fixture(scope="module")
def sky():
    return Sky()


def count_dev(sky, mi, mode=A.GF_ACC_FP64, want_w=True):
    T = torch()
    res = A.gf_count_result()
    per = T.zeros(sky.np, dtype=T.float64, device="cuda")
    w = T.zeros(sky.np * sky.K, dtype=T.float64, device="cuda") if want_w else None
    G.check(G.lib().gf_skygen_count(C.byref(sky.sp), C.byref(sky.models_d[mi]), C.c_void_p(sky.pix_dev.data_ptr()),
                                    None, mode, C.byref(res), C.c_void_p(per.data_ptr()),
                                    C.c_void_p(w.data_ptr()) if w is not None else None, stream_ptr()),
            "gf_skygen_count")
    return res, per.cpu().numpy(), (w.cpu().numpy() if w is not None else None)


@pytest.mark.parametrize("mi", [0, 1, 2])
def test_count_matches_oracle(sky, mi):
    res, per, w = count_dev(sky, mi)
    ref = A.gf_count_result()
    rper = np.zeros(sky.np)
    rw = np.zeros(sky.np * sky.K)
    O.lib().gfo_count_bins(C.byref(sky.sp), C.byref(sky.models_h[mi]), C.cast(sky.pix, C.c_void_p), None,
                           C.byref(ref), O.ptr(rper), O.ptr(rw))
    assert res.nstarsExpectedToGenerate == pytest.approx(ref.nstarsExpectedToGenerate, rel=1e-12)
    assert res.nstarsExpected == pytest.approx(ref.nstarsExpected, rel=1e-12)
    assert res.ncells == ref.ncells
    assert np.array_equal(w, rw)
    np.testing.assert_allclose(per, rper, rtol=1e-12, atol=0)
    # the reference's own cell-by-cell float expectations summed in the
    # crawl order of kernel<0> agree to the stated 1e-6
    crawl = A.gf_count_result()
    O.lib().gfo_count_crawl(C.byref(sky.sp), C.byref(sky.models_h[mi]), C.cast(sky.pix, C.c_void_p), None,
                            C.byref(crawl), None)
    assert crawl.ncells == res.ncells
    assert res.nstarsExpectedToGenerate == pytest.approx(crawl.nstarsExpectedToGenerate, rel=1e-6)
    assert res.nstarsExpected == pytest.approx(crawl.nstarsExpected, rel=1e-6)


def test_count_precision_variants(sky):
    r64, _, _ = count_dev(sky, 0, A.GF_ACC_FP64, False)
    rk, _, _ = count_dev(sky, 0, A.GF_ACC_KAHAN32, False)
    rf, _, _ = count_dev(sky, 0, A.GF_ACC_FP32, False)
    assert rk.nstarsExpectedToGenerate == pytest.approx(r64.nstarsExpectedToGenerate, rel=1e-5)
    assert rf.nstarsExpectedToGenerate == pytest.approx(r64.nstarsExpectedToGenerate, rel=1e-4)


def _cols_equal(a, b, n):
    for c in ("hidden", "projIdx", "projXY", "lb", "DM", "absmag", "XYZ", "comp", "Am", "AmInf"):
        bad = a.arr(c)[:, :n] != b.arr(c)[:, :n]
        assert not bad.any(), "%s: %d of %d differ" % (c, bad.sum(), bad.size)


@pytest.mark.parametrize("mi", [0, 2])
def test_compat_draw_bit_exact(sky, mi):
    state = gfx.mwc_state(42, 65536)
    drng = DevRng(state)
    # a capacity that is too small returns the needed size and leaves the streams alone
    nst = C.c_uint64()
    small = DeviceTable(16)
    scols = small.skygen_cols()
    rc = G.lib().gf_skygen_draw_compat(C.byref(sky.sp), C.byref(sky.models_d[mi]), C.c_void_p(sky.pix_dev.data_ptr()),
                                       None, C.byref(sky.foot), C.byref(drng.rng), C.byref(scols), 16, C.byref(nst),
                                       stream_ptr())
    assert rc == -2 and nst.value > 16
    assert np.array_equal(drng.state(), state)
    cap = nst.value
    dev = DeviceTable(cap)
    cols = dev.skygen_cols()
    G.check(G.lib().gf_skygen_draw_compat(C.byref(sky.sp), C.byref(sky.models_d[mi]),
                                          C.c_void_p(sky.pix_dev.data_ptr()), None, C.byref(sky.foot),
                                          C.byref(drng.rng), C.byref(cols), cap, C.byref(nst), stream_ptr()),
            "gf_skygen_draw_compat")
    host = HostTable(cap)
    hcols = host.skygen_cols()
    n = O.lib().gfo_draw_compat(C.byref(sky.sp), C.byref(sky.models_h[mi]), C.cast(sky.pix, C.c_void_p), None,
                                C.byref(sky.foot), O.ptr(state), 65536, C.byref(hcols), cap)
    assert n == nst.value and n > 1000
    _cols_equal(dev.to_host(), host, n)
    assert np.array_equal(drng.state(), state)


def test_cumulative_table_draw_bit_exact_and_chunked(sky):
    mi = 0
    _, _, w = count_dev(sky, mi)
    T = torch()
    wd = T.from_numpy(w).cuda()
    state = gfx.mwc_state(7, 65536)
    drng = DevRng(state)
    plan = C.c_void_p()
    nst = C.c_uint64()
    G.check(G.lib().gf_skygen_plan(C.byref(sky.sp), C.byref(sky.models_d[mi]), C.c_void_p(sky.pix_dev.data_ptr()),
                                   None, C.c_void_p(wd.data_ptr()), C.byref(drng.rng), C.byref(plan), C.byref(nst),
                                   stream_ptr()), "gf_skygen_plan")
    U = sky.np * sky.K
    counts = np.zeros(U, dtype=np.uint32)
    O.lib().gfo_plan(O.ptr(w), U, O.ptr(state), 65536, O.ptr(counts))
    adv = state.copy()
    O.lib().gfo_plan_advance(O.ptr(adv), 65536, U)
    total = int(counts.sum())
    assert total == nst.value
    # the plan moves the streams past every unit; the emission leaves them
    assert np.array_equal(drng.state(), adv)
    offs = np.zeros(U + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(counts)
    # emit in two chunks on the device, one on the oracle
    dev = DeviceTable(total)
    cut = total // 3
    cols = dev.skygen_cols()
    G.check(G.lib().gf_skygen_emit(plan, 0, cut, C.byref(sky.foot), C.byref(drng.rng), C.byref(cols), stream_ptr()))
    dev2 = DeviceTable(total - cut)
    cols2 = dev2.skygen_cols()
    G.check(G.lib().gf_skygen_emit(plan, cut, total, C.byref(sky.foot), C.byref(drng.rng), C.byref(cols2),
                                   stream_ptr()))
    host = HostTable(total)
    hcols = host.skygen_cols()
    O.lib().gfo_emit(C.byref(sky.sp), C.byref(sky.models_h[mi]), C.cast(sky.pix, C.c_void_p), None,
                     C.byref(sky.foot), O.ptr(offs), U, 0, total, O.ptr(state), 65536, C.byref(hcols))
    a, b = dev.to_host(), dev2.to_host()
    joined = HostTable(total)
    for c in joined.buf:
        joined.arr(c)[:, :cut] = a.arr(c)[:, :cut]
        joined.arr(c)[:, cut:] = b.arr(c)[:, : total - cut]
    _cols_equal(joined, host, total)
    assert np.array_equal(drng.state(), adv)
    G.lib().gf_skygen_plan_free(plan)


def test_draw_totals_are_poisson_consistent(sky):
    # (generated - expected)/sqrt(expected), as skygenHost::drawSources reports
    for mi in range(3):
        res, _, w = count_dev(sky, mi)
        T = torch()
        wd = T.from_numpy(w).cuda()
        drng = DevRng(gfx.mwc_state(100 + mi, 65536))
        plan, nst = C.c_void_p(), C.c_uint64()
        G.check(G.lib().gf_skygen_plan(C.byref(sky.sp), C.byref(sky.models_d[mi]),
                                       C.c_void_p(sky.pix_dev.data_ptr()), None, C.c_void_p(wd.data_ptr()),
                                       C.byref(drng.rng), C.byref(plan), C.byref(nst), stream_ptr()))
        G.lib().gf_skygen_plan_free(plan)
        exp = res.nstarsExpectedToGenerate
        assert abs(nst.value - exp) / np.sqrt(exp) < 5


def test_cumulative_draw_matches_reference_order_draw_statistically(sky):
    """The per-(pixel, DM-bin) cumulative-table draw and the reference's
    per-cell Poisson draw sample the same distribution (Poisson
    superposition): KS on m, M, DM and l, b."""
    mi = 0
    res, _, w = count_dev(sky, mi)
    T = torch()
    wd = T.from_numpy(w).cuda()
    drng = DevRng(gfx.mwc_state(5, 65536))
    plan, nst = C.c_void_p(), C.c_uint64()
    G.check(G.lib().gf_skygen_plan(C.byref(sky.sp), C.byref(sky.models_d[mi]), C.c_void_p(sky.pix_dev.data_ptr()),
                                   None, C.c_void_p(wd.data_ptr()), C.byref(drng.rng), C.byref(plan), C.byref(nst),
                                   stream_ptr()))
    total = nst.value
    a = DeviceTable(total)
    cols = a.skygen_cols()
    G.check(G.lib().gf_skygen_emit(plan, 0, total, None, C.byref(drng.rng), C.byref(cols), stream_ptr()))
    G.lib().gf_skygen_plan_free(plan)
    cap = 3 * total + 1000
    b = DeviceTable(cap)
    bcols = b.skygen_cols()
    nb = C.c_uint64()
    drng2 = DevRng(gfx.mwc_state(6, 65536))
    G.check(G.lib().gf_skygen_draw_compat(C.byref(sky.sp), C.byref(sky.models_d[mi]),
                                          C.c_void_p(sky.pix_dev.data_ptr()), None, None, C.byref(drng2.rng),
                                          C.byref(bcols), cap, C.byref(nb), stream_ptr()))
    ha, hb = a.to_host(), b.to_host()
    na, nbv = total, nb.value
    assert abs(na - nbv) < 6 * np.sqrt(res.nstarsExpectedToGenerate)
    va = ha.arr("hidden")[0, :na] == 0
    vb = hb.arr("hidden")[0, :nbv] == 0
    for name, f in (("m", lambda h, n: h.arr("DM")[0, :n] + h.arr("absmag")[0, :n]),
                    ("M", lambda h, n: h.arr("absmag")[0, :n]),
                    ("DM", lambda h, n: h.arr("DM")[0, :n]),
                    ("l", lambda h, n: h.arr("lb")[0, :n]),
                    ("b", lambda h, n: h.arr("lb")[1, :n])):
        p = stats.ks_2samp(f(ha, na)[va], f(hb, nbv)[vb]).pvalue
        assert p > 1e-4, (name, p)
