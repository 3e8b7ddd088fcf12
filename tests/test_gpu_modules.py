"""Per-star module kernels on the GPU against the oracle, and the fused chain
against the unfused sequence."""
import ctypes as C
import copy

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import _abi as A
from galfast_amd.table import DeviceTable, HostTable
import oracle as O
from devutil import DevRng, dev_array, stream_ptr, sync
import gfx

pytestmark = pytest.mark.gpu

N = 300_000     # 1172 chunks of 256 rows
SEED = 77


class Fixture:
    def __init__(self, n=N, nbands=5):
        self.host = gfx.synthetic_stars(n, seed=3, nbands=nbands)
        self.dev = DeviceTable(n, self.host.spec)
        self.dev.upload(self.host)
        self.state = gfx.mwc_state(SEED, 65536)
        self.drng = DevRng(self.state)
        self.n = n
        self.nbands = nbands
        # photometry tables shared by device and oracle
        rows = gfx.synthetic_photosys(ncolors=nbands - 1)
        ntex, nF, nM, tc, colors, flags = gfx.photometry_tables(rows, nbands - 1)
        sz = nF * nM * 4
        self.phot_host_arrays = [colors[i * sz:(i + 1) * sz] for i in range(ntex)] + \
                                [flags[i * sz:(i + 1) * sz] for i in range(ntex)]
        self.phot_dev_arrays = [dev_array(a) for a in self.phot_host_arrays]
        red = [1.8739, 1.3788, 1.0, 0.7583, 0.5376, 0.4]
        self.phot_host = gfx.photometry_params(ntex, nF, nM, tc,
                                               [a.ctypes.data for a in self.phot_host_arrays[:ntex]],
                                               [a.ctypes.data for a in self.phot_host_arrays[ntex:]],
                                               nbands - 1, 2, red)
        self.phot_dev = gfx.photometry_params(ntex, nF, nM, tc,
                                              [t.data_ptr() for t in self.phot_dev_arrays[:ntex]],
                                              [t.data_ptr() for t in self.phot_dev_arrays[ntex:]],
                                              nbands - 1, 2, red)
        M, phi = gfx.seeded_lf()
        (sec, sec_tc), (cum, cum_tc), (inv, inv_tc) = gfx.multiples_tables(M, phi)
        self.mult_host_arrays = [sec, cum, inv]
        self.mult_dev_arrays = [dev_array(a) for a in self.mult_host_arrays]
        tcs = [sec_tc, cum_tc, inv_tc]

        def mult(ptrs, algo):
            p = A.gf_multiples_data()
            p.secProb = gfx.tex1d(ptrs[0], len(sec), tcs[0])
            p.cumLF = gfx.tex1d(ptrs[1], len(cum), tcs[1])
            p.invCumLF = gfx.tex1d(ptrs[2], len(inv), tcs[2])
            p.algo = algo
            return p
        self.mult = lambda dev, algo: mult([t.data_ptr() for t in self.mult_dev_arrays] if dev else
                                           [a.ctypes.data for a in self.mult_host_arrays], algo)
        kx, ky, koff = gfx.photoerr_knots(nbands)
        self.pe_host_arrays = [kx, ky]
        self.pe_dev_arrays = [dev_array(kx), dev_array(ky)]

        def perr(ptrs):
            p = A.gf_photoerr_data()
            p.nbands = nbands
            for i, k in enumerate(koff):
                p.koff[i] = k
            p.kx, p.ky = ptrs
            return p
        self.perr = lambda dev: perr([t.data_ptr() for t in self.pe_dev_arrays] if dev else
                                     [a.ctypes.data for a in self.pe_host_arrays])

    def rows(self):
        return A.gf_rows(0, 256, self.n, 0)

    def oracle_chain(self, **mods):
        ch = A.gf_chain()
        for k, v in mods.items():
            if k in ("gfeh_comps", "mult_comps", "phot_comps", "gal2other_coordsys"):
                setattr(ch, k, v)
            else:
                setattr(ch, k, C.pointer(v))
        t = self.host.table()
        rc = O.lib().gfo_module_chain(C.byref(self.rows()), C.byref(ch), O.ptr(self.state), 65536, C.byref(t))
        assert rc == 0

    def compare(self, cols, exact=True, rtol=0.0, atol=0.0):
        sync()
        got = self.dev.to_host()
        vis = self.host.arr("hidden")[0] == 0
        for c in cols:
            a, b = got.arr(c)[:, vis], self.host.arr(c)[:, vis]
            if exact:
                bad = (a != b) & ~(np.isnan(a) & np.isnan(b))
                assert not bad.any(), "%s: %d of %d differ, e.g. %r vs %r" % (
                    c, bad.sum(), bad.size, a[bad][:3], b[bad][:3])
            else:
                np.testing.assert_allclose(a, b, rtol=rtol, atol=atol, err_msg=c)
        assert np.array_equal(self.drng.state(), self.state), "stream states differ"


@pytest.fixture
def fx():
    return Fixture()


def test_FeH(fx):
    p = gfx.feh_params()
    G.check(G.lib().gf_os_FeH_kernel(C.byref(fx.rows()), C.byref(p), C.byref(fx.drng.rng), fx.dev.col("comp"),
                                     fx.dev.col("hidden"), fx.dev.col("XYZ"), fx.dev.col("FeH"), stream_ptr()))
    fx.oracle_chain(feh=p)
    fx.compare(["FeH"])


def test_GaussianFeH(fx):
    p = A.gf_GaussianFeH_data()
    p.ncomp = 3
    for i, (m, s, w) in enumerate([(-0.5, 0.2, 0.6), (-1.0, 0.3, 0.3), (-1.5, 0.4, 0.1)]):
        p.mean[i], p.sigma[i] = m, s
    cs = np.cumsum([0.6, 0.3, 0.1])
    for i in range(3):
        p.pcum[i] = cs[i] / cs[-1]
    comps = A.bitmap(0, 1, 2)
    G.check(G.lib().gf_os_GaussianFeH_kernel(C.byref(fx.rows()), comps, C.byref(p), C.byref(fx.drng.rng),
                                             fx.dev.col("comp"), fx.dev.col("hidden"), fx.dev.col("XYZ"),
                                             fx.dev.col("FeH_comp"), fx.dev.col("FeH"), stream_ptr()))
    fx.oracle_chain(gfeh=p, gfeh_comps=comps)
    fx.compare(["FeH", "FeH_comp"])


def test_Bond2010(fx):
    p = gfx.bond_params()
    G.check(G.lib().gf_os_Bond2010_kernel(C.byref(fx.rows()), C.byref(p), C.byref(fx.drng.rng), fx.dev.col("comp"),
                                          fx.dev.col("hidden"), fx.dev.col("XYZ"), fx.dev.col("vcyl"), stream_ptr()))
    fx.oracle_chain(bond=p)
    fx.compare(["vcyl"])


@pytest.mark.parametrize("coordsys", [0, 1])
def test_vel2pm(fx, coordsys):
    p = gfx.vel2pm_params(coordsys)
    G.check(G.lib().gf_os_vel2pm_kernel(C.byref(fx.rows()), C.byref(p), fx.dev.col("lb"), fx.dev.col("XYZ"),
                                        fx.dev.col("vcyl"), fx.dev.col("pm"), fx.dev.col("hidden"), stream_ptr()))
    fx.oracle_chain(v2pm=p)
    fx.compare(["pm"])


def test_gal2other(fx):
    G.check(G.lib().gf_os_gal2other_kernel(C.byref(fx.rows()), 1, fx.dev.col("hidden"), fx.dev.col("lb"),
                                           fx.dev.col("radec"), stream_ptr()))
    fx.oracle_chain(gal2other_coordsys=1)
    fx.compare(["radec"])


def test_photometry(fx):
    comps = A.bitmap(0, 1, 2)
    Mr = fx.dev.col("absmag")
    G.check(G.lib().gf_os_photometry_kernel(C.byref(fx.rows()), comps, C.byref(fx.phot_dev), fx.dev.col("Am"),
                                            fx.dev.col("photoFlags"), fx.dev.col("DM"), Mr, 1, fx.dev.col("mags"),
                                            fx.dev.col("FeH"), fx.dev.col("comp"), fx.dev.col("hidden"),
                                            stream_ptr()))
    fx.oracle_chain(phot=fx.phot_host, phot_comps=comps)
    fx.compare(["mags", "photoFlags"])


def quad_layout(colors, flags, ntex, nF, nM):
    """The catalog's lookup-quad copy of the colour and flag texels: per
    lower corner (i, j) the four corners' float4s of every table and each
    corner's flag bits (bit 4t+c for colour 4t+c)."""
    sz = nF * nM * 4
    col = np.stack([colors[t * sz:(t + 1) * sz].reshape(nM, nF, 4) for t in range(ntex)])   # t, j, i, c
    flg = np.stack([flags[t * sz:(t + 1) * sz].reshape(nM, nF, 4) for t in range(ntex)])
    bits = np.zeros((nM, nF), dtype=np.uint16)
    for t in range(ntex):
        for c in range(4):
            bits |= ((flg[t, :, :, c] != 0).astype(np.uint16) << np.uint16(4 * t + c))
    j, i = np.meshgrid(np.arange(nM), np.arange(nF), indexing="ij")
    i1, j1 = np.minimum(i + 1, nF - 1), np.minimum(j + 1, nM - 1)
    corners = [(j, i), (j, i1), (j1, i), (j1, i1)]
    qd = np.stack([np.stack([col[t][cj, ci] for cj, ci in corners], axis=2) for t in range(ntex)], axis=2)
    fb = np.stack([bits[cj, ci] for cj, ci in corners], axis=2)
    return np.ascontiguousarray(qd, dtype=np.float32).ravel(), np.ascontiguousarray(fb).ravel()


@pytest.mark.parametrize("nbands,system", [(5, False), (7, False), (7, True)])
def test_photometry_quad_layout(nbands, system):
    """The photometry kernel reading the lookup-quad copy gives the oracle's
    bits (the oracle samples the reference's separate texel tables), for one
    and two colour tables, lone stars and binaries."""
    fx = Fixture(nbands=nbands)
    rows = gfx.synthetic_photosys(ncolors=nbands - 1)
    ntex, nF, nM, tc, colors, flags = gfx.photometry_tables(rows, nbands - 1)
    qd, fb = quad_layout(colors, flags, ntex, nF, nM)
    dq, dfb = dev_array(qd), dev_array(fb)
    pd = fx.phot_dev      # this test's own fixture
    pd.texels, pd.flagbits = dq.data_ptr(), dfb.data_ptr()
    comps = A.bitmap(0, 1, 2)
    if system:
        md, mh = fx.mult(True, 1), fx.mult(False, 1)
        G.check(G.lib().gf_os_unresolvedMultiples_kernel(C.byref(fx.rows()), comps, C.byref(md),
                                                         C.byref(fx.drng.rng), 2, fx.dev.col("absmag"),
                                                         fx.dev.col("absmagSys"), fx.dev.col("absmagSysNcomp"),
                                                         fx.dev.col("comp"), fx.dev.col("hidden"), stream_ptr()))
    Mr, nabs = (fx.dev.col("absmagSys"), 2) if system else (fx.dev.col("absmag"), 1)
    G.check(G.lib().gf_os_photometry_kernel(C.byref(fx.rows()), comps, C.byref(pd), fx.dev.col("Am"),
                                            fx.dev.col("photoFlags"), fx.dev.col("DM"), Mr, nabs,
                                            fx.dev.col("mags"), fx.dev.col("FeH"), fx.dev.col("comp"),
                                            fx.dev.col("hidden"), stream_ptr()))
    if system:
        fx.oracle_chain(mult=mh, mult_comps=comps, phot=fx.phot_host, phot_comps=comps)
        fx.compare(["absmagSys", "mags", "photoFlags"])
    else:
        fx.oracle_chain(phot=fx.phot_host, phot_comps=comps)
        fx.compare(["mags", "photoFlags"])


@pytest.mark.parametrize("algo", [1, 2, 3])
def test_unresolvedMultiples(fx, algo):
    comps = A.bitmap(0, 1, 2)
    pd, ph = fx.mult(True, algo), fx.mult(False, algo)
    G.check(G.lib().gf_os_unresolvedMultiples_kernel(C.byref(fx.rows()), comps, C.byref(pd), C.byref(fx.drng.rng),
                                                     2, fx.dev.col("absmag"), fx.dev.col("absmagSys"),
                                                     fx.dev.col("absmagSysNcomp"), fx.dev.col("comp"),
                                                     fx.dev.col("hidden"), stream_ptr()))
    fx.oracle_chain(mult=ph, mult_comps=comps)
    fx.compare(["absmag", "absmagSys", "absmagSysNcomp"])


def test_photometric_errors(fx):
    fx.dev.buf["mags"][0].copy_(__import__("torch").from_numpy(
        np.random.default_rng(1).uniform(14, 26, fx.dev.buf["mags"][0].numel()).astype(np.float32)))
    fx.host.buf["mags"][0][:] = fx.dev.buf["mags"][0].cpu().numpy()
    pd, ph = fx.perr(True), fx.perr(False)
    G.check(G.lib().gf_photometric_errors_kernel(C.byref(fx.rows()), C.byref(pd), C.byref(fx.drng.rng),
                                                 fx.dev.col("mags"), fx.dev.col("obsMags"), fx.dev.col("hidden"),
                                                 stream_ptr()))
    fx.oracle_chain(perr=ph)
    fx.compare(["obsMags"])


def _irregular_knots(nbands, rng):
    """Error curves whose knots the interval guess often misses: irregular
    spacing, one band of a single knot, one of two, one without a curve."""
    kx, ky, koff = [], [], [0]
    for b in range(nbands):
        if b == 2:
            m = np.array([20.0])
        elif b == 3:
            m = np.array([12.0, 27.0])
        elif b == nbands - 1:
            m = np.zeros(0)
        else:
            m = np.sort(rng.uniform(10.0, 30.0, 300 + 100 * b))
            m[0], m[-1] = 10.0, 30.0
        kx.append(m)
        ky.append(0.02 + np.exp(0.8 * (m - 24.0)))
        koff.append(koff[-1] + len(m))
    return np.concatenate(kx), np.concatenate(ky), koff


@pytest.mark.parametrize("irregular", [False, True])
def test_photometric_errors_interval_records(fx, irregular):
    """The kernel with gf_photoerr_records' interval records gives the bits
    of the evaluation from the knots (the oracle's)."""
    import torch
    rng = np.random.default_rng(7)
    mags = rng.uniform(8, 32, fx.dev.buf["mags"][0].numel()).astype(np.float32)
    fx.dev.buf["mags"][0].copy_(torch.from_numpy(mags))
    fx.host.buf["mags"][0][:] = mags
    nb = fx.perr(False).nbands
    if irregular:
        kx, ky, koff = _irregular_knots(nb, rng)
    else:
        kx, ky, koff = gfx.photoerr_knots(nb)
    koffa = np.array(koff, dtype=np.int32)
    kq, gx0, gs = np.zeros(4 * len(kx)), np.zeros(18), np.zeros(18)
    G.check(G.lib().gf_photoerr_records(nb, koffa.ctypes.data, kx.ctypes.data, ky.ctypes.data, kq.ctypes.data,
                                        gx0.ctypes.data, gs.ctypes.data), "gf_photoerr_records")
    dx, dy, dq = dev_array(kx), dev_array(ky), dev_array(kq)

    def par(ptrs):
        p = A.gf_photoerr_data()
        p.nbands = nb
        for i, k in enumerate(koff):
            p.koff[i] = k
        p.kx, p.ky = ptrs[0], ptrs[1]
        if len(ptrs) > 2:
            p.kq = ptrs[2]
            for b in range(nb):
                p.gx0[b], p.gscale[b] = gx0[b], gs[b]
        return p
    pd = par([dx.data_ptr(), dy.data_ptr(), dq.data_ptr()])
    ph = par([kx.ctypes.data, ky.ctypes.data])
    G.check(G.lib().gf_photometric_errors_kernel(C.byref(fx.rows()), C.byref(pd), C.byref(fx.drng.rng),
                                                 fx.dev.col("mags"), fx.dev.col("obsMags"), fx.dev.col("hidden"),
                                                 stream_ptr()))
    fx.oracle_chain(perr=ph)
    fx.compare(["obsMags"])


def test_stream_limit_is_enforced(fx):
    p = gfx.feh_params()
    rows = A.gf_rows(0, 256, 65536 * 256 + 1, 0)
    rc = G.lib().gf_os_FeH_kernel(C.byref(rows), C.byref(p), C.byref(fx.drng.rng), fx.dev.col("comp"),
                                  fx.dev.col("hidden"), fx.dev.col("XYZ"), fx.dev.col("FeH"), stream_ptr())
    assert rc == -3


def test_fused_chain_equals_unfused_and_oracle():
    fx = Fixture()
    feh, bond, v2pm = gfx.feh_params(), gfx.bond_params(), gfx.vel2pm_params(1)
    comps = A.bitmap(0, 1, 2)
    pe_d, pe_h = fx.perr(True), fx.perr(False)
    ch = A.gf_chain()
    ch.feh, ch.bond, ch.v2pm = C.pointer(feh), C.pointer(bond), C.pointer(v2pm)
    ch.phot, ch.phot_comps = C.pointer(fx.phot_dev), comps
    ch.perr = C.pointer(pe_d)
    ch.gal2other_coordsys = 1
    t = fx.dev.table()
    G.check(G.lib().gf_module_chain(C.byref(fx.rows()), C.byref(ch), C.byref(fx.drng.rng), C.byref(t),
                                    stream_ptr()), "gf_module_chain")
    fx.oracle_chain(feh=feh, bond=bond, phot=fx.phot_host, phot_comps=comps, perr=pe_h, v2pm=v2pm,
                    gal2other_coordsys=1)
    fx.compare(["FeH", "vcyl", "mags", "photoFlags", "obsMags", "pm", "radec"])


def _gfeh_params():
    p = A.gf_GaussianFeH_data()
    p.ncomp = 3
    for i, (m, s, w) in enumerate([(-0.5, 0.2, 0.6), (-1.0, 0.3, 0.3), (-1.5, 0.4, 0.1)]):
        p.mean[i], p.sigma[i] = m, s
    cs = np.cumsum([0.6, 0.3, 0.1])
    for i in range(3):
        p.pcum[i] = cs[i] / cs[-1]
    return p


@pytest.mark.parametrize("mult_comps", [(0, 1, 2), (0, 2)])
def test_fused_chain_gaussianfeh_multiples_equals_oracle(mult_comps):
    """GaussianFeH + unresolvedMultiples + Bond2010 + photometric errors (the
    config-4 kind of chain) in one fused chain against the oracle.  With 5 %
    hidden rows some 16-row tiles of the draw pass have every row taking part
    in unresolvedMultiples and some not (and, with component 1 left out,
    most not), so both of its store paths run."""
    L = G.lib()
    fx = Fixture()
    comps = A.bitmap(0, 1, 2)
    mcomps = A.bitmap(*mult_comps)
    gfeh, bond = _gfeh_params(), gfx.bond_params()
    md, mh = fx.mult(True, 1), fx.mult(False, 1)
    pe_d, pe_h = fx.perr(True), fx.perr(False)
    ch = A.gf_chain()
    ch.gfeh, ch.gfeh_comps = C.pointer(gfeh), comps
    ch.mult, ch.mult_comps = C.pointer(md), mcomps
    ch.bond = C.pointer(bond)
    ch.perr = C.pointer(pe_d)
    G.check(L.gf_module_chain(C.byref(fx.rows()), C.byref(ch), C.byref(fx.drng.rng), C.byref(fx.dev.table()),
                              stream_ptr()), "gf_module_chain")
    fx.oracle_chain(gfeh=gfeh, gfeh_comps=comps, mult=mh, mult_comps=mcomps, bond=bond, perr=pe_h)
    fx.compare(["FeH", "FeH_comp", "absmag", "absmagSys", "absmagSysNcomp", "vcyl", "obsMags"])
