"""Pins of the oracle's random-number modules against the distributions the
reference's formulas fix (no reference test holds their outputs).

For fixed inputs each module's output is a known distribution: FeH
(modules/FeH_gpu.cu.h:55-86) a two-Gaussian mixture around the
height-dependent disk mean or one Gaussian for the halo; GaussianFeH
(modules/GaussianFeH_gpu.cu.h:55-72) the configured mixture; Bond2010
(modules/Bond2010_gpu.cu.h:106-216) a trivariate Gaussian whose moments are
the modfun polynomials, rotated for the halo; unresolvedMultiples
(modules/unresolvedMultiples_gpu.cu.h:80-123) a companion fraction and a
luminosity sum; photometric errors a Gaussian of the tabulated sigma(m).
Each test checks the oracle's samples against that distribution and also
checks that the same test rejects a plausible mistake (a wrong sign, a
swapped weight or index), so it has the power to catch one."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O
import gfx

N = 120_000
P_OK = 1e-4       # a correct model passes the KS test above this p-value
P_BAD = 1e-8      # a planted mistake must fall below this one


def table(n, comp, xyz=(300.0, -200.0, 900.0), nbands=5, absmag=None, mags=None):
    t = HostTable(n, catalog_spec(nbands=nbands, nabsmag=2))
    t.arr("comp")[0] = comp
    t.arr("hidden")[0] = 0
    for i in range(3):
        t.arr("XYZ")[i] = xyz[i]
    if absmag is not None:
        t.arr("absmag")[0] = absmag
    if mags is not None:
        t.arr("mags")[:] = mags
    return t


def run(t, seed=5, **mods):
    ch = A.gf_chain()
    for k, v in mods.items():
        if k.endswith("_comps") or k == "gal2other_coordsys":
            setattr(ch, k, v)
        else:
            setattr(ch, k, C.pointer(v))
    st = gfx.mwc_state(seed)
    rows = A.gf_rows(0, 256, t.n, 0)
    assert O.lib().gfo_module_chain(C.byref(rows), C.byref(ch), O.ptr(st), 65536, C.byref(t.table())) == 0
    return t


def transformed(p, xyz):
    v = np.array(xyz, dtype=np.float64) + np.array(p.T[:], dtype=np.float64)
    M = np.array([p.M[i][:] for i in range(3)], dtype=np.float64)
    return M @ v


def mixture_cdf(means, sigmas, weights):
    w = np.asarray(weights, dtype=np.float64)
    w = w / w.sum()
    return lambda x: sum(wi * stats.norm.cdf(x, m, s) for wi, m, s in zip(w, means, sigmas))


def ks_p(x, cdf):
    return stats.kstest(x, cdf).pvalue


# ---------------------------------------------------------------------------
# FeH (Ivezic et al. 2008 as in the reference)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("comp", [0, 1])
def test_feh_disk_is_the_height_dependent_mixture(comp):
    p = gfx.feh_params()
    xyz = (300.0, -200.0, 900.0)
    t = run(table(N, comp, xyz), feh=p)
    feh = t.arr("FeH")[0].astype(np.float64)
    z = transformed(p, xyz)[2]
    aZ = p.muInf + p.DeltaMu * np.exp(-abs(z) / p.Hmu) - 0.067
    means = [aZ + p.offs[0], aZ + p.offs[1]]
    good = mixture_cdf(means, [p.sigma[0], p.sigma[1]], [p.A[0], p.A[1]])
    assert ks_p(feh, good) > P_OK
    # planted mistakes: weights swapped; the height term with the wrong sign
    assert ks_p(feh, mixture_cdf(means, [p.sigma[0], p.sigma[1]], [p.A[1], p.A[0]])) < P_BAD
    aZbad = p.muInf - p.DeltaMu * np.exp(-abs(z) / p.Hmu) - 0.067
    assert ks_p(feh, mixture_cdf([aZbad + p.offs[0], aZbad + p.offs[1]], [0.2, 0.2], [p.A[0], p.A[1]])) < P_BAD


def test_feh_halo_is_one_gaussian_and_other_components_untouched():
    p = gfx.feh_params()
    t = table(N, 2)
    t.arr("comp")[0][: N // 10] = 5          # a component FeH does not handle
    t.arr("FeH")[0][:] = 99.0
    run(t, feh=p)
    feh = t.arr("FeH")[0]
    assert np.all(feh[: N // 10] == 99.0)
    x = feh[N // 10:].astype(np.float64)
    assert ks_p(x, stats.norm(p.offs[2], p.sigma[2]).cdf) > P_OK
    assert ks_p(x, stats.norm(p.offs[2], p.sigma[0]).cdf) < P_BAD


# ---------------------------------------------------------------------------
# GaussianFeH
# ---------------------------------------------------------------------------
def test_gaussian_feh_mixture_and_component_index():
    p = A.gf_GaussianFeH_data()
    p.ncomp = 3
    mix = [(-0.5, 0.2, 0.6), (-1.0, 0.3, 0.3), (-1.6, 0.4, 0.1)]
    cs = np.cumsum([w for _, _, w in mix])
    for i, (m, s, _) in enumerate(mix):
        p.mean[i], p.sigma[i], p.pcum[i] = m, s, cs[i] / cs[-1]
    t = run(table(N, 1), gfeh=p, gfeh_comps=A.bitmap(1))
    feh = t.arr("FeH")[0].astype(np.float64)
    at = t.arr("FeH_comp")[0]
    counts = np.bincount(at, minlength=3)
    w = np.array([w for _, _, w in mix])
    assert stats.chisquare(counts, w / w.sum() * N).pvalue > P_OK
    # each component's draws are its own Gaussian
    for i, (m, s, _) in enumerate(mix):
        assert ks_p(feh[at == i], stats.norm(m, s).cdf) > P_OK
    # planted mistake: component index off by one
    assert ks_p(feh[at == 1], stats.norm(mix[0][0], mix[0][1]).cdf) < P_BAD


# ---------------------------------------------------------------------------
# Bond2010 kinematics
# ---------------------------------------------------------------------------
def modfun(R2, Z, v):
    return v[0] + v[1] * abs(Z) ** v[2] + v[3] * R2 ** (0.5 * v[4])


def test_bond_disk_moments_and_marginals():
    p = gfx.bond_params()
    xyz = (300.0, -200.0, 900.0)
    t = run(table(N, 0, xyz), bond=p)
    v = t.arr("vcyl").astype(np.float64)
    x, y, z = transformed(p, xyz)
    R2, Z = 1e-6 * (x * x + y * y), 1e-3 * z

    def moments(d):
        m = np.array([modfun(R2, Z, d.m[i][:]) for i in range(3)])
        m[1] = min(m[1], 0.0)
        s = [modfun(R2, Z, d.ss[i][:]) for i in range(6)]
        return m, s

    (m1, s1), (m2, s2) = moments(p.disk_1), moments(p.disk_2)
    # vR and vZ: the two variants share them
    assert ks_p(v[0], stats.norm(m1[0], s1[0]).cdf) > P_OK
    assert ks_p(v[2], stats.norm(m1[2], s1[5]).cdf) > P_OK
    # vPhi: fk of disk_1, 1 - fk of disk_2
    good = mixture_cdf([m1[1], m2[1]], [s1[3], s2[3]], [p.fk, 1 - p.fk])
    assert ks_p(v[1], good) > P_OK
    assert ks_p(v[1], mixture_cdf([m1[1], m2[1]], [s1[3], s2[3]], [1 - p.fk, p.fk])) < P_BAD
    # no correlations in the default ellipsoid
    c = np.corrcoef(v)
    assert abs(c[0, 2]) < 5 / np.sqrt(N)
    # planted mistake: |Z| exponent applied to R instead
    assert ks_p(v[0], stats.norm(m1[0], 40 + 5 * R2 ** 0.75).cdf) < P_BAD


def test_bond_disk_covariance_terms():
    p = gfx.bond_params()
    # give the disk a full covariance: cov(R,Phi)=300, cov(R,Z)=-200, cov(Phi,Z)=150 (km/s)^2
    for d in (p.disk_1, p.disk_2):
        gfx._farr(d.ss[1], [300, 0, 0, 0, 0])
        gfx._farr(d.ss[2], [-200, 0, 0, 0, 0])
        gfx._farr(d.ss[4], [150, 0, 0, 0, 0])
    p.fk = 1.0   # disk_1 only
    t = run(table(N, 0), bond=p)
    v = t.arr("vcyl").astype(np.float64)
    x, y, z = transformed(p, (300.0, -200.0, 900.0))
    R2, Z = 1e-6 * (x * x + y * y), 1e-3 * z
    s = [modfun(R2, Z, p.disk_1.ss[i][:]) for i in range(6)]
    cov = np.array([[s[0] ** 2, s[1], s[2]], [s[1], s[3] ** 2, s[4]], [s[2], s[4], s[5] ** 2]])
    got = np.cov(v)
    # sampling error of a covariance ~ sqrt((s_i^2 s_j^2 + c_ij^2) / N)
    err = np.sqrt((np.outer(np.diag(cov), np.diag(cov)) + cov ** 2) / N)
    assert np.all(np.abs(got - cov) < 5 * err), (got, cov)
    # planted mistake: the R-Z covariance with the wrong sign
    bad = cov.copy()
    bad[0, 2] = bad[2, 0] = -cov[0, 2]
    assert np.abs(got - bad)[0, 2] > 10 * err[0, 2]


def test_bond_halo_rotation():
    p = gfx.bond_params(halo_sigma=(141, 75, 85))
    xyz = (300.0, -200.0, 2500.0)
    t = run(table(N, 2, xyz), bond=p)
    v = t.arr("vcyl").astype(np.float64)
    x, y, z = transformed(p, xyz)
    R2, Z = 1e-6 * (x * x + y * y), 1e-3 * z
    r = np.sqrt(R2 + Z * Z)
    ca, sa = np.sqrt(R2) / r, Z / r
    sr, sth, sphi = 141.0, 75.0, 85.0
    # (vR, vPhi, vZ) = (ca vr - sa vth, vphi, sa vr + ca vth)
    cov = np.array([[ca * ca * sr ** 2 + sa * sa * sth ** 2, 0, ca * sa * (sr ** 2 - sth ** 2)],
                    [0, sphi ** 2, 0],
                    [ca * sa * (sr ** 2 - sth ** 2), 0, sa * sa * sr ** 2 + ca * ca * sth ** 2]])
    got = np.cov(v)
    err = np.sqrt((np.outer(np.diag(cov), np.diag(cov)) + cov ** 2) / N)
    assert np.all(np.abs(got - cov) < 5 * err), (got, cov)
    assert np.abs(v.mean(axis=1)).max() < 5 * 141 / np.sqrt(N)
    # planted mistake: rotation angle with the wrong sign flips cov(vR, vZ)
    assert abs(got[0, 2] + cov[0, 2]) > 10 * err[0, 2]


# ---------------------------------------------------------------------------
# unresolvedMultiples
# ---------------------------------------------------------------------------
def multiples(algo, fraction=0.3):
    M, phi = gfx.seeded_lf()
    (sec, sec_tc), (cum, cum_tc), (inv, inv_tc) = gfx.multiples_tables(M, phi, fraction=fraction)
    keep = [sec, cum, inv]
    p = A.gf_multiples_data()
    p.secProb = gfx.tex1d(sec.ctypes.data, len(sec), sec_tc)
    p.cumLF = gfx.tex1d(cum.ctypes.data, len(cum), cum_tc)
    p.invCumLF = gfx.tex1d(inv.ctypes.data, len(inv), inv_tc)
    p.algo = algo
    return p, keep, (cum, cum_tc, inv, inv_tc)


@pytest.mark.parametrize("algo", [1, 2, 3])
def test_multiples_fraction_and_luminosity_sum(algo):
    frac = 0.3
    p, keep, _ = multiples(algo, frac)
    M1 = np.float32(8.0)
    t = run(table(N, 0, absmag=M1), mult=p, mult_comps=A.bitmap(0))
    n2 = t.arr("absmagSysNcomp")[0]
    sysm = t.arr("absmagSys")
    Mtot = t.arr("absmag")[0].astype(np.float64)
    assert set(np.unique(n2)) <= {1, 2}
    k = int((n2 == 2).sum())
    assert stats.binomtest(k, N, frac).pvalue > P_OK
    assert stats.binomtest(k, N, 1 - frac).pvalue < P_BAD
    assert np.all(sysm[0] == M1)
    single = n2 == 1
    assert np.all(Mtot[single] == np.float32(M1)) or np.allclose(Mtot[single], M1, atol=2e-6)
    M2 = sysm[1][~single].astype(np.float64)
    expect = -2.5 * np.log10(10 ** (-0.4 * M1) + 10 ** (-0.4 * M2))
    assert np.allclose(Mtot[~single], expect, atol=2e-5)
    if algo == 3:
        assert np.all(M2 == M1)
    if algo == 1:
        assert np.all(M2 >= M1)


def test_multiples_companion_magnitudes_follow_the_lf():
    p, keep, (cum, cum_tc, inv, inv_tc) = multiples(2, 0.5)
    t = run(table(N, 0, absmag=np.float32(8.0)), mult=p, mult_comps=A.bitmap(0))
    n2 = t.arr("absmagSysNcomp")[0]
    M2 = t.arr("absmagSys")[1][n2 == 2].astype(np.float64)
    # algo 2: M2 = invCumLF(u), u uniform, so cumLF(M2) is uniform
    xs = (np.arange(len(cum)) - 0.0) / cum_tc[1] + cum_tc[0]
    u = np.interp(M2, xs, cum.astype(np.float64))
    assert ks_p(u, stats.uniform(0, 1).cdf) > P_OK
    assert ks_p(u ** 0.8, stats.uniform(0, 1).cdf) < P_BAD


# ---------------------------------------------------------------------------
# photometric errors
# ---------------------------------------------------------------------------
def test_photometric_errors_are_gaussian_with_the_tabulated_sigma():
    nb = 5
    kx, ky, koff4 = gfx.photoerr_knots(nb - 1)
    L = koff4[1]
    koff = [0, L, 2 * L, 2 * L, 3 * L, 4 * L]     # band 2 has no error curve
    p = A.gf_photoerr_data()
    p.nbands = nb
    for i, k in enumerate(koff):
        p.koff[i] = k
    p.kx, p.ky = kx.ctypes.data, ky.ctypes.data
    rng = np.random.default_rng(3)
    mags = rng.uniform(18, 25, (nb, N)).astype(np.float32)
    t = run(table(N, 0, mags=mags), perr=p)
    obs = t.arr("obsMags").astype(np.float64)
    for b in range(nb):
        if koff[b + 1] == koff[b]:
            assert np.array_equal(t.arr("obsMags")[b], mags[b])
            continue
        m = mags[b].astype(np.float64)
        sig = 0.02 + np.exp(0.8 * (m - 24.0))
        zres = (obs[b] - m) / sig
        assert ks_p(zres, stats.norm().cdf) > P_OK
        # planted mistake: sigma of the neighbouring magnitude bin edge (m + 0.5)
        zbad = (obs[b] - m) / (0.02 + np.exp(0.8 * (m + 0.5 - 24.0)))
        assert ks_p(zbad, stats.norm().cdf) < P_BAD
