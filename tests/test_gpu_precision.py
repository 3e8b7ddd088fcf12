"""Precision study on the config-1 count grid (config 2).

The count pass's expectation of every (pixel, DM bin) pair,
W = rho * norm * D^3 * dA * POGSON * dm * dM * sum of LF over the bin's
absolute-magnitude window, is computed on the GPU with the window sum taken
from an FP64 prefix table (the production mode), a Kahan-compensated FP32
prefix and a naive FP32 prefix (skygen.cu.h:582-599 sums in FP32 and
FP64).  Each is compared with the FP64 oracle on every tenth pixel of the
config-1 sky (full sky |b| > 10 deg, 0.5 deg pixels, r 14-24): maximum
relative error of W and of the expected total, and the fraction of
(pixel, DM bin) units whose Poisson draw differs from the FP64 plan when
the same streams draw from each variant's W.  The FP64 GPU values and the
streams after the FP64 plan are bit-identical to the oracle's.

With GALFAST_REPORT set to a directory the numbers are also written there
as precision_c1.json."""
import ctypes as C
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import _abi as A, configs
from galfast_amd.catalog import Catalog
import oracle as O
import replay as R
from devutil import DevRng, dev_array, stream_ptr, torch

pytestmark = pytest.mark.gpu

STEP = 10          # every STEP-th pixel of the sky
MODES = {"fp64": A.GF_ACC_FP64, "kahan32": A.GF_ACC_KAHAN32, "fp32": A.GF_ACC_FP32}


@pytest.fixture(scope="module")
def c1(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("c1p"))
    cat = Catalog(configs.write_c1(d), seed=42)
    view = R.host_view(G.lib(), cat._h)
    allpix = (A.gf_pencil_beam * view.sp.npixels).from_address(view.pixels)
    pix = (A.gf_pencil_beam * ((view.sp.npixels + STEP - 1) // STEP))(*[allpix[i] for i in
                                                                        range(0, view.sp.npixels, STEP)])
    sp = A.gf_skygen_params.from_buffer_copy(view.sp)
    sp.npixels = len(pix)
    yield cat, view, sp, pix
    cat.close()


def oracle_w(sp, m, pix):
    """FP64 oracle expectations, the pixels split over threads."""
    K = sp.nm + sp.nM - 1
    w = np.zeros(sp.npixels * K)
    nth = min(16, os.cpu_count() or 1)
    bounds = np.linspace(0, sp.npixels, nth + 1).astype(int)

    def part(i):
        a, b = bounds[i], bounds[i + 1]
        if b <= a:
            return
        s = A.gf_skygen_params.from_buffer_copy(sp)
        s.npixels = int(b - a)
        sub = (A.gf_pencil_beam * (b - a)).from_buffer_copy(pix, a * C.sizeof(A.gf_pencil_beam))
        ww = np.zeros((b - a) * K)
        res = A.gf_count_result()
        O.lib().gfo_count_bins(C.byref(s), C.byref(m), C.cast(sub, C.c_void_p), None, C.byref(res), None,
                               O.ptr(ww))
        w[a * K:b * K] = ww

    with ThreadPoolExecutor(nth) as ex:
        list(ex.map(part, range(nth)))
    return w


def test_count_precision_study_c1(c1):
    cat, view, sp, pix = c1
    T = torch()
    K = sp.nm + sp.nM - 1
    U = sp.npixels * K
    pix_dev = T.from_numpy(np.frombuffer(bytes(pix), dtype=np.uint8).copy()).cuda()
    state0 = np.ctypeslib.as_array(view.rng_state0, shape=(3 * view.nstreams,)).copy()
    report = {"workload": "config-1 count grid, every %d-th pixel: %d pixels x %d DM bins x %d models"
                          % (STEP, sp.npixels, K, view.nmodels), "models": []}
    for mi in range(view.nmodels):
        mh = view.models[mi]
        lf = np.ctypeslib.as_array(C.cast(mh.lf.data, C.POINTER(C.c_float)), shape=(mh.lf.n,)).copy()
        lf_dev = dev_array(lf)
        md = A.gf_model.from_buffer_copy(mh)
        md.lf.data = lf_dev.data_ptr()
        wo = oracle_w(sp, mh, pix)
        pos = wo > 0
        entry = {"model": mi, "units": int(U), "units_with_stars": int(pos.sum()), "expected_fp64": float(wo[pos].sum())}
        plans = {}
        for name, mode in MODES.items():
            res = A.gf_count_result()
            w = T.zeros(U, dtype=T.float64, device="cuda")
            G.check(G.lib().gf_skygen_count(C.byref(sp), C.byref(md), C.c_void_p(pix_dev.data_ptr()), None, mode,
                                            C.byref(res), None, C.c_void_p(w.data_ptr()), stream_ptr()),
                    "gf_skygen_count")
            wd = w.cpu().numpy()
            rel = np.abs(wd[pos] - wo[pos]) / wo[pos]
            st = state0.copy()
            counts = np.zeros(U, dtype=np.uint32)
            O.lib().gfo_plan(O.ptr(wd), U, O.ptr(st), view.nstreams, O.ptr(counts))
            O.lib().gfo_plan_advance(O.ptr(st), view.nstreams, U)
            plans[name] = counts
            entry[name] = {"max_rel_err_W": float(rel.max()), "rel_err_total": float(abs(wd[pos].sum() - wo[pos].sum())
                                                                                    / wo[pos].sum()),
                           "bit_exact_W": bool(np.array_equal(wd, wo))}
            if name == "fp64":
                # device plan from the same streams: identical counts and streams
                drng = DevRng(state0.copy())
                plan = C.c_void_p()
                n = C.c_uint64()
                wdev = dev_array(wd)
                G.check(G.lib().gf_skygen_plan(C.byref(sp), C.byref(md), C.c_void_p(pix_dev.data_ptr()), None,
                                               C.c_void_p(wdev.data_ptr()), C.byref(drng.rng), C.byref(plan),
                                               C.byref(n), stream_ptr()), "gf_skygen_plan")
                G.lib().gf_skygen_plan_free(plan)
                entry["fp64"]["plan_total_equals_oracle"] = int(n.value) == int(counts.sum())
                entry["fp64"]["streams_bit_exact"] = bool(np.array_equal(drng.state(), st))
        # the drawn counts against the FP64 expectations: per DM bin (summed
        # over pixels) and per pixel (summed over DM bins), chi-square per
        # degree of freedom and the largest deviation in sigma
        n64 = plans["fp64"].astype(np.float64).reshape(sp.npixels, K)
        e64 = wo.reshape(sp.npixels, K)
        hist = {}
        for axis, what in ((0, "per_dm_bin"), (1, "per_pixel")):
            N, E = n64.sum(axis=axis), e64.sum(axis=axis)
            ok = E > 5.0
            z = (N[ok] - E[ok]) / np.sqrt(E[ok])
            hist[what] = {"bins": int(ok.sum()), "chi2_per_dof": float((z ** 2).mean()),
                          "max_abs_sigma": float(np.abs(z).max())}
        tot_n, tot_e = float(n64.sum()), float(e64.sum())
        hist["total_sigma"] = (tot_n - tot_e) / np.sqrt(tot_e)
        entry["fp64"]["drawn_vs_expected"] = hist
        assert 0.8 < hist["per_dm_bin"]["chi2_per_dof"] < 1.25 and 0.8 < hist["per_pixel"]["chi2_per_dof"] < 1.25
        assert abs(hist["total_sigma"]) < 5
        for name in ("kahan32", "fp32"):
            entry[name]["drawn_bin_disagreement"] = float((plans[name] != plans["fp64"]).mean())
            entry[name]["drawn_bin_disagreement_units_with_stars"] = float(
                (plans[name][pos] != plans["fp64"][pos]).mean())
        report["models"].append(entry)
        # FP64 on the GPU is the oracle's arithmetic.  The FP32 variants' W
        # is a float product times a difference of two prefix values, so a
        # narrow window far along the prefix loses digits to cancellation;
        # the totals stay close
        assert entry["fp64"]["bit_exact_W"]
        assert entry["fp64"]["plan_total_equals_oracle"] and entry["fp64"]["streams_bit_exact"]
        assert entry["kahan32"]["rel_err_total"] < 1e-4
        assert entry["fp32"]["rel_err_total"] < 1e-2
    print(json.dumps(report))
    out = os.environ.get("GALFAST_REPORT")
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "precision_c1.json"), "w") as f:
            json.dump(report, f, indent=1)


# J2000 Galactic <-> equatorial constants (modules/module_lib.h:62-71)
D2R = np.pi / 180.0
ANGP, L0, CE, SE = D2R * 192.859508333, D2R * 32.932, 0.88998740217659689, 0.45598511375586859


def pm_radec_fp64(lb, xyz, vcyl, par):
    """vel2pm (vcyl2pm, vel2pm_gpu.cu.h:221-251, and transform_pm :134-164)
    and gal2other (galequ, gal2other_gpu.cu.h:39-60) evaluated entirely in
    FP64 from the same inputs."""
    l, b = lb[0] * D2R, lb[1] * D2R
    X = par.Rg - xyz[0].astype(np.float64)
    Y = -xyz[1].astype(np.float64)
    Z = xyz[2].astype(np.float64)
    vr, vphi, vz0 = (v.astype(np.float64) for v in vcyl)
    rho = np.hypot(X, Y)
    cphi, sphi = X / rho, Y / rho
    px = -sphi * vphi + cphi * vr - par.u0
    py = cphi * vphi + sphi * vr - (par.v0 + par.vLSR)
    pz = vz0 - par.w0
    vx, vy, vz = -px, -py, pz
    cl, sl, cb, sb = np.cos(l), np.sin(l), np.cos(b), np.sin(b)
    vl = -vx * sl + vy * cl
    tmp = vx * cl + vy * sl
    vrad = cb * tmp + vz * sb
    vb = -sb * tmp + vz * cb
    D = np.sqrt((8000.0 - X) ** 2 + Y ** 2 + Z ** 2)
    pml, pmb = vl / (4.74 * D * 1e-3), vb / (4.74 * D * 1e-3)
    # equatorial proper motions
    sl0, cl0 = np.sin(l - L0), np.cos(l - L0)
    t1 = cb * SE - CE * sb * sl0
    t2 = CE * sb - cb * SE * sl0
    t4 = cb * cl0
    t3 = sb * SE + cb * CE * sl0
    cd = np.sqrt(1.0 - t3 * t3)
    vlc = pml / cb
    pma = (-CE * cl0 * pmb + cb * t1 * vlc) / (t4 * t4 + t2 * t2) * cd
    pmd = (t1 * pmb + cb * CE * cl0 * vlc) / cd
    ra = np.mod(np.arctan2(cb * cl0, sb * CE - cb * SE * sl0) + ANGP, 2 * np.pi)
    dec = np.arcsin(cb * CE * sl0 + sb * SE)
    return np.array([pma, pmd, vrad]), np.array([ra / D2R, dec / D2R]), np.abs(t3) > 0.9999999


def test_pm_radec_precision_c1(tmp_path):
    """Proper motions and equatorial coordinates of config-1 stars against
    an FP64 evaluation of the same formulas from the same inputs (lb, XYZ,
    vcyl): the float pipeline's error, reported beside the count grid's."""
    from galfast_amd.catalog import Catalog
    from test_gpu_catalog import fetch
    from galfast_amd.table import catalog_spec
    cat = Catalog(configs.write_c1(str(tmp_path)), seed=42, chunk_rows=30_000_000)
    cat.count()
    view = R.host_view(G.lib(), cat._h)
    par = A.gf_vel2pm_data.from_buffer_copy(view.chain.v2pm.contents)
    spec = {k: v for k, v in catalog_spec(nbands=view.nbands, nabsmag=view.nabsmag).items()
            if k in ("lb", "XYZ", "vcyl", "pm", "radec", "hidden")}
    take = []

    def on_chunk(t, n, comp):
        if len(take) < 3:   # the first chunk of each model
            take.append(fetch(t, min(n, 2_000_000), spec))

    cat.run(on_chunk)
    cat.close()
    pm_err, pm_rel, pm_med, ra_err, dec_err, nstars = [], [], [], [], [], 0
    for h in take:
        vis = h.arr("hidden")[0] == 0
        pm64, rd64, pole = pm_radec_fp64(h.arr("lb")[:, vis], h.arr("XYZ")[:, vis], h.arr("vcyl")[:, vis], par)
        ok = ~pole
        pm = h.arr("pm")[:, vis].astype(np.float64)
        d = np.abs(pm[:, ok] - pm64[:, ok])
        pm_err.append(d.max(axis=1))
        rel = d / np.maximum(np.abs(pm64[:, ok]), 1e-3)
        pm_rel.append(np.quantile(rel, 0.999, axis=1))
        pm_med.append(np.median(rel, axis=1))
        rd = h.arr("radec")[:, vis]
        dra = np.abs(rd[0] - rd64[0])
        dra = np.minimum(dra, 360.0 - dra)
        ra_err.append(dra.max())
        dec_err.append(np.abs(rd[1] - rd64[1]).max())
        nstars += int(vis.sum())
    pm_err = np.max(pm_err, axis=0)
    pm_rel = np.max(pm_rel, axis=0)
    pm_med = np.max(pm_med, axis=0)
    rep = {"stars": nstars,
           "pm_max_abs_err_mas_yr": [float(pm_err[0]), float(pm_err[1])],
           "vrad_max_abs_err_km_s": float(pm_err[2]),
           "pm_rel_err_q999": [float(pm_rel[0]), float(pm_rel[1]), float(pm_rel[2])],
           "pm_rel_err_median": [float(pm_med[0]), float(pm_med[1]), float(pm_med[2])],
           "ra_max_abs_err_deg": float(np.max(ra_err)), "dec_max_abs_err_deg": float(np.max(dec_err))}
    print(json.dumps(rep))
    # the reference's float formulation (l, b narrowed to float radians,
    # float velocity sums): errors far below the 0.1 mas/yr its text output
    # prints, relative errors of float size except for the smallest motions
    assert max(rep["pm_max_abs_err_mas_yr"]) < 0.01 and rep["vrad_max_abs_err_km_s"] < 0.01
    assert max(rep["pm_rel_err_median"]) < 1e-6
    assert rep["ra_max_abs_err_deg"] < 1e-12 and rep["dec_max_abs_err_deg"] < 1e-12
    out = os.environ.get("GALFAST_REPORT")
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, "precision_pm_c1.json"), "w") as f:
            json.dump(rep, f, indent=1)
