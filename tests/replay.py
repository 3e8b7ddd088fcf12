"""Replay a catalog job on the oracle (CPU).

Given the host view of a ``gf_catalog`` (``gf_catalog_host_view_get``), run
the same steps as ``gf_catalog_run`` on the oracle: per model the
(pixel, DM-bin) expectations, the Poisson plan, the emission in the same
output chunks and the module chain in the same 65,536 x 256-row sub-batches.
Used by the end-to-end parity test, by smoke() and as bench.py's CPU
baseline.  Test infrastructure.
"""
import ctypes as C
import time

import numpy as np

from galfast_amd import _abi as A
from galfast_amd.table import HostTable, catalog_spec
import oracle as O


class gf_catalog_host_view(C.Structure):
    _fields_ = [("sp", A.gf_skygen_params), ("pixels", C.c_void_p), ("nmodels", C.c_int32),
                ("nbands", C.c_int32), ("models", C.POINTER(A.gf_model)), ("foot", A.gf_footprint),
                ("chain", A.gf_chain), ("nabsmag", C.c_int32), ("nstreams", C.c_uint32),
                ("rng_state0", C.POINTER(C.c_uint32)), ("ext", A.gf_extinction),
                ("model_sp", C.POINTER(A.gf_skygen_params))]


def model_sp(view, mi, npixels=None):
    """The binning of model mi (its absolute-magnitude range), optionally
    for a subset of npixels pixels."""
    sp = A.gf_skygen_params.from_buffer_copy(view.model_sp[mi])
    if npixels is not None:
        sp.npixels = npixels
    return sp


def _ext(view):
    """The view's extinction tables (None without extinction)."""
    e = view.ext
    return C.byref(e) if (e.beam.data or e.north.data or e.south.data) else None


def host_view(lib, handle):
    lib.gf_catalog_host_view_get.argtypes = [C.c_void_p, C.POINTER(gf_catalog_host_view)]
    v = gf_catalog_host_view()
    rc = lib.gf_catalog_host_view_get(handle, C.byref(v))
    assert rc == 0
    return v


def replay(view, chunk_rows=30_000_000, pixel_index=None, on_chunk=None, time_budget=None):
    """Run the job on the oracle.  pixel_index: optional subset of pixels.
    Returns dict(generated, visible, seconds, stopped_early)."""
    o = O.lib()
    ns = view.nstreams
    state = np.ctypeslib.as_array(view.rng_state0, shape=(3 * ns,)).copy()
    allpix = (A.gf_pencil_beam * view.sp.npixels).from_address(view.pixels)
    if pixel_index is None:
        pix = allpix
    else:
        pix = (A.gf_pencil_beam * len(pixel_index))(*[allpix[i] for i in pixel_index])
    spec = catalog_spec(nbands=max(view.nbands, 1), nabsmag=max(view.nabsmag, 1))
    sub = 65536 * 256
    t0 = time.perf_counter()
    generated = visible = 0
    stopped = False
    foot = A.gf_footprint.from_buffer_copy(view.foot)
    for mi in range(view.nmodels):
        m = view.models[mi]
        sp = model_sp(view, mi, len(pix))
        K = sp.nm + sp.nM - 1
        U = sp.npixels * K
        res = A.gf_count_result()
        wpk = np.zeros(U)
        o.gfo_count_bins(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), _ext(view), C.byref(res), None,
                         O.ptr(wpk))
        counts = np.zeros(U, dtype=np.uint32)
        base = state.copy()          # the streams at the plan: the emission's too
        o.gfo_plan(O.ptr(wpk), U, O.ptr(base), ns, O.ptr(counts))
        o.gfo_plan_advance(O.ptr(state), ns, U)
        offs = np.zeros(U + 1, dtype=np.uint64)
        offs[1:] = np.cumsum(counts, dtype=np.uint64)
        total = int(offs[-1])
        generated += total
        chunk = max(1, min(chunk_rows, total))
        for r0 in range(0, total, chunk):
            r1 = min(total, r0 + chunk)
            n = r1 - r0
            t = HostTable(n, spec)
            cols = t.skygen_cols()
            o.gfo_emit(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), _ext(view), C.byref(foot), O.ptr(offs),
                       U, r0, r1, O.ptr(base), ns, C.byref(cols))
            tab = t.table()
            for b0 in range(0, n, sub):
                ks = A.gf_rows(b0, 256, min(n, b0 + sub), 0)
                rc = o.gfo_module_chain(C.byref(ks), C.byref(view.chain), O.ptr(state), ns, C.byref(tab))
                assert rc == 0
            visible += int((t.arr("hidden")[0] == 0).sum())
            if on_chunk is not None:
                on_chunk(t, n, m.comp)
            if time_budget is not None and time.perf_counter() - t0 > time_budget:
                stopped = True
                break
        if stopped:
            break
    return {"generated": generated, "visible": visible, "seconds": time.perf_counter() - t0,
            "stopped_early": stopped}


def replay_reference(view, on_chunk=None, capacity=None):
    """The catalog's reference draw mode on the oracle, model by model: the
    crawl (sp.nthreads threads on streams 0 ..), in batches of at most
    `capacity` stars (all at once by default) drawn from the streams' states
    at the model's start while the count pass leaves the streams past the
    whole draw; per batch the order of drawSources (sort by hidden, l, b,
    DM, M, Am; truncate at the first hidden row), the clipper, then the
    module chain over the visible rows.  Returns dict(generated, visible)."""
    o = O.lib()
    ns = view.nstreams
    state = np.ctypeslib.as_array(view.rng_state0, shape=(3 * ns,)).copy()
    pix = (A.gf_pencil_beam * view.sp.npixels).from_address(view.pixels)
    spec = catalog_spec(nbands=max(view.nbands, 1), nabsmag=max(view.nabsmag, 1))
    foot = A.gf_footprint.from_buffer_copy(view.foot)
    sub = 65536 * 256
    generated = visible = 0
    for mi in range(view.nmodels):
        m = view.models[mi]
        sp = model_sp(view, mi)
        st0 = state.copy()
        n = o.gfo_draw_compat(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), _ext(view), None, O.ptr(state), ns,
                              None, 0)
        generated += n
        cap = max(1, min(capacity or n, n))
        for r0 in range(0, n, cap):
            r1 = min(n, r0 + cap)
            drawn = HostTable(r1 - r0, spec)
            dcols = drawn.skygen_cols()
            assert o.gfo_draw_window(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), _ext(view), None,
                                     O.ptr(st0), ns, r0, r1, C.byref(dcols)) == r1 - r0
            t = HostTable(r1 - r0, spec)
            cols = t.skygen_cols()
            nv = o.gfo_order(C.byref(dcols), r1 - r0, C.byref(cols))
            o.gfo_clip(C.byref(sp), C.byref(foot), C.byref(cols), nv)
            tab = t.table()
            for b0 in range(0, nv, sub):
                ks = A.gf_rows(b0, 256, min(nv, b0 + sub), 0)
                assert o.gfo_module_chain(C.byref(ks), C.byref(view.chain), O.ptr(state), ns, C.byref(tab)) == 0
            visible += int((t.arr("hidden")[0, :nv] == 0).sum())
            if on_chunk is not None and nv:
                on_chunk(t, nv, m.comp)
    return {"generated": generated, "visible": visible}


def _shift(cols, row0):
    """Copy of a gf_skygen_cols / gf_table whose columns start row0 rows later."""
    out = type(cols).from_buffer_copy(cols)
    for name, _ in type(cols)._fields_:
        c = getattr(out, name)
        if c.ptr:
            c.ptr = c.ptr + row0 * c.elem
    return out


def request_streams(view, npix):
    """Streams of a server request: S = 64-rounded (pixels x DM bins), at
    least 64 and at most 65,536; the state is the first S streams of the
    pipeline seed's sequence, planes [a | c | x]."""
    K = view.sp.nm + view.sp.nM - 1
    S = min(65536, (max(64, npix * K) + 63) // 64 * 64)
    full = np.ctypeslib.as_array(view.rng_state0, shape=(3 * view.nstreams,))
    ns = view.nstreams
    st = np.concatenate([full[j * ns: j * ns + S] for j in range(3)]).copy()
    return S, st


def replay_request(view):
    """One server request on the oracle, given the host view of a catalog
    opened on the request's own beam, faint limit and seed: every model is
    planned (in model order) on the request's S streams, each from the
    streams as the previous models' plans left them, then every model is
    emitted, model after model, into one table; then the module chain runs
    over the request's 256-row chunks, chunk c on stream c.
    Returns (HostTable, expected)."""
    o = O.lib()
    npix = view.sp.npixels
    pix = (A.gf_pencil_beam * npix).from_address(view.pixels)
    S, state = request_streams(view, npix)
    foot = A.gf_footprint.from_buffer_copy(view.foot)
    plans = []
    expected = 0.0
    for mi in range(view.nmodels):
        m = view.models[mi]
        sp = model_sp(view, mi)
        K = sp.nm + sp.nM - 1
        U = npix * K
        res = A.gf_count_result()
        wpk = np.zeros(U)
        o.gfo_count_bins(C.byref(sp), C.byref(m), C.cast(pix, C.c_void_p), _ext(view), C.byref(res), None,
                         O.ptr(wpk))
        expected += float(wpk[wpk > 0].sum())
        counts = np.zeros(U, dtype=np.uint32)
        base = state.copy()
        o.gfo_plan(O.ptr(wpk), U, O.ptr(base), S, O.ptr(counts))
        o.gfo_plan_advance(O.ptr(state), S, U)
        offs = np.zeros(U + 1, dtype=np.uint64)
        offs[1:] = np.cumsum(counts, dtype=np.uint64)
        plans.append((offs, base, sp, U))
    total = int(sum(int(p[-1]) for p, _, _, _ in plans))
    spec = catalog_spec(nbands=max(view.nbands, 1), nabsmag=max(view.nabsmag, 1))
    t = HostTable(total, spec)
    row = 0
    for mi, (offs, base, sp, U) in enumerate(plans):
        n = int(offs[-1])
        if n:
            cols = _shift(t.skygen_cols(), row)
            o.gfo_emit(C.byref(sp), C.byref(view.models[mi]), C.cast(pix, C.c_void_p), _ext(view), C.byref(foot),
                       O.ptr(offs), U, 0, n, O.ptr(base), S, C.byref(cols))
        row += n
    if total:
        assert (total + 255) // 256 <= S, "replay needs at most one chunk per stream"
        ks = A.gf_rows(0, 256, total, 0)
        rc = o.gfo_module_chain(C.byref(ks), C.byref(view.chain), O.ptr(state), S, C.byref(t.table()))
        assert rc == 0
    return t, expected


def unwritten_columns(view):
    """The columns no module of the job writes: they hold whatever their
    memory held, on the device and in the oracle's buffers alike."""
    ch = view.chain
    unwritten = {"FeH_comp"} if not ch.gfeh else set()
    for ptr, cols in ((ch.mult, ("absmagSys", "absmagSysNcomp")), (ch.bond, ("vcyl",)), (ch.v2pm, ("pm",)),
                      (ch.phot, ("mags", "photoFlags")), (ch.perr, ("obsMags",)), (ch.feh or ch.gfeh, ("FeH",))):
        if not ptr:
            unwritten |= set(cols)
    if not ch.gal2other_coordsys:
        unwritten.add("radec")
    return unwritten
