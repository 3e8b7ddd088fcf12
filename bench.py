"""Catalog throughput of the MI355X pipeline on one GPU.

A step is one complete catalog job of the chosen workload: the count pass
over every (pixel, DM-bin) pair of every model, the Poisson plan, the star
draw and the fused per-star module chain, chunk by chunk, with all inputs
(pixels, luminosity functions, colour tables, error curves, RNG streams)
resident in HBM before the timed region.  ``value`` is visible catalog
stars per second.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1]

The CPU baseline replays a bounded sample of the same job (blocks of sky
pixels drawn at random) on the oracle, one replay per host core, and once
on a single core.  The serving record times the request server on the
config-3 stream of footprint requests beside one catalog job per request;
the c4 sub-record repeats the catalog measurement on config 4.
"""
import argparse
import json
import os
import sys
import tempfile
import time

# the catalog's four HIP streams each on a hardware queue of their own
# (galfast_amd/__init__.py): at least 8 queues for this process, whatever
# lower value the environment holds; set before anything initialises HIP
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    "c1": "full sky |b|>10 deg, 0.5 deg Lambert pixels, r 14-24, dm 0.05, Mr 3-15; thin+thick disk + "
          "power-law halo; Ivezic [Fe/H]; Bond2010 kinematics (halo 140/105/75 km/s); equatorial "
          "proper motions and coordinates; 5 synthetic bands with sigma(m)=0.02+exp(0.8(m-24)) errors; "
          "3e7-row output chunks",
    "c0": "10x10 deg patch at (l,b)=(90,60), 1 deg pixels, r 15-21, dm 0.05, Mr 3-15; thin+thick disk + "
          "halo; Ivezic [Fe/H]; 5 synthetic bands",
    "c4": "l 60-100, b 20-50 (~970 deg^2), 0.25 deg pixels, r 14-27.5, dm 0.05, Mr 3-15; thin+thick disk + "
          "halo; analytic exponential dust disk (h=100 pc, 0.5 mag/kpc) as north/south FITS extinction cubes; "
          "unresolved binaries (fraction 0.5, LF companions); Ivezic [Fe/H]; 6 synthetic bands with errors",
}

PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per row of each kernel: every column a kernel must read,
# and every column it writes, once.  Hidden rows cost the 4-byte hidden flag
# in the module kernels.
# emission, per generated row: the 10 skygen columns (64 B) written; the
# unit data (offsets, start state, pixel) is read once per unit, not per row
EMIT_STARS_BYTES_PER_ROW = 64
STAGE_NAMES = ["k_chain_draws", "k_chain_rows<FEH>", "k_chain_rows<BOND>", "k_chain_rows<PHOT>",
               "k_chain_rows<PERR>", "k_chain_rows<ASTRO>"]


def stage_bytes(nbands, nabsmag, present):
    """(bytes per visible row, bytes per hidden row) of the module-chain
    launches for a job with nbands bands, nabsmag system components (2 or
    more: unresolved multiples) and the row passes in `present`."""
    nb = nbands
    mult = nabsmag > 1
    npos = sum(1 for k in ("k_chain_rows<FEH>", "k_chain_rows<BOND>", "k_chain_rows<PERR>") if k in present)
    # draw pass: read hidden + comp of every row; per visible row write the
    # 8-byte stream position of each random-number module; unresolved
    # multiples draw their companions in it: read absmag, write absmagSys
    # and its count (and, with more than two planes, the system's absmag)
    split = nabsmag == 2
    draws = 4 + 4 + 8 * npos + ((4 + 4 * nabsmag + 4 + (0 if split else 4)) if mult else 0)
    mr = 4 * nabsmag if mult else 4
    return {
        "k_chain_draws": (draws, 4 + 4),
        # metallicity rows: read hidden, comp, XYZ, 1 position; write FeH;
        # with two magnitude planes also the system's magnitude: read
        # absmagSys and its count, write absmag
        "k_chain_rows<FEH>": (4 + 4 + 12 + 8 + 4 + ((4 * nabsmag + 4 + 4) if mult and split else 0), 4),
        # kinematics rows: read hidden, comp, XYZ, 1 position; write vcyl
        "k_chain_rows<BOND>": (4 + 4 + 12 + 8 + 12, 4),
        # photometry rows: read hidden, comp, FeH, the absolute magnitude(s),
        # DM, Am; write mags, photoFlags
        "k_chain_rows<PHOT>": (4 + 4 + 4 + mr + 4 + 4 + 4 * nb + 4, 4),
        # photometric-error rows: read hidden, mags, 1 position; write obsMags
        "k_chain_rows<PERR>": (4 + 4 * nb + 8 + 4 * nb, 4),
        # astrometry rows: read hidden, lb, XYZ, vcyl; write pm, radec
        "k_chain_rows<ASTRO>": (4 + 16 + 12 + 12 + 12 + 16, 4),
    }


def table_row_bytes(nbands, nabsmag, present):
    """Bytes of one output row: the skygen columns and the columns of the
    modules that run."""
    b = 64 + 4 + 4 * nbands + 4                       # skygen, FeH, mags, photoFlags
    if nabsmag > 1:
        b += 4 * nabsmag + 4                          # absmagSys, absmagSysNcomp
    if "k_chain_rows<PERR>" in present:
        b += 4 * nbands
    if "k_chain_rows<BOND>" in present:
        b += 12
    if "k_chain_rows<ASTRO>" in present:
        b += 12 + 16                                  # pm, radec
    return b


def host_cpu_info():
    """The host's processor model (/proc/cpuinfo), logical CPU count (nproc),
    the CPUs this process may run on (affinity) and the cgroup's CPU quota
    (cpu.max, in CPUs; None when unlimited or unreadable)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff, "cgroup_cpus": quota}


def default_cpu_threads():
    """Every core this process can use: the affinity set, capped by the
    cgroup's CPU quota."""
    info = host_cpu_info()
    n = info["affinity"]
    if info["cgroup_cpus"]:
        n = min(n, max(1, int(info["cgroup_cpus"])))
    return max(1, n)


def cpu_baseline(cat, workload, budget_s, threads):
    """The oracle (a single-threaded C++ restatement of the reference's CPU
    path) timed on the host: once on one thread, and on `threads` threads
    at once (the oracle calls release the GIL, so they run in parallel).
    Each thread replays its own sample of pixels drawn at random (fixed
    seed) from the whole sky, so the sample's star mix is the job's."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import galfast_amd as G
    import replay as R
    view = R.host_view(G.lib(), cat._h)
    npix = view.sp.npixels
    block = max(1, min(npix // max(threads, 1), 400))
    perm = np.random.default_rng(12345).permutation(npix)

    def run(i, budget):
        idx = sorted(int(v) for v in perm[i * block:(i + 1) * block])
        return R.replay(view, chunk_rows=30_000_000, pixel_index=idx, time_budget=budget), len(idx)

    one, n1 = run(0, budget_s)
    rate1 = one["visible"] / one["seconds"] if one["seconds"] > 0 else 0.0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(lambda i: run(i, budget_s), range(threads)))
    wall = time.perf_counter() - t0
    vis = sum(r["visible"] for r, _ in res)
    host = host_cpu_info()
    return {"value": vis / wall, "unit": "stars/s", "cores": threads, "kind": "port",
            "host": host,
            "sample": "%d threads, each replaying %d pixels drawn at random from the %d of the sky on the oracle "
                      "(%d visible stars in %.1f s wall); host: %s, nproc %s, %s CPUs in the affinity set, "
                      "cgroup quota %s CPUs" % (threads, block, npix, vis, wall, host["model"], host["nproc"],
                                                host["affinity"], host["cgroup_cpus"]),
            "single_thread": {"value": rate1, "cores": 1,
                              "sample": "%d pixels (%d visible stars, %.1f s)" % (n1, one["visible"],
                                                                                    one["seconds"])}}


def _serve_phase(srv, reqs, rate, id_offset=0):
    """Submit reqs to the open server srv, their ids shifted by id_offset,
    with Poisson arrivals at `rate` req/s (None: all at once), and wait for
    them; the phase's record (latency = done - submit on the server's
    clock, worker times as the phase's share of the server's totals)."""
    import numpy as np
    from galfast_amd.server import Server
    n0 = len(srv.results)
    st0 = srv.stats()
    rng = np.random.default_rng(1)
    t = Server.clock()
    for r in reqs:
        if rate:
            t += rng.exponential(1.0 / rate)
            while True:
                dt = t - Server.clock()
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.001))
        srv.submit(r[0] + id_offset, *r[1:])
    srv.wait()
    res = srv.results[n0:]
    st = srv.stats()
    lat = np.array([x.t_done - x.t_submit for x in res])
    wall = max(x.t_done for x in res) - min(x.t_submit for x in res)
    stars = sum(x.visible for x in res)
    nb = len(set((x.t_start, x.batch_size) for x in res))
    return {"requests": len(res), "failed": sum(1 for x in res if x.status != 0), "stars": stars,
            "rows": sum(x.generated for x in res), "wall_s": wall, "stars_per_s": stars / wall,
            "p50_ms": 1000.0 * float(np.percentile(lat, 50)), "p99_ms": 1000.0 * float(np.percentile(lat, 99)),
            "mean_batch": len(res) / max(nb, 1), "host_prep_s": st.t_prep - st0.t_prep,
            "worker_batch_s": st.t_batch - st0.t_batch, "worker_device_wait_s": st.t_wait - st0.t_wait,
            "worker_deliver_s": st.t_deliver - st0.t_deliver}


WARM_IDS = 10 ** 8   # id offset of a warm-up pass over a request stream


def _serve_once(conf, reqs, rate, warm, full_warm=False, **kw):
    """Serve reqs on a new server after `warm` warm-up requests (and, with
    full_warm, after a first pass over reqs themselves under other ids, so
    that the workers' batch buffers have reached the sizes the stream's
    batches need when the timed pass starts)."""
    from galfast_amd.server import Server
    srv = Server(conf, None, hbm_budget=200 << 30, **kw)
    try:
        _serve_phase(srv, warm, None)
        if full_warm:
            _serve_phase(srv, reqs, None, WARM_IDS)
        return _serve_phase(srv, reqs, rate)
    finally:
        srv.close()


def _plain_jobs(reqs, seed, tmpdir):
    """The plain GPU path: every request as its own catalog job (its own
    configuration load, tables, 65,536-stream seeding, count, draw and
    modules; nothing kept between requests).  The configuration files a
    user would already have on disk are written before the clock starts.
    Returns stars, seconds, per-request latencies."""
    import numpy as np
    from galfast_amd import configs
    from galfast_amd.catalog import Catalog
    import galfast_amd as G
    confs = []
    for rid, l, b, r, m1, rseed in reqs:
        d = os.path.join(tmpdir, "job%d" % rid)
        os.makedirs(d, exist_ok=True)
        confs.append(configs.write_c3(d, seed, m1=float(np.float32(m1)), beam=(l, b, r)))
    stars, lat = 0, []
    t_all = time.perf_counter()
    for conf, (rid, l, b, r, m1, rseed) in zip(confs, reqs):
        t0 = time.perf_counter()
        cat = Catalog(conf, seed=rseed, chunk_rows=1 << 22)
        cat.count()
        st = cat.run()
        G.lib().gf_sync(None)
        cat.close()
        lat.append(time.perf_counter() - t0)
        stars += st.visible
    return stars, time.perf_counter() - t_all, lat


def output_bench(args):
    """The output stages on the config-1 workload: the two device kernels
    alone on the job's first 3e7-row chunk (gf_fits_pack_rows and
    gf_text_format_rows over the rows as drawn, hidden ones included), and
    whole jobs written through the FITS and text writers to /dev/null
    (device packing or formatting, the copy to pinned host memory and the
    host write; the count pass is not timed)."""
    import ctypes as C
    import torch
    import galfast_amd as G
    from galfast_amd import _abi as A, configs
    from galfast_amd.catalog import CHUNK_FN, Catalog, gf_run_stats, _declare

    L = G.lib()
    _declare()
    d = tempfile.mkdtemp(prefix="galfast_out_")
    cmd = configs.WRITERS["c1"](d, args.seed)
    # a sky-count map of the observed magnitudes: 1-degree (l, b) beams,
    # half-magnitude bins from 10 to 30, one coadd
    configs.add_countsmap(cmd, dict(COUNTSMAP_C1))
    res = {}

    def grab(user, table, n, comp):
        if not res:
            torch.cuda.synchronize()   # the catalog's streams idle while the kernels are timed
            res.update(_pack_and_format(L, A, table.contents, n))
        return 0

    cat = Catalog(cmd, seed=args.seed, chunk_rows=30_000_000)
    cat.count()
    st = gf_run_stats()
    cb = CHUNK_FN(grab)
    G.check(L.gf_catalog_run(cat._h, cb, None, C.byref(st)), "gf_catalog_run")
    for kind in ("fits", "text", "countsmap"):
        cat.count()
        t0 = time.perf_counter()
        st, rows = cat.write("/dev/null", kind)
        dt = time.perf_counter() - t0
        res["%s_file" % kind] = {"rows": rows, "s": dt, "stars_per_s": rows / dt}
    cat.close()
    return res


COUNTSMAP_C1 = {"xy": "lb", "z": "obsSYN", "x0": 0.5, "x1": 359.5, "y0": -89.5, "y1": 89.5, "z0": 10, "z1": 30,
                "dz": 0.5, "coadd.any": "0 0 0 0 0"}

# the config-1 output columns: table column, FITS type, text flag, width, precision
OUTPUT_C1 = [("lb", "D", 1, 13, 8), ("radec", "D", 1, 13, 8), ("XYZ", "E", 1, 10, 2), ("DM", "E", 1, 7, 3),
             ("absmag", "E", 1, 7, 3), ("comp", "J", 0, 3, -1), ("FeH", "E", 1, 6, 3), ("vcyl", "E", 1, 7, 1),
             ("pm", "E", 1, 8, 2), ("Am", "E", 0, 6, 3), ("AmInf", "E", 0, 6, 3), ("obsMags", "E", 1, 7, 3),
             ("mags", "E", 1, 7, 3), ("photoFlags", "J", 1, 4, -1)]


def _pack_and_format(L, A, table, n, reps=3):
    """Best of `reps` timed calls of each output kernel on one chunk."""
    import ctypes as C
    import torch
    import galfast_amd as G

    class gf_outcol(C.Structure):
        _fields_ = [("col", A.gf_col), ("map", C.c_void_p), ("nmap", C.c_uint32), ("type", C.c_char),
                    ("flags", C.c_char), ("width", C.c_int16), ("prec", C.c_int32)]

    cols = (gf_outcol * len(OUTPUT_C1))()
    rec = 0
    for i, (name, t, fl, w, p) in enumerate(OUTPUT_C1):
        c = getattr(table, name)
        cols[i].col = c
        cols[i].type = t.encode()
        cols[i].flags = bytes([fl])
        cols[i].width = w
        cols[i].prec = p
        rec += c.width * c.elem
    L.gf_fits_pack_rows.argtypes = [C.c_void_p, C.c_int, A.gf_col, C.c_uint64, C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64), C.c_void_p]
    L.gf_text_format_rows.argtypes = [C.c_void_p, C.c_int, A.gf_col, C.c_uint64, C.c_void_p, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_void_p]

    def best(call):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            G.check(call(), "output kernel")
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    out = {}
    buf = torch.empty(n * rec, dtype=torch.uint8, device="cuda")
    nv = C.c_uint64()
    dt = best(lambda: L.gf_fits_pack_rows(cols, len(cols), table.hidden, n, buf.data_ptr(), n, C.byref(nv), None))
    # every row's hidden flag and each visible row's columns read, each record written
    moved = 4 * n + 2 * nv.value * rec
    out["fits_pack"] = {"rows": n, "records": nv.value, "record_bytes": rec, "ms": 1e3 * dt,
                        "GBps": moved / dt / 1e9}
    del buf
    cap = n * 320
    tbuf = torch.empty(cap, dtype=torch.uint8, device="cuda")
    nb = C.c_uint64()
    dt = best(lambda: L.gf_text_format_rows(cols, len(cols), table.hidden, n, tbuf.data_ptr(), cap, C.byref(nb),
                                            C.byref(nv), None))
    out["text_format"] = {"rows": n, "lines": nv.value, "bytes": nb.value, "ms": 1e3 * dt,
                          "lines_per_s": nv.value / dt, "text_GBps": nb.value / dt / 1e9}
    del tbuf
    # the countsMap binning of the same rows (COUNTSMAP_C1's map)
    p = A.gf_countsmap_params()
    p.x0, p.dx, p.y0, p.dy, p.z0, p.dz = 0.5, 1.0, -89.5, 1.0, 10.0, 0.5
    p.nx, p.ny, p.nz = 360, 180, 41
    p.zmag_width, p.ncoadd, p.equalarea = 5, 1, 0
    co = torch.zeros(5, dtype=torch.float32, device="cuda")
    p.coadd_offsets = co.data_ptr()
    zw = 6
    counts = torch.zeros(360 * 180 * 41 * zw, dtype=torch.int32, device="cuda")
    beams = torch.zeros(360 * 180, dtype=torch.uint8, device="cuda")
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    dt = best(lambda: L.gf_countsmap_bin(C.byref(p), table.lb, table.obsMags, table.hidden, n, counts.data_ptr(),
                                         beams.data_ptr(), total.data_ptr(), None))
    # read: every row's hidden flag, each visible row's (l, b) and 5 magnitudes
    out["countsmap_bin"] = {"rows": n, "ms": 1e3 * dt, "rows_per_s": n / dt,
                            "GBps": (4 * n + (16 + 20) * int(total.item()) // (5 * reps)) / dt / 1e9}
    return out


def serving(args):
    """Request-server measurement on the serving workload: the whole request
    set served with continuous batching (all submitted at once), Poisson
    arrivals at each rate for the latency percentiles, the whole set served
    one at a time with resident tables (unbatched: no batching, no overlap),
    and a sample of the requests run as separate catalog jobs (the plain GPU
    path)."""
    from galfast_amd import configs
    d = tempfile.mkdtemp(prefix="galfast_serve_")
    conf = configs.write_c3(d, args.seed)
    reqs = configs.c3_requests(args.serve_requests, seed=args.seed)
    warm = configs.c3_requests(16, seed=args.seed + 1, id0=10 ** 7)
    out = {"workload": "%d requests: circular beams, area log-uniform 0.01-100 deg^2, centres uniform with "
                       "|b|>5 deg, r 14-m1 with m1 uniform in [20,25], seed per request id; config-1 models "
                       "and modules, 0.25 deg pixels; HBM budget 200 GB" % len(reqs)}
    import numpy as np
    # The phases run from the quietest host state to the noisiest: the plain
    # jobs open and free a catalog (gigabytes of device buffers) per request,
    # and for seconds afterwards the host runs the server's CPU work (beam
    # pixelization, stream seeding) at a third of its speed, so they go last.
    # One batching server serves every batched phase, after a warm-up pass
    # over the stream: device allocations (which cost seconds on a device
    # that has just freed tens of gigabytes) happen before the timed passes,
    # as in a long-running server.
    from galfast_amd.server import Server
    srv = Server(conf, None, hbm_budget=200 << 30, max_batch=args.serve_batch, workers=2)
    try:
        _serve_phase(srv, warm, None)
        _serve_phase(srv, reqs, None, WARM_IDS)
        out["batched"] = _serve_phase(srv, reqs, None)
        # Poisson arrivals: at each rate the first min(all, rate x 10 s) requests
        out["poisson"] = {}
        for i, rate in enumerate(args.serve_rates):
            n = max(20, min(len(reqs), int(rate * args.serve_seconds)))
            out["poisson"]["%g" % rate] = _serve_phase(srv, reqs[:n], rate, (i + 2) * WARM_IDS)
    finally:
        srv.close()
    out["unbatched"] = _serve_once(conf, reqs, None, warm, full_warm=True, plain=True)
    # plain: one catalog job per request, on every k-th request (a bounded
    # sample of the same request stream)
    k = max(1, len(reqs) // args.plain_jobs)
    sample = reqs[::k][: args.plain_jobs]
    _plain_jobs(warm[:2], args.seed, d)
    pstars, psec, plat = _plain_jobs(sample, args.seed, d)
    out["plain_job"] = {"requests": len(sample), "stars": pstars, "wall_s": psec, "stars_per_s": pstars / psec,
                        "p50_ms": 1000.0 * float(np.percentile(plat, 50)),
                        "p99_ms": 1000.0 * float(np.percentile(plat, 99)),
                        "sample": "every %d-th request of the stream" % k}
    out["served_vs_plain"] = out["batched"]["stars_per_s"] / out["plain_job"]["stars_per_s"]
    out["served_vs_unbatched"] = out["batched"]["stars_per_s"] / out["unbatched"]["stars_per_s"]
    return out


def footprint_cover_bench(reps=20):
    """Sky-grid coverage of the config-1 footprint (|b| >= 10 deg, 0.5 deg
    cells, both pole grids): gf_footprint_cover, wall clock over `reps`
    launches on the null stream."""
    import ctypes as C
    import math
    import galfast_amd as G
    from galfast_amd import _abi as A
    L = G.lib()
    f = A.gf_footprint()
    f.nregion = 1
    r = f.r[0]
    r.kind, r.l0, r.l1, r.b0, r.b1, r.bmin = 1, 0.0, 360.0, -90.0, 90.0, 10.0
    dx = 0.5 * math.pi / 180
    n = int(math.ceil(2 * math.sqrt(2.0) / dx))
    buf = C.c_void_p()
    G.check(L.gf_malloc(C.byref(buf), 2 * n * n), "gf_malloc")
    try:
        G.check(L.gf_footprint_cover(C.byref(f), dx, n, buf, None), "gf_footprint_cover")
        G.check(L.gf_sync(None), "gf_sync")
        t0 = time.perf_counter()
        for _ in range(reps):
            G.check(L.gf_footprint_cover(C.byref(f), dx, n, buf, None), "gf_footprint_cover")
        G.check(L.gf_sync(None), "gf_sync")
        dt = (time.perf_counter() - t0) / reps
    finally:
        L.gf_free(buf)
    return {"cells": 2 * n * n, "samples": 128 * n * n, "ms": 1e3 * dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=sorted(WORKLOADS))
    ap.add_argument("--overlap", type=int, default=1, choices=(0, 1),
                    help="0: every launch of the catalog on one stream (isolated kernel durations)")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="threads of the CPU baseline (default: every core this process may use)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the config-4 sub-record")
    ap.add_argument("--c4-steps", type=int, default=3)
    ap.add_argument("--output-bench", action="store_true",
                    help="also time the output stages (FITS and text) on the config-1 workload")
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-batch", type=int, default=256)
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[10.0, 100.0, 1000.0])
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="arrival window per Poisson rate")
    ap.add_argument("--plain-jobs", type=int, default=40, help="requests timed as separate catalog jobs")
    ap.add_argument("--serve-only", action="store_true",
                    help="only serve the request stream with continuous batching (all submitted at once) and "
                         "print that record (the command the serving profiles are taken with)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (the reference kernel for scaling "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4}))
        return
    if args.serve_only:
        import galfast_amd as G
        from galfast_amd import configs
        G.check(G.lib().gf_init(0), "gf_init")
        d = tempfile.mkdtemp(prefix="galfast_serve_")
        conf = configs.write_c3(d, args.seed)
        reqs = configs.c3_requests(args.serve_requests, seed=args.seed)
        warm = configs.c3_requests(16, seed=args.seed + 1, id0=10 ** 7)
        print(json.dumps({"batched": _serve_once(conf, reqs, None, warm, full_warm=True, max_batch=args.serve_batch,
                                                 workers=2)}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the sky would shard by "
              "pixels across GPUs); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.output_bench:
        import torch   # device buffers of the output timing; torch takes the device first
        torch.cuda.init()
    import galfast_amd as G
    G.check(G.lib().gf_init(0), "gf_init")
    out, cat = measure(args, args.workload, args.steps, args.warmup, args.overlap)
    if args.workload == "c1":
        out["footprint_cover"] = footprint_cover_bench()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cat, args.workload, args.cpu_budget,
                                           args.cpu_threads or default_cpu_threads())
    cat.close()
    if args.output_bench:
        out["output"] = output_bench(args)
    if not args.no_serve:
        sv = serving(args)
        out["serving"] = sv
        top = sv["poisson"][max(sv["poisson"], key=float)] if sv["poisson"] else None
        out["p50_ms"] = top["p50_ms"] if top else None
        out["p99_ms"] = top["p99_ms"] if top else None
    if args.workload == "c1" and not args.no_c4:
        out["c4"] = c4_bench(args)
    print(json.dumps(out))


def measure(args, workload, steps, warmup, overlap=1):
    """Time `steps` catalog steps of `workload` after `warmup` untimed ones.
    Returns (the bench record without the CPU and serving parts, the open
    catalog)."""
    import galfast_amd as G
    from galfast_amd import configs
    from galfast_amd.catalog import Catalog
    d = tempfile.mkdtemp(prefix="galfast_bench_%s_" % workload)
    cmd = configs.WRITERS[workload](d, args.seed)
    cat = Catalog(cmd, seed=args.seed, chunk_rows=30_000_000)
    cat.set("overlap", overlap)
    info = cat.info()

    def step():
        t0 = time.perf_counter()
        expected = cat.count()
        t1 = time.perf_counter()
        st = cat.run()
        return expected, t1 - t0, st

    for _ in range(warmup):
        step()
    G.lib().gf_sync(None)
    t0 = time.perf_counter()
    acc = {"visible": 0, "generated": 0, "t_count": 0.0, "t_plan": 0.0, "t_emit": 0.0, "t_chain": 0.0,
           "n_emit": 0, "n_chain": 0, "expected": 0.0, "t_stage": [0.0] * len(STAGE_NAMES),
           "t_emit_stars": 0.0, "t_emit_units": 0.0}
    for _ in range(steps):
        expected, tc, st = step()
        acc["visible"] += st.visible
        acc["generated"] += st.generated
        acc["t_count"] += tc
        acc["t_plan"] += st.t_plan
        acc["t_emit"] += st.t_emit
        acc["t_emit_stars"] += st.t_emit_stars
        acc["t_emit_units"] += st.t_emit_units
        acc["t_chain"] += st.t_chain
        acc["n_emit"] += st.n_emit
        acc["n_chain"] += st.n_chain
        acc["expected"] += expected
        for q in range(len(STAGE_NAMES)):
            acc["t_stage"][q] += st.t_stage[q]
    G.lib().gf_sync(None)
    elapsed = time.perf_counter() - t0
    value = acc["visible"] / elapsed

    # roofline of the dominant kernel, from HIP-event durations recorded on
    # the stream each launch runs on
    hidden = acc["generated"] - acc["visible"]
    present = {STAGE_NAMES[q] for q in range(len(STAGE_NAMES)) if acc["t_stage"][q] > 0}
    sbytes = stage_bytes(info.nbands, info.nabsmag, present)
    kernels = {"k_emit_stars": (acc["t_emit_stars"], acc["n_emit"], EMIT_STARS_BYTES_PER_ROW * acc["generated"])}
    for q, name in enumerate(STAGE_NAMES):
        if name in present:
            bv, bh = sbytes[name]
            kernels[name] = (acc["t_stage"][q], acc["n_chain"], bv * acc["visible"] + bh * hidden)
    name = max(kernels, key=lambda k: kernels[k][0])
    t, n, bytes_total = kernels[name]
    # HBM bytes per launch of that kernel from the committed rocprofv3
    # counter runs of this benchmark, if measured
    traffic = None
    tj = os.path.join(ROOT, "profiles", "traffic.json" if workload == "c1" else "traffic_%s.json" % workload)
    if os.path.exists(tj):
        with open(tj) as f:
            tr = json.load(f)
        if name in tr:
            traffic = tr[name]["bytes"]
    per_launch_s = t / max(n, 1)
    achieved = bytes_total / max(n, 1) / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
    # vector-instruction issue of the same kernel: its VALU instructions per
    # launch from the committed counter runs over the launch duration
    # measured here, as SIMD cycles per instruction (1,024 SIMDs at 2.4 GHz;
    # a wave64 FMA occupies a SIMD for 4 cycles)
    valu_issue = None
    vj = os.path.join(ROOT, "profiles", "valu.json" if workload == "c1" else "valu_%s.json" % workload)
    if os.path.exists(vj):
        with open(vj) as f:
            vt = json.load(f)
        if name in vt and vt[name]["valu_insts"] > 0:
            valu_issue = {"kernel": name, "valu_insts_per_launch": vt[name]["valu_insts"],
                          "simd_cycles_per_valu_inst": 1024 * 2.4e9 * per_launch_s / vt[name]["valu_insts"],
                          "isolated_simd_cycles_per_valu_inst": vt[name]["simd_cycles_per_valu"]}
    # the count integral's vector arithmetic, from the same counter runs
    # (SQ_INSTS_VALU_FLOPS_FP32 / _FP64 over the isolated launch)
    count_valu = None
    if os.path.exists(vj):
        with open(vj) as f:
            vt = json.load(f)
        ck = next((k for k in vt if k.startswith("k_count<")), None)
        if ck and "fp32_tflops" in vt[ck]:
            count_valu = {"kernel": ck, "fp32_tflops": vt[ck]["fp32_tflops"], "fp64_tflops": vt[ck]["fp64_tflops"],
                          "peak_fp32_tflops": 157.3, "peak_fp64_tflops": 78.6,
                          "simd_cycles_per_valu_inst": vt[ck]["simd_cycles_per_valu"],
                          "source": os.path.relpath(vj, ROOT)}
    row_b = table_row_bytes(info.nbands, info.nabsmag, present)
    rows_per_s = acc["generated"] / elapsed
    out = {
        "metric": "catalog stars/sec (full pipeline) on 1 GPU vs plain; p50/p99 request latency",
        "value": value,
        "unit": "stars/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1000.0 * elapsed / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "%s: %s" % (workload, WORKLOADS[workload]),
                   "pixels": int(info.npixels), "nm": info.nm, "nM": info.nM,
                   "stars_per_step": acc["visible"] // steps,
                   "rows_per_step": acc["generated"] // steps,
                   "expected_per_step": acc["expected"] / steps},
        "breakdown_ms_per_step": dict(
            [(k, 1000.0 * acc[k] / steps) for k in ("t_count", "t_plan", "t_emit", "t_chain")] +
            [("k_emit_stars", 1000.0 * acc["t_emit_stars"] / steps),
             ("k_tile_units", 1000.0 * acc["t_emit_units"] / steps)] +
            [(STAGE_NAMES[q], 1000.0 * acc["t_stage"][q] / steps) for q in range(len(STAGE_NAMES))]),
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_total / max(n, 1)},
        # the whole job against HBM: every output row written once
        "pipeline_roofline": {"bytes_per_row": row_b, "achieved": rows_per_s * row_b / 1e9,
                              "peak": PEAK_HBM_GBS, "unit": "GB/s",
                              "frac": rows_per_s * row_b / 1e9 / PEAK_HBM_GBS},
        "valu_issue": valu_issue,
        "count_valu": count_valu,
    }
    return out, cat


def c4_bench(args):
    """The deep config-4 job (extinction, binaries, 6 bands): a few timed
    steps with its own roofline; the drawn stars per model and apparent-
    magnitude bin beside the count's expectation per model; and the same
    footprint served as 200 rect sub-requests tiling it."""
    import numpy as np
    import galfast_amd as G
    rec, cat = measure(args, "c4", args.c4_steps, 1, 1)
    out = {k: rec[k] for k in ("value", "unit", "ms_per_step", "steps", "config", "breakdown_ms_per_step",
                               "roofline", "pipeline_roofline", "valu_issue")}
    # expected stars per model (count) and drawn stars per model and 1-mag
    # bin of the primary's r (untimed pass)
    cat.count()
    ids = cat.component_ids()
    per = {"thin disk": 0, "thick disk": 1, "halo": 2}
    bins = np.arange(14.0, 28.0, 1.0)
    hist = np.zeros((len(ids), len(bins) - 1), dtype=np.int64)
    L = G.lib()
    import ctypes as C
    L.gf_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]

    def col(c, n, dt, plane=0):
        a = np.empty(n, dtype=dt)
        G.check(L.gf_memcpy_d2h(a.ctypes.data, c.ptr + plane * c.pitch, n * c.elem, None), "copy")
        return a

    def on_chunk(t, n, comp):
        vis = col(t.hidden, n, np.int32) == 0
        m = col(t.DM, n, np.float32) + col(t.absmagSys, n, np.float32, 0)
        h, _ = np.histogram(m[vis], bins=bins)
        hist[comp] += h

    cat.run(on_chunk)
    out["per_model"] = {name: {"expected": float(cat.per_model[i].nstarsExpected),
                               "drawn_visible_per_mag": dict(("%g-%g" % (bins[j], bins[j + 1]), int(hist[i][j]))
                                                             for j in range(len(bins) - 1))}
                        for name, i in per.items() if i < len(ids)}
    out["count_note"] = ("the luminosity function spans Mr 3-15 (the workload's LF tabulation), so at r = 27.5 "
                         "the faintest stars it holds are seen to DM 12.5 (3.2 kpc); the dust disk removes the "
                         "near-plane volume; over 970 deg^2 at b 20-50 this gives %.3g stars, not the ~1e9 "
                         "the workload text estimates for an LF that continues to fainter dwarfs"
                         % rec["config"]["expected_per_step"])
    cat.close()
    out["served_200"] = c4_served(args)
    return out


def c4_served(args):
    """Config 4 as 200 Galactic rect sub-requests (20 x 10 tiles of 2 x 3
    deg) submitted at once to one server, after a warm-up pass over the same
    tiles; stars/s, latency, batches."""
    import numpy as np
    from galfast_amd import configs
    from galfast_amd.server import Server
    d = tempfile.mkdtemp(prefix="galfast_c4serve_")
    conf = configs.write_c4(d, args.seed)
    tiles = configs.c4_tiles()
    srv = Server(conf, None, hbm_budget=200 << 30, max_batch=args.serve_batch, workers=2)
    try:
        warm = configs.c4_tiles(2, 1, l=(60.0, 62.0), b=(20.0, 23.0), id0=10 ** 7)
        for r in warm:
            srv.submit(*r)
        srv.wait()
        # a warm-up pass over the tiles under other ids: the workers' batch
        # buffers reach their sizes before the timed pass
        _serve_phase(srv, tiles, None, WARM_IDS)
        nwarm = len(srv.results)
        for r in tiles:
            srv.submit(*r)
        srv.wait()
        res = srv.results[nwarm:]
    finally:
        srv.close()
    lat = np.array([x.t_done - x.t_submit for x in res])
    wall = max(x.t_done for x in res) - min(x.t_submit for x in res)
    stars = sum(x.visible for x in res)
    return {"requests": len(res), "failed": sum(1 for x in res if x.status != 0), "stars": stars,
            "rows": sum(x.generated for x in res), "expected": sum(x.expected for x in res),
            "wall_s": wall, "stars_per_s": stars / wall,
            "p50_ms": 1000.0 * float(np.percentile(lat, 50)), "p99_ms": 1000.0 * float(np.percentile(lat, 99)),
            "batches": len(set(x.batch for x in res)),
            "tiles": "20 x 10 Galactic rects of 2 x 3 deg over l 60-100, b 20-50, r to 27.5, seed per tile"}


if __name__ == "__main__":
    main()
