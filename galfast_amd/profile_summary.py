"""Summaries of rocprofv3 runs of the catalog benchmark.

    python -m galfast_amd.profile_summary --kt DIR [--fetch DIR --write DIR
        --copy-fetch DIR --copy-write DIR] --out profiles --tag catalog_c1
        --program "python3 bench.py --steps 1 --warmup 0 --no-cpu --no-serve"

--kt is a ``rocprofv3 --kernel-trace --stats --output-format csv`` output
directory; --fetch / --write are ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE``
runs of the same command; --copy-fetch / --copy-write are the same counters
over ``bench.py --copy-calibration`` (a plain 1 GiB device copy), which fix
the factor between each counter and bytes.  Writes <tag>_kernel_stats.md,
<tag>_traffic.md (when the counter runs are given) and traffic.json (HBM
bytes per launch of each kernel, read by bench.py for roofline.traffic).

--pmc DIR... are further counter runs of the same command (any SQ_* /
GRBM_GUI_ACTIVE counters, at most 8 SQ counters a run); with SQ_WAVES and
SQ_INSTS_VALU among them they give <tag>_instruction_mix.md and valu.json:
the vector instructions each launch issues and, over the kernel-trace
duration, the SIMD cycles per vector instruction: 1,024 SIMDs (256 CUs x 4)
at 2.4 GHz times the launch's duration, over its vector instructions.  A
wave64 FP32 or FP64 FMA occupies a SIMD for 4 cycles (16 lanes a cycle), so
a launch near 4 cycles per instruction or below keeps the vector units
issuing all the time.
"""
import argparse
import csv
import glob
import json
import os
import re

COPY_BYTES = 1 << 30            # bench.py --copy-calibration copies 1 GiB
SIMD_GCYCLES = 256 * 4 * 2.4   # 1,024 SIMDs at 2.4 GHz: G SIMD-cycles per second
ROW_GROUPS = {"0": "FEH", "1": "BOND", "2": "PHOT", "3": "PERR", "4": "ASTRO"}


def short_name(full):
    """'void (anonymous namespace)::k_chain_rows<1, 1>(...)' -> 'k_chain_rows<PHOT>'."""
    s = full.replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    m = re.match(r"([A-Za-z_:][\w:]*)(<[^()]*>)?", s)
    if not m:
        return s
    name, targs = m.group(1), m.group(2)
    if name.endswith("k_chain_rows") or name.endswith("k_b_rows"):
        g = targs.strip("<>").split(",")[0].strip() if targs else ""
        return "%s<%s>" % (name, ROW_GROUPS.get(g, g))
    return name


def _csv(d, suffix):
    f = glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)
    if not f:
        raise SystemExit("no *%s under %s" % (suffix, d))
    with open(f[0]) as fh:
        return list(csv.DictReader(fh))


def kernel_stats(d):
    out = {}
    for r in _csv(d, "kernel_stats.csv"):
        k = short_name(r["Name"])
        e = out.setdefault(k, {"calls": 0, "total_ns": 0.0})
        e["calls"] += int(r["Calls"])
        e["total_ns"] += float(r["TotalDurationNs"])
    for e in out.values():
        e["avg_ns"] = e["total_ns"] / max(e["calls"], 1)
    return out


def counter_per_dispatch(d, counter, reduce=None):
    """{kernel: (mean over dispatches, dispatches)} of a counter summed (or,
    for GRBM_GUI_ACTIVE, the maximum) over the instances of one dispatch."""
    reduce = reduce or ("max" if counter.startswith("GRBM_") else "sum")
    acc = {}
    for r in _csv(d, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        k = short_name(r["Kernel_Name"])
        e = acc.setdefault(k, {})
        v = float(r["Counter_Value"])
        i = r["Dispatch_Id"]
        e[i] = (max(e[i], v) if reduce == "max" else e[i] + v) if i in e else v
    return {k: (sum(v.values()) / len(v), len(v)) for k, v in acc.items()}


def counters_in(d):
    return sorted({r["Counter_Name"] for r in _csv(d, "counter_collection.csv")})


def instruction_mix(a, ks, kt_cmd):
    per = {}
    pmc_cmds = []
    for d in a.pmc:
        names = counters_in(d)
        pmc_cmds.append("rocprofv3 --pmc %s --output-format csv -d %s -o pmc -- %s" % (" ".join(names), d, a.program))
        for c in names:
            for k, (v, n) in counter_per_dispatch(d, c).items():
                per.setdefault(k, {})[c] = v
    kernels = [k for k in sorted(ks, key=lambda k: -ks[k]["total_ns"]) if k in per and per[k].get("SQ_WAVES")]
    classes = [c for c in sorted({c for k in kernels for c in per[k]})
               if c.startswith("SQ_INSTS_") and "_FLOPS_" not in c]
    flops = [c for c in ("SQ_INSTS_VALU_FLOPS_FP32", "SQ_INSTS_VALU_FLOPS_FP64") if any(c in per[k] for k in kernels)]
    lines = ["# Instruction mix per launch: %s" % a.tag, "", "Commands:", ""] + ["* `%s`" % c for c in pmc_cmds] + \
            ["* durations: `%s`" % kt_cmd, "",
             "SIMD cycles per VALU instruction = 1,024 SIMDs x 2.4 GHz x mean kernel-trace duration / "
             "SQ_INSTS_VALU per launch (a wave64 FMA occupies a SIMD for 4 cycles: near 4 or below, the vector "
             "units issue all the time).  Counts are per wave (64 lanes).", "",
             ] + (["FP32 / FP64 TFLOP/s = SQ_INSTS_VALU_FLOPS_FP32 / _FP64 (floating-point operations per wave "
                   "instruction, an FMA counting 2) x 64 lanes x waves / mean duration; MI355X vector peaks: FP32 "
                   "157.3 TFLOP/s (packed FP32 math), FP64 78.6 TFLOP/s.", ""] if flops else []) + [
             "| kernel | waves / launch | VALU / launch | mean us | SIMD cycles / VALU | " +
             "".join("%s TFLOP/s | " % c[-4:] for c in flops) +
             " | ".join(c.replace("SQ_INSTS_", "") + " / wave" for c in classes) + " |",
             "|---|---|---|---|---|" + "---|" * (len(flops) + len(classes))]
    out = {"_simd_gcycles_per_s": SIMD_GCYCLES, "_program": a.program}
    for k in kernels:
        e = per[k]
        w = e["SQ_WAVES"]
        valu = e.get("SQ_INSTS_VALU", 0.0)
        avg = ks[k]["avg_ns"]
        cpi = SIMD_GCYCLES * avg / valu if valu else 0.0
        out[k] = {"waves": w, "valu_insts": valu, "avg_ns": avg, "simd_cycles_per_valu": cpi}
        tf = []
        for c in flops:
            t = e.get(c, 0.0) * 64.0 / avg / 1e3 if avg else 0.0   # TFLOP/s
            out[k][c[-4:].lower() + "_tflops"] = t
            tf.append("%.2f | " % t)
        lines.append("| `%s` | %.0f | %.3g | %.1f | %.1f | " % (k, w, valu, avg / 1e3, cpi) + "".join(tf) +
                     " | ".join("%.1f" % (e.get(c, 0.0) / w) for c in classes) + " |")
    with open(os.path.join(a.out, "%s_instruction_mix.md" % a.tag), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(a.out, "valu.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kt", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-fetch")
    ap.add_argument("--copy-write")
    ap.add_argument("--out", required=True)
    ap.add_argument("--tag", required=True)
    ap.add_argument("--program", required=True, help="the profiled program (after rocprofv3's --)")
    ap.add_argument("--workload", default="")
    ap.add_argument("--pmc", nargs="*", default=[])
    ap.add_argument("--kt-isolated", help="kernel trace of the same program with --overlap 0, whose durations "
                    "the instruction-mix table uses")
    ap.add_argument("--program-isolated", default="")
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)

    ks = kernel_stats(a.kt)
    total = sum(e["total_ns"] for e in ks.values())
    kt_cmd = "rocprofv3 --kernel-trace --stats --output-format csv -d %s -o kt -- %s" % (a.kt, a.program)
    lines = ["# Kernel trace: %s" % a.tag, "", "Command: `%s`" % kt_cmd, ""]
    if a.workload:
        lines += ["Workload: %s" % a.workload, ""]
    lines += ["| kernel | calls | total ms | avg us | share |", "|---|---|---|---|---|"]
    for k, e in sorted(ks.items(), key=lambda kv: -kv[1]["total_ns"]):
        lines.append("| `%s` | %d | %.2f | %.1f | %.1f%% |" % (k, e["calls"], e["total_ns"] / 1e6, e["avg_ns"] / 1e3,
                                                            100.0 * e["total_ns"] / total))
    with open(os.path.join(a.out, "%s_kernel_stats.md" % a.tag), "w") as f:
        f.write("\n".join(lines) + "\n")
    if a.pmc:
        if a.kt_isolated:
            kt_cmd = "rocprofv3 --kernel-trace --stats --output-format csv -d %s -o kt -- %s" % (
                a.kt_isolated, a.program_isolated or a.program)
            instruction_mix(a, kernel_stats(a.kt_isolated), kt_cmd)
        else:
            instruction_mix(a, ks, kt_cmd)

    if not (a.fetch and a.write and a.copy_fetch and a.copy_write):
        return
    fetch = counter_per_dispatch(a.fetch, "FETCH_SIZE")
    write = counter_per_dispatch(a.write, "WRITE_SIZE")
    cf = counter_per_dispatch(a.copy_fetch, "FETCH_SIZE")
    cw = counter_per_dispatch(a.copy_write, "WRITE_SIZE")
    copy = "__amd_rocclr_copyBuffer"
    f_fetch = COPY_BYTES / (cf[copy][0] * 1024.0)
    f_write = COPY_BYTES / (cw[copy][0] * 1024.0)
    traffic = {"_factors": {"fetch_bytes_per_KiB_count": f_fetch * 1024.0, "write_bytes_per_KiB_count": f_write * 1024.0,
                            "copy_fetch_KiB": cf[copy][0], "copy_write_KiB": cw[copy][0], "copy_bytes": COPY_BYTES},
               "_program": a.program}
    pmc = "rocprofv3 --pmc %s --output-format csv -d %s -o %s -- %s"
    lines = ["# HBM traffic per launch: %s" % a.tag, "", "Commands:", "",
             "* `%s`" % (pmc % ("FETCH_SIZE", a.fetch, "fetch", a.program)),
             "* `%s`" % (pmc % ("WRITE_SIZE", a.write, "write", a.program)),
             "* `%s`" % (pmc % ("FETCH_SIZE", a.copy_fetch, "cf", "python3 bench.py --copy-calibration")),
             "* `%s`" % (pmc % ("WRITE_SIZE", a.copy_write, "cw", "python3 bench.py --copy-calibration")),
             "", "Calibration: `bench.py --copy-calibration` copies 1 GiB with `hipMemcpy` "
             "(kernel `%s`): FETCH_SIZE reads %.0f KiB and WRITE_SIZE %.0f KiB for 1,048,576 KiB moved each way, "
             "so bytes = FETCH_SIZE x %.3f KiB and WRITE_SIZE x %.3f KiB." % (copy, cf[copy][0], cw[copy][0],
                                                                          f_fetch, f_write), "",
             "| kernel | dispatches | read MB / launch | written MB / launch | total MB / launch | avg us | "
             "GB/s |", "|---|---|---|---|---|---|---|"]
    for k in sorted(set(fetch) & set(write), key=lambda k: -ks.get(k, {"total_ns": 0})["total_ns"]):
        rb = fetch[k][0] * 1024.0 * f_fetch
        wb = write[k][0] * 1024.0 * f_write
        avg = ks.get(k, {}).get("avg_ns")
        traffic[k] = {"read_bytes": rb, "write_bytes": wb, "bytes": rb + wb, "dispatches": fetch[k][1],
                      "avg_ns": avg}
        gbs = (rb + wb) / avg if avg else 0.0
        lines.append("| `%s` | %d | %.1f | %.1f | %.1f | %s | %.0f |" % (
            k, fetch[k][1], rb / 1e6, wb / 1e6, (rb + wb) / 1e6, "%.1f" % (avg / 1e3) if avg else "-", gbs))
    with open(os.path.join(a.out, "%s_traffic.md" % a.tag), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(a.out, "traffic.json"), "w") as f:
        json.dump(traffic, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
