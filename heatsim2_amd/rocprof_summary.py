"""Summaries of rocprofv3 CSV output (``-f csv``) for the ADI sweep kernels.

    python -m heatsim2_amd.rocprof_summary trace DIR
        per-kernel launch statistics (the *_kernel_stats.csv of --stats) and
        mean device time per sweep (x, y, z) from *_kernel_trace.csv
    python -m heatsim2_amd.rocprof_summary traffic --fetch DIR --write DIR \
            --copy-fetch DIR --copy-write DIR --copy-bytes B --cells N
        HBM bytes per sweep launch from FETCH_SIZE / WRITE_SIZE (KiB per
        dispatch), each scaled by what the same counter reads for a plain
        device copy of B bytes
    python -m heatsim2_amd.rocprof_summary kernels --trace DIR [--fetch DIR
            --write DIR --copy-fetch DIR --copy-write DIR --copy-bytes B]
            [--match SUBSTRING]
        every kernel (or those whose name holds SUBSTRING): launches, total
        and mean device time, and with the counter runs the copy-scaled HBM
        bytes per launch and in total

Kernels are attributed to sweeps by name (the strided kernel's third
template argument is its axis).  The last line of the traffic
summary is ``json: {...}`` with the per-sweep traffic in bytes; bench.py
reads it.
"""
import argparse
import collections
import csv
import glob
import json
import os
import re
import sys

BYTES_PER_CELL = {"x": 16.0, "y": 24.0, "z": 24.0}
# y from the x sweep's right-hand side (the strided kernel's fifth template
# argument): reads R_y and writes T**, no T^n
Y_FROM_X_BYTES = 16.0


def _one(directory, suffix):
    hits = sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit("no *%s under %s" % (suffix, directory))
    return hits[0]


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def short_name(name):
    name = name.replace("(anonymous namespace)::", "")
    name = re.sub(r"\(.*", "", name).replace("void ", "").replace("hs2::", "")
    return name.strip()


_STRIDED = re.compile(r"k_sweep_strided<[^,<>]+,\s*\d+,\s*(\d+)")
_STRIDED_RY = re.compile(r"k_sweep_strided<[^,<>]+,\s*\d+,\s*1,\s*(?:true|false),\s*true>")


def sweep_of(name):
    """'x', 'y' or 'z' for a sweep kernel's name, else None.  The strided
    kernel carries its axis (1: y, 0: z) as its third template argument."""
    if "k_sweep_x" in name:
        return "x"
    m = _STRIDED.search(name)
    if m:
        return "y" if m.group(1) == "1" else "z"
    return None


def algo_bytes_per_cell(sweep, name):
    """Algorithmic FP64 field bytes per cell of a sweep kernel instantiation."""
    if sweep == "y" and _STRIDED_RY.search(name):
        return Y_FROM_X_BYTES
    return BYTES_PER_CELL[sweep]


def sweep_kernels(dispatches):
    """dispatches: [(dispatch_id, kernel_name, value)] -> {"x"|"y"|"z":
    (kernel_name, [values])} of the most frequently launched instantiation
    per sweep (the source-free step kernel; the flash step's source variant
    is excluded)."""
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for _, name, value in sorted(dispatches):
        s = sweep_of(name)
        if s:
            per[s][name].append(value)
    return {s: max(by.items(), key=lambda kv: len(kv[1])) for s, by in per.items()}


def sweep_of_dispatches(dispatches):
    """dispatches -> {"x"|"y"|"z": [values]} (see sweep_kernels)."""
    return {s: v for s, (_, v) in sweep_kernels(dispatches).items()}


def trace_summary(directory, skip=0, count=0):
    """Per-kernel statistics of a kernel trace, and per sweep the mean device
    time of its most launched instantiation: over every launch, and with
    count > 0 also over launches skip .. skip + count - 1 in dispatch order
    (the timed run of a bench, apart from its warm-up and timing passes)."""
    lines = []
    stats = _rows(_one(directory, "_kernel_stats.csv"))
    lines.append("%-58s %7s %11s %11s %11s" % ("kernel", "calls", "avg_us", "min_us", "max_us"))
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:12]:
        lines.append("%-58s %7d %11.2f %11.2f %11.2f" % (
            short_name(r["Name"])[:58], int(r["Calls"]), float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3))
    trace = _rows(_one(directory, "_kernel_trace.csv"))
    disp = [(int(r["Dispatch_Id"]), r["Kernel_Name"],
             (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3) for r in trace]
    per = sweep_of_dispatches(disp)
    lines.append("")
    lines.append("per sweep: mean device time of the most launched instantiation")
    for s in ("x", "y", "z"):
        v = per.get(s, [])
        if v:
            lines.append("  sweep_%s: %d launches, mean %.2f us, min %.2f us" % (
                s, len(v), sum(v) / len(v), min(v)))
    if count > 0:
        lines.append("timed run: launches %d .. %d of each" % (skip, skip + count - 1))
        for s in ("x", "y", "z"):
            v = per.get(s, [])[skip:skip + count]
            if v:
                lines.append("  sweep_%s: %d launches, mean %.2f us, min %.2f us" % (
                    s, len(v), sum(v) / len(v), min(v)))
        per = {s: v[skip:skip + count] for s, v in per.items()}
    return "\n".join(lines), {s: (sum(v) / len(v) if v else None) for s, v in per.items()}


def _counter(directory, counter):
    rows = _rows(_one(directory, "_counter_collection.csv"))
    disp = collections.defaultdict(float)
    names = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        disp[d] += float(r["Counter_Value"])
        names[d] = r["Kernel_Name"]
    return [(d, names[d], v) for d, v in disp.items()]


def _copy_reading(directory, counter):
    """Mean reading of the most frequently launched kernel (the copy)."""
    disp = _counter(directory, counter)
    by = collections.defaultdict(list)
    for _, name, v in disp:
        by[name].append(v)
    name, vals = max(by.items(), key=lambda kv: len(kv[1]))
    return short_name(name), sum(vals) / len(vals), len(vals)


def traffic_summary(a):
    lines = []
    cname, cf, nf = _copy_reading(a.copy_fetch, "FETCH_SIZE")
    _, cw, nw = _copy_reading(a.copy_write, "WRITE_SIZE")
    ff = a.copy_bytes / (cf * 1024.0)
    fw = a.copy_bytes / (cw * 1024.0)
    lines.append("calibration: %s over %d B: FETCH_SIZE %.0f KiB (n=%d), WRITE_SIZE %.0f KiB (n=%d)"
                 % (cname, a.copy_bytes, cf, nf, cw, nw))
    lines.append("scale factors: fetch x%.3f, write x%.3f" % (ff, fw))
    fk = sweep_kernels(_counter(a.fetch, "FETCH_SIZE"))
    fetch = {s: v for s, (_, v) in fk.items()}
    write = sweep_of_dispatches(_counter(a.write, "WRITE_SIZE"))
    lines.append("")
    lines.append("%-8s %12s %12s %14s %14s %14s %10s" % (
        "sweep", "FETCH_KiB", "WRITE_KiB", "read_MB", "write_MB", "traffic_MB", "algo_MB"))
    out = {}
    for s in ("x", "y", "z"):
        if not fetch.get(s) or not write.get(s):
            continue
        f = sum(fetch[s]) / len(fetch[s])
        w = sum(write[s]) / len(write[s])
        rb, wb = f * 1024.0 * ff, w * 1024.0 * fw
        algo = algo_bytes_per_cell(s, fk[s][0]) * a.cells
        out["sweep_" + s] = rb + wb
        lines.append("%-8s %12.0f %12.0f %14.1f %14.1f %14.1f %10.1f" % (
            "sweep_" + s, f, w, rb / 1e6, wb / 1e6, (rb + wb) / 1e6, algo / 1e6))
    lines.append("")
    lines.append("json: " + json.dumps({"cells": a.cells, "fetch_scale": round(ff, 4),
                                        "write_scale": round(fw, 4),
                                        "traffic_bytes": {k: round(v) for k, v in out.items()}}))
    return "\n".join(lines)


def _per_kernel(dispatches):
    by = collections.defaultdict(list)
    for _, name, v in dispatches:
        by[short_name(name)].append(v)
    return by


def kernels_summary(a):
    trace = _rows(_one(a.trace, "_kernel_trace.csv"))
    dur = _per_kernel([(int(r["Dispatch_Id"]), r["Kernel_Name"],
                        (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
                       for r in trace])
    have_pmc = bool(a.fetch and a.write and a.copy_fetch and a.copy_write and a.copy_bytes)
    fetch = write = {}
    ff = fw = 1.0
    lines = []
    if have_pmc:
        _, cf, _ = _copy_reading(a.copy_fetch, "FETCH_SIZE")
        _, cw, _ = _copy_reading(a.copy_write, "WRITE_SIZE")
        ff = a.copy_bytes / (cf * 1024.0)
        fw = a.copy_bytes / (cw * 1024.0)
        fetch = _per_kernel(_counter(a.fetch, "FETCH_SIZE"))
        write = _per_kernel(_counter(a.write, "WRITE_SIZE"))
        lines.append("counter scale factors (device copy of %d B): fetch x%.3f, write x%.3f"
                     % (a.copy_bytes, ff, fw))
    hdr = "%-60s %7s %10s %10s" % ("kernel", "calls", "total_ms", "mean_us")
    if have_pmc:
        hdr += " %11s %11s %10s" % ("read_MB", "write_MB", "total_GB")
        if a.cells:
            hdr += " %8s" % "B/cell"
    lines.append(hdr)
    rows = sorted(dur.items(), key=lambda kv: -sum(kv[1]))
    for name, v in rows:
        if a.match and a.match not in name:
            continue
        line = "%-60s %7d %10.2f %10.2f" % (name[:60], len(v), sum(v) / 1e3, sum(v) / len(v))
        if have_pmc and name in fetch and name in write:
            rb = sum(fetch[name]) / len(fetch[name]) * 1024.0 * ff
            wb = sum(write[name]) / len(write[name]) * 1024.0 * fw
            line += " %11.1f %11.1f %10.2f" % (rb / 1e6, wb / 1e6, (rb + wb) * len(v) / 1e9)
            if a.cells:
                line += " %8.1f" % ((rb + wb) / a.cells)
        lines.append(line)
    return "\n".join(lines)


def read_traffic(path):
    """The json line of a committed traffic summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("json: "):
                    return json.loads(line[6:])
    except OSError:
        return None
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--skip", type=int, default=0, help="launches of each sweep before the timed run")
    t.add_argument("--count", type=int, default=0, help="launches of each sweep in the timed run")
    r = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        r.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    r.add_argument("--copy-bytes", type=int, required=True)
    r.add_argument("--cells", type=int, required=True)
    k = sub.add_parser("kernels")
    k.add_argument("--trace", required=True)
    k.add_argument("--cells", type=float, default=None,
                   help="mean cells one launch works on: adds HBM bytes per cell")
    for name in ("fetch", "write", "copy_fetch", "copy_write"):
        k.add_argument("--" + name.replace("_", "-"), dest=name, default=None)
    k.add_argument("--copy-bytes", type=int, default=None)
    k.add_argument("--match", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print(trace_summary(a.dir, a.skip, a.count)[0])
    elif a.cmd == "kernels":
        print(kernels_summary(a))
    else:
        print(traffic_summary(a))


if __name__ == "__main__":
    sys.exit(main())
