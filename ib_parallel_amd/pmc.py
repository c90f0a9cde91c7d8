"""Per-kernel HBM traffic and FP64 work from rocprofv3 counter runs, and a
markdown summary of a --kernel-trace --stats run.

Usage (on a machine with the GPU, each counter set in a run of its own):

    B="python3 -u bench.py --steps 4 --warmup 2 --no-cpu --batch 0 --adjoint-steps 0 \\
       --adjoint-fsi-steps 0 --window-steps 0 --pmc-copy-mb 256"
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- $B
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- $B
    rocprofv3 --pmc SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 SQ_INSTS_VALU_FMA_F64 \\
        SQ_INSTS_VALU_TRANS_F64 SQ_INSTS_VALU_MFMA_MOPS_F64 --output-format csv -d OUT/f64 -o run -- $B
    python -m ib_parallel_amd.pmc traffic OUT/fetch OUT/write --copy-mb 256 --f64 OUT/f64 \\
        --json profiles/pmc_traffic.json
    python -m ib_parallel_amd.pmc stats OUT/trace        # a --kernel-trace --stats run
    python -m ib_parallel_amd.pmc rates OUT/trace profiles/pmc_traffic.json   # GB/s and GFLOP/s
    python -m ib_parallel_amd.pmc ranges OUT/trace   # a --marker-trace run: the step's phases

FETCH_SIZE and WRITE_SIZE are read in KiB per dispatch.  The device-to-device
copy of known size that --pmc-copy-mb adds to the run calibrates them: the
copy reads and writes exactly copy_mb MiB, so bytes = counter * (copy bytes /
counter of the copy).  bench.py issues that copy right after its warm-up
steps, so only the dispatches after the copy (the timed steps; the one-time
set-up comes before it) are averaged.  FP64 operations per dispatch = 64 *
(ADD + MUL + TRANS + 2 FMA) wavefront instructions (all 64 lanes counted,
masked ones included) + 512 * SQ_INSTS_VALU_MFMA_MOPS_F64 (the matrix-core
FP64 flops, counted in units of 512) when that counter was collected.
"""
import argparse
import csv
import glob
import json
import os
import re
from collections import defaultdict


def read_counters(folder, counters):
    """{kernel name: {dispatch id: {counter: value}}} from rocprofv3
    counter_collection CSVs; the pseudo-counter "ns" is the dispatch's
    duration."""
    files = glob.glob(os.path.join(folder, "**", "*counter_collection*.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection CSV under {folder}")
    per = defaultdict(lambda: defaultdict(lambda: defaultdict(float)))
    for f in files:
        with open(f) as fh:
            for row in csv.DictReader(fh):
                c = row.get("Counter_Name", "").strip()
                if c not in counters:
                    continue
                name = row.get("Kernel_Name", "")
                key = int(row.get("Dispatch_Id") or row.get("Correlation_Id") or 0)
                per[name][key][c] += float(row["Counter_Value"])  # summed over dimensions of one dispatch
                if row.get("Start_Timestamp") and row.get("End_Timestamp"):
                    per[name][key]["ns"] = float(row["End_Timestamp"]) - float(row["Start_Timestamp"])
    return per


def copy_marker(runs):
    """Dispatch id of the calibration copy: the longest dispatch of a copy
    kernel (the runtime also copies small status blocks with the same kernel)."""
    best = None
    for k, disp in runs.items():
        if "copy" not in k.lower():
            continue
        for d, v in disp.items():
            if best is None or v.get("ns", 0.0) > best[1]:
                best = (d, v.get("ns", 0.0))
    return None if best is None else best[0]


def short(name):
    m = re.search(r"(k_[A-Za-z0-9_]+(?:<[^>]*>)?)", name)
    return m.group(1) if m else name.split("(")[0][:80]


def _mean(vals):
    return sum(vals) / len(vals) if vals else 0.0


def traffic(args):
    fetch = read_counters(args.fetch_dir, {"FETCH_SIZE"})
    write = read_counters(args.write_dir, {"WRITE_SIZE"})
    copy_kib = args.copy_mb * 1024.0
    cands = [k for k in fetch if "copy" in k.lower()]
    if not cands:
        raise SystemExit("no copy kernel in the counter run (pass --pmc-copy-mb to bench.py)")

    def copy_of(runs, counter):
        # the dispatch of a copy kernel whose counter is closest to the copy size
        best = None
        for k in cands:
            for d, v in runs.get(k, {}).items():
                x = v.get(counter, 0.0)
                if best is None or abs(x - copy_kib) < abs(best[2] - copy_kib):
                    best = (k, d, x)
        return best
    ck, cdf, cf = copy_of(fetch, "FETCH_SIZE")
    _, cdw, cw = copy_of(write, "WRITE_SIZE")
    sf = copy_kib / cf if cf > 0 else None
    sw = copy_kib / cw if cw > 0 else None
    f64 = read_counters(args.f64, {"SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64",
                                   "SQ_INSTS_VALU_TRANS_F64", "SQ_INSTS_VALU_MFMA_MOPS_F64"}) if args.f64 else {}
    # the marker dispatch of each run: its calibration copy; later dispatches are the timed steps
    f64_mark = copy_marker(f64) if f64 else None
    out = {"copy": {"kernel": ck, "bytes": copy_kib * 1024.0, "fetch_kib": cf, "write_kib": cw,
                    "scale_fetch": sf, "scale_write": sw},
           "window": "dispatches after the calibration copy (the timed steps)", "kernels": {}}
    names = defaultdict(list)
    for k in set(fetch) | set(write):
        if "copy" not in k.lower():
            names[short(k)].append(k)
    for sk, full in sorted(names.items()):
        fv = [v["FETCH_SIZE"] for k in full for d, v in fetch.get(k, {}).items() if d > cdf]
        wv = [v["WRITE_SIZE"] for k in full for d, v in write.get(k, {}).items() if d > cdw]
        if not fv and not wv:
            continue
        mf, mw = _mean(fv), _mean(wv)
        ent = {"dispatches": max(len(fv), len(wv)), "fetch_kib": mf, "write_kib": mw,
               "bytes_per_launch": (mf * (sf or 1.0) + mw * (sw or 1.0)) * 1024.0}
        if f64:
            ops = []
            for k in full:
                for d, v in f64.get(k, {}).items():
                    if f64_mark is None or d > f64_mark:
                        ops.append(64.0 * (v["SQ_INSTS_VALU_ADD_F64"] + v["SQ_INSTS_VALU_MUL_F64"]
                                           + v["SQ_INSTS_VALU_TRANS_F64"] + 2.0 * v["SQ_INSTS_VALU_FMA_F64"])
                                   + 512.0 * v["SQ_INSTS_VALU_MFMA_MOPS_F64"])
            ent["fp64_ops_per_launch"] = _mean(ops)
        out["kernels"][sk] = ent
    os.makedirs(os.path.dirname(os.path.abspath(args.json)), exist_ok=True)
    with open(args.json, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(json.dumps(out["copy"]))
    print("| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB/launch (calibrated) | FP64 Mop/launch |")
    print("|---|---|---|---|---|---|")
    for k, v in sorted(out["kernels"].items(), key=lambda kv: -kv[1]["bytes_per_launch"] * kv[1]["dispatches"]):
        fo = v.get("fp64_ops_per_launch")
        print(f"| {k} | {v['dispatches']} | {v['fetch_kib']:.1f} | {v['write_kib']:.1f} | "
              f"{v['bytes_per_launch'] / 1e6:.3f} | {'' if fo is None else f'{fo / 1e6:.3f}'} |")


def _trace_dispatches(trace_dir):
    files = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {trace_dir}")
    disp = []
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            disp.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["End_Timestamp"]) - float(r["Start_Timestamp"])))
    return sorted(disp)


def after_copy_durations(trace_dir):
    """{kernel name: [ns, ...]} of the dispatches after the calibration copy
    in a --kernel-trace run."""
    disp = _trace_dispatches(trace_dir)
    copies = [d for d in disp if "copy" in d[1].lower()]
    mark = max(copies, key=lambda d: d[2])[0] if copies else -1
    agg = defaultdict(list)
    for d, name, ns in disp:
        if d > mark and "copy" not in name.lower():
            agg[name].append(ns)
    return agg


def window_durations(trace_dir, first, until):
    """{kernel name: [ns, ...]} of the dispatches from the first one of the
    kernel named `first` (short name) up to, not including, the next one of
    the kernel named `until` (to the end of the run when there is none)."""
    disp = _trace_dispatches(trace_dir)
    i0 = next((i for i, d in enumerate(disp) if short(d[1]) == first), None)
    if i0 is None:
        raise SystemExit(f"no dispatch of {first}")
    i1 = next((i for i in range(i0 + 1, len(disp)) if until and short(disp[i][1]) == until), len(disp))
    agg = defaultdict(list)
    for _, name, ns in disp[i0:i1]:
        agg[name].append(ns)
    return agg


def rates(args):
    """Markdown table joining a --kernel-trace run (durations of the
    dispatches after the calibration copy) with a traffic JSON (calibrated
    HBM bytes and FP64 operations per launch): achieved HBM GB/s and FP64
    GFLOP/s per kernel."""
    with open(args.traffic_json) as fh:
        tr = json.load(fh)["kernels"]
    agg = defaultdict(list)
    for name, v in after_copy_durations(args.trace_dir).items():
        agg[short(name)] += v
    print("| kernel | calls | average us | HBM MB/launch | HBM GB/s | FP64 Mflop/launch | FP64 GFLOP/s |")
    print("|---|---|---|---|---|---|---|")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[: args.top]:
        us = sum(v) / len(v) / 1e3
        t = tr.get(k, {})
        b, f = t.get("bytes_per_launch"), t.get("fp64_ops_per_launch")
        bs = " | " if b is None else f"{b / 1e6:.3f} | {b / us / 1e3:.0f}"
        fs = " | " if f is None else f"{f / 1e6:.2f} | {f / us / 1e3:.0f}"
        print(f"| {k} | {len(v)} | {us:.2f} | {bs} | {fs} |")


def ranges(args):
    """Markdown table of the library's host-side ranges (roctx push/pop)
    from a rocprofv3 --marker-trace run: calls, total and mean duration per
    range name; with --after-copy only the ranges that start after the
    calibration copy's dispatch ends (needs --kernel-trace in the same run)."""
    files = glob.glob(os.path.join(args.trace_dir, "**", "*marker_api_trace.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no marker_api_trace.csv under {args.trace_dir}")
    t_copy = None
    if args.after_copy:
        kt = glob.glob(os.path.join(args.trace_dir, "**", "*kernel_trace.csv"), recursive=True)
        if kt:
            with open(kt[0]) as fh:
                copies = [(float(r["End_Timestamp"]) - float(r["Start_Timestamp"]), float(r["End_Timestamp"]))
                          for r in csv.DictReader(fh) if "copy" in r["Kernel_Name"].lower()]
            if copies:
                t_copy = max(copies)[1]
    agg = defaultdict(list)
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            name = r.get("Function") or r.get("Message") or r.get("Name") or ""
            t0, t1 = float(r["Start_Timestamp"]), float(r["End_Timestamp"])
            if t_copy is not None and t0 < t_copy:
                continue
            agg[name].append(t1 - t0)
    print("| range | calls | total ms | mean us |")
    print("|---|---|---|---|")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"| {k} | {len(v)} | {sum(v) / 1e6:.3f} | {sum(v) / len(v) / 1e3:.1f} |")


def stats(args):
    """Markdown table of a rocprofv3 --kernel-trace --stats run: the
    whole-process kernel_stats.csv, or with --after-copy the kernel trace
    restricted to the dispatches after the calibration copy (the timed
    steps of a bench.py run with --pmc-copy-mb), or with --first/--until
    the window between two kernels' dispatches (a phase of the run)."""
    if args.after_copy or args.first:
        agg = (window_durations(args.trace_dir, args.first, args.until) if args.first
               else after_copy_durations(args.trace_dir))
        rows = [{"Name": k, "Calls": len(v), "TotalDurationNs": sum(v), "AverageNs": sum(v) / len(v),
                 "MinNs": min(v), "MaxNs": max(v)} for k, v in agg.items()]
    else:
        files = glob.glob(os.path.join(args.trace_dir, "**", "*kernel_stats.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no kernel_stats.csv under {args.trace_dir}")
        with open(files[0]) as fh:
            rows = list(csv.DictReader(fh))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | calls | total ms | average us | min us | max us | share |")
    print("|---|---|---|---|---|---|---|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[: args.top]:
        t = float(r["TotalDurationNs"])
        print(f"| {short(r['Name'])} | {r['Calls']} | {t / 1e6:.3f} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | {t / tot:.3f} |")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--copy-mb", type=float, required=True)
    t.add_argument("--f64", default=None, help="folder of the FP64 instruction counter run")
    t.add_argument("--json", required=True)
    s = sub.add_parser("stats")
    s.add_argument("trace_dir")
    s.add_argument("--top", type=int, default=25)
    s.add_argument("--after-copy", action="store_true", help="only the dispatches after the calibration copy")
    s.add_argument("--first", default=None, help="only the dispatches from the first one of this kernel")
    s.add_argument("--until", default=None, help="... up to the next one of this kernel")
    r = sub.add_parser("rates")
    r.add_argument("trace_dir")
    r.add_argument("traffic_json")
    r.add_argument("--top", type=int, default=25)
    g = sub.add_parser("ranges")
    g.add_argument("trace_dir")
    g.add_argument("--after-copy", action="store_true", help="only the ranges after the calibration copy")
    args = ap.parse_args()
    {"traffic": traffic, "stats": stats, "rates": rates, "ranges": ranges}[args.cmd](args)


if __name__ == "__main__":
    main()
