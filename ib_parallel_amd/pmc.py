"""HBM traffic per kernel launch from two rocprofv3 counter runs.

Usage (on a machine with the GPU, each counter in a run of its own):

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- \\
        python bench.py --steps 2 --warmup 0 --no-cpu --pmc-copy-mb 256
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- \\
        python bench.py --steps 2 --warmup 0 --no-cpu --pmc-copy-mb 256
    python -m ib_parallel_amd.pmc OUT/fetch OUT/write --copy-mb 256 --json profiles/pmc_traffic.json

FETCH_SIZE and WRITE_SIZE are read in KiB per dispatch.  The device-to-device
copy of known size that --pmc-copy-mb adds to the same run calibrates them:
the copy reads and writes exactly copy_mb MiB, so bytes = counter * (copy
bytes / counter of the copy).  The JSON gives, per kernel, the mean counters
per dispatch, the two scale factors and the calibrated HBM bytes per launch.
"""
import argparse
import csv
import glob
import json
import os
import re
from collections import defaultdict


def read_counters(folder, counter):
    """{kernel name: [value per dispatch]} from a rocprofv3 counter_collection CSV."""
    files = glob.glob(os.path.join(folder, "**", "*counter_collection*.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection CSV under {folder}")
    per = defaultdict(lambda: defaultdict(float))
    for f in files:
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name", "").strip() != counter:
                    continue
                name = row.get("Kernel_Name", "")
                key = (row.get("Dispatch_Id") or row.get("Correlation_Id") or str(len(per[name])))
                per[name][key] += float(row["Counter_Value"])  # summed over dimensions of one dispatch
    return {k: list(v.values()) for k, v in per.items()}


def short(name):
    m = re.search(r"(k_[A-Za-z0-9_]+(?:<[^>]*>)?)", name)
    return m.group(1) if m else name.split("(")[0][:80]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("fetch_dir")
    ap.add_argument("write_dir")
    ap.add_argument("--copy-mb", type=float, required=True)
    ap.add_argument("--json", required=True)
    args = ap.parse_args()
    fetch = read_counters(args.fetch_dir, "FETCH_SIZE")
    write = read_counters(args.write_dir, "WRITE_SIZE")
    copy_kib = args.copy_mb * 1024.0
    # the copy: the kernel whose mean FETCH_SIZE is closest to the copy size
    cands = [k for k in fetch if "copy" in k.lower()]
    if not cands:
        raise SystemExit("no copy kernel in the counter run (pass --pmc-copy-mb to bench.py)")
    ck = min(cands, key=lambda k: abs(max(fetch[k]) - copy_kib))
    cf = max(fetch[ck])
    cw = max(write.get(ck, [0.0]))
    sf = copy_kib / cf if cf > 0 else None
    sw = copy_kib / cw if cw > 0 else None
    out = {"copy": {"kernel": ck, "bytes": copy_kib * 1024.0, "fetch_kib": cf, "write_kib": cw,
                    "scale_fetch": sf, "scale_write": sw}, "kernels": {}}
    for k in sorted(set(fetch) | set(write)):
        if k == ck:
            continue
        fv, wv = fetch.get(k, []), write.get(k, [])
        mf = sum(fv) / len(fv) if fv else 0.0
        mw = sum(wv) / len(wv) if wv else 0.0
        out["kernels"][short(k)] = {
            "dispatches": max(len(fv), len(wv)), "fetch_kib": mf, "write_kib": mw,
            "bytes_per_launch": (mf * (sf or 1.0) + mw * (sw or 1.0)) * 1024.0}
    os.makedirs(os.path.dirname(os.path.abspath(args.json)), exist_ok=True)
    with open(args.json, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(json.dumps(out["copy"]))
    for k, v in sorted(out["kernels"].items(), key=lambda kv: -kv[1]["bytes_per_launch"])[:12]:
        print(f"{k:60s} {v['dispatches']:6d} {v['bytes_per_launch'] / 1e6:12.3f} MB/launch")


if __name__ == "__main__":
    main()
