"""Summaries of rocprofv3 output for the committed profiles/ directory.

    python -m espec_amd.profiling stats DIR OUT.txt --command "..."
        kernel table (calls, mean/min/max duration, share) from
        DIR/*kernel_stats.csv

    python -m espec_amd.profiling pmc FETCH_DIR WRITE_DIR OUT.json --command "..."
        HBM bytes per dispatch per kernel from two --pmc runs (FETCH_SIZE,
        then WRITE_SIZE, KiB per dispatch), each scaled by what the counter
        reads for the copy kernel (copy_probe moves a known number of bytes:
        its argument n float4 read and n float4 written per launch, given by
        --copy-bytes, default 1 GiB each way)

bench.py reads the JSON's per-kernel "traffic_bytes" for its roofline.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys


def _one(path_glob):
    files = sorted(glob.glob(path_glob, recursive=True))
    if not files:
        raise SystemExit(f"no file matches {path_glob}")
    return files[0]


def short_name(name):
    """Template-qualified kernel name without namespaces and argument list."""
    n = name.strip().replace("(anonymous namespace)::", "").replace("espec::", "")
    n = re.sub(r"\(.*\)$", "", n)
    return re.sub(r"^void ", "", n)


def stats(args):
    path = _one(os.path.join(args.dir, "**", "*kernel_stats.csv"))
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = [f"# rocprofv3 --kernel-trace --stats summary", f"# command: {args.command}", "",
             f"{'kernel':64s} {'calls':>7s} {'mean_us':>9s} {'min_us':>8s} {'max_us':>8s} "
             f"{'total_ms':>9s} {'share%':>7s}"]
    for r in rows[: args.top]:
        lines.append(f"{short_name(r['Name'])[:64]:64s} {int(r['Calls']):7d} "
                     f"{float(r['AverageNs']) / 1e3:9.3f} {float(r['MinNs']) / 1e3:8.3f} "
                     f"{float(r['MaxNs']) / 1e3:8.3f} {float(r['TotalDurationNs']) / 1e6:9.3f} "
                     f"{float(r['Percentage']):7.2f}")
    open(args.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def _counter_means(d, counter):
    path = _one(os.path.join(d, "**", "*counter_collection.csv"))
    acc = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        k = short_name(r["Kernel_Name"])
        s, c = acc.get(k, (0.0, 0))
        acc[k] = (s + float(r["Counter_Value"]), c + 1)
    return {k: (s / c, c) for k, (s, c) in acc.items()}


def pmc(args):
    fetch = _counter_means(args.fetch_dir, "FETCH_SIZE")
    write = _counter_means(args.write_dir, "WRITE_SIZE")
    probe = [k for k in fetch if k.startswith("copy_probe")]
    if not probe or probe[0] not in write:
        raise SystemExit("copy_probe dispatches missing: run bench.py (it calls the copy kernel)")
    p = probe[0]
    fs = args.copy_bytes / (fetch[p][0] * 1024.0)
    ws = args.copy_bytes / (write[p][0] * 1024.0)
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        kernels[k] = {
            "dispatches": fetch[k][1],
            "fetch_kib_per_dispatch": fetch[k][0],
            "write_kib_per_dispatch": write[k][0],
            "traffic_bytes": fetch[k][0] * 1024.0 * fs + write[k][0] * 1024.0 * ws,
        }
    out = {
        "command": args.command,
        "counters": "FETCH_SIZE and WRITE_SIZE (KiB per dispatch), one rocprofv3 --pmc run each",
        "calibration": {"kernel": p, "bytes_each_way": args.copy_bytes,
                        "fetch_scale": fs, "write_scale": ws},
        "kernels": kernels,
    }
    json.dump(out, open(args.out, "w"), indent=1)
    for k, v in kernels.items():
        print(f"{k[:70]:70s} {v['traffic_bytes'] / 1e6:10.3f} MB/dispatch")
    print(f"scales: fetch {fs:.3f} write {ws:.3f}")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m espec_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("out")
    a.add_argument("--command", required=True)
    a.add_argument("--top", type=int, default=12)
    b = sub.add_parser("pmc")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("out")
    b.add_argument("--command", required=True)
    b.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    args = ap.parse_args(argv)
    stats(args) if args.cmd == "stats" else pmc(args)


if __name__ == "__main__":
    main(sys.argv[1:])
