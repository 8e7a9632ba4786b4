"""Turns rocprofv3 CSV output of profile_driver.py runs into the short text
summaries kept with the repository.

    python3 xsbench_amd/profile_summary.py traffic --fetch D1 --write D2 --hit D3 \\
        --meta D1/meta.json --kernel k_lookup_ids --command "..." > summary.txt
    python3 xsbench_amd/profile_summary.py stats D --command "..." > summary.txt

traffic: per-dispatch FETCH_SIZE and WRITE_SIZE (KiB) of the named kernel,
each scaled by the factor that makes the same counter read 2^30 bytes for the
plain copy kernel (k_copy, which reads and writes exactly copy_bytes), plus
the L2 hit rate TCC_HIT / (TCC_HIT + TCC_MISS) and the gather efficiency
(algorithmic bytes / HBM bytes).  stats: the kernel-trace statistics table.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys


def _counter_csv(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        sys.exit(f"no counter_collection.csv under {d}")
    return files[0]


def per_dispatch(d, kernel, counter):
    """mean over dispatches of `kernel` (substring of the name) of `counter`"""
    vals = collections.defaultdict(float)
    for r in csv.DictReader(open(_counter_csv(d))):
        if kernel in r["Kernel_Name"] and r["Counter_Name"] == counter:
            vals[r["Dispatch_Id"]] += float(r["Counter_Value"])
    if not vals:
        sys.exit(f"{counter} of {kernel} not found under {d}")
    return sum(vals.values()) / len(vals), len(vals)


def traffic(args):
    meta = json.load(open(args.meta))
    k = meta["kernel_id"]
    copy_bytes = meta["copy_bytes"]
    f_lookup, nd = per_dispatch(args.fetch, args.kernel, "FETCH_SIZE")
    f_copy, _ = per_dispatch(args.fetch, "k_copy", "FETCH_SIZE")
    w_lookup, _ = per_dispatch(args.write, args.kernel, "WRITE_SIZE")
    w_copy, _ = per_dispatch(args.write, "k_copy", "WRITE_SIZE")
    fetch_factor = copy_bytes / (f_copy * 1024.0)
    write_factor = copy_bytes / (w_copy * 1024.0)
    fetch_b = f_lookup * 1024.0 * fetch_factor
    write_b = w_lookup * 1024.0 * write_factor
    hbm = fetch_b + write_b
    out = [f"# HBM traffic per dispatch of {args.kernel}, schedule k={k}, workload "
           f"{meta['workload']}, {meta['precision']}",
           f"# command: {args.command}",
           f"# dispatches averaged: {nd}",
           f"# calibration: k_copy of {copy_bytes} B reads FETCH_SIZE {f_copy:.0f} KiB "
           f"(factor {fetch_factor:.4f}) and WRITE_SIZE {w_copy:.0f} KiB (factor {write_factor:.4f})",
           f"k{k}.fetch_bytes_per_launch = {fetch_b:.6e}",
           f"k{k}.write_bytes_per_launch = {write_b:.6e}",
           f"k{k}.hbm_bytes_per_launch = {hbm:.6e}"]
    if args.hit:
        hit, _ = per_dispatch(args.hit, args.kernel, "TCC_HIT_sum")
        miss, _ = per_dispatch(args.hit, args.kernel, "TCC_MISS_sum")
        out.append(f"k{k}.l2_hit_rate = {hit / (hit + miss):.4f}")
    if "balg_bytes_per_lookup" in meta:
        alg = meta["balg_bytes_per_lookup"] * meta["lookups"]
        out.append(f"k{k}.algorithmic_bytes_per_launch = {alg:.6e}")
        out.append(f"k{k}.gather_efficiency = {alg / hbm:.4f}")
        out.append(f"k{k}.lookup_ms = {meta['lookup_ms']:.4f}")
        out.append(f"k{k}.hbm_gbs = {hbm / (meta['lookup_ms'] * 1e-3) / 1e9:.1f}")
    print("\n".join(out))


def stats(args):
    files = glob.glob(os.path.join(args.dir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        sys.exit(f"no kernel_stats.csv under {args.dir}")
    rows = list(csv.DictReader(open(files[0])))
    print(f"# rocprofv3 kernel-trace statistics\n# command: {args.command}")
    print(f"{'kernel':60s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}")
    for r in rows:
        name = r["Name"]
        name = name if len(name) <= 60 else name[:57] + "..."
        print(f"{name:60s} {int(r['Calls']):6d} {float(r['AverageNs']) / 1e3:10.1f} "
              f"{float(r['TotalDurationNs']) / 1e6:10.3f} {float(r['Percentage']):6.2f}")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--hit", default=None)
    t.add_argument("--meta", required=True)
    t.add_argument("--kernel", default="k_lookup_ids")
    t.add_argument("--command", default="")
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", default="")
    args = ap.parse_args()
    traffic(args) if args.cmd == "traffic" else stats(args)


if __name__ == "__main__":
    main()
