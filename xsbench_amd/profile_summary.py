"""Turns rocprofv3 CSV output of profile_driver.py runs into the short text
summaries kept with the repository.

    python3 xsbench_amd/profile_summary.py roofline DIR --command "..." > summary.txt
    python3 xsbench_amd/profile_summary.py stats DIR/trace --command "..." > summary.txt

roofline: DIR holds one profile_driver.py workload -- its meta.json (from the
--kernel-trace run) and the counter passes rd/ (TCC_EA0_RDREQ, _32B,
TCC_BUBBLE), wr/ (TCC_EA0_WRREQ, _64B, TCC_HIT, TCC_MISS), dram/
(TCC_EA0_RDREQ_DRAM, TCC_READ_SECTORS, TCC_REQ, TCC_READ) and optionally sq/
and tcp/ (TCP_TOTAL_CACHE_ACCESSES, TA_FLAT_READ_WAVEFRONTS, TA_TA_BUSY,
GRBM_GUI_ACTIVE: 64-byte L1 accesses per vector load instruction and the
share of cycles the 256 texture-address units were busy).
HBM bytes per launch count every request to the memory side by its size:
reads 32 B (RDREQ_32B), 128 B (TCC_BUBBLE) or 64 B (the rest), scaled by the
factor that makes the same formula give the plain copy kernel's known bytes;
writes 64 B (WRREQ_64B) or 32 B (the rest).  The kernel's time is the trace's
average duration; the compulsory bytes are the distinct 128-byte lines its
lookups read (xsb_hip_lines_*).
stats: the kernel-trace statistics table.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

HBM_PEAK = 8.0e12


def _one(pattern):
    files = glob.glob(pattern, recursive=True)
    if not files:
        sys.exit(f"nothing matches {pattern}")
    return files[0]


def counters(d, kernel):
    """{counter: mean over dispatches of `kernel` (a name prefix)}, dispatches"""
    vals = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(_one(os.path.join(d, "**", "*counter_collection.csv")))):
        name = r["Kernel_Name"]
        if name.startswith(kernel) or name.startswith("void xsb::" + kernel) or \
                ("xsb::" + kernel + "<") in name or name.split("(")[0].endswith(kernel):
            vals[r["Counter_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
    if not vals:
        sys.exit(f"no counters of {kernel} under {d}")
    out = {c: sum(v.values()) / len(v) for c, v in vals.items()}
    return out, max(len(v) for v in vals.values())


def trace_ms(d, kernel):
    """average duration (ms) of `kernel` in the kernel-trace statistics"""
    for r in csv.DictReader(open(_one(os.path.join(d, "**", "*kernel_stats.csv")))):
        name = r["Name"]
        if ("xsb::" + kernel) in name or name.split("(")[0].endswith(kernel):
            return float(r["AverageNs"]) / 1e6, int(r["Calls"])
    sys.exit(f"{kernel} not in the kernel statistics under {d}")


def read_bytes(c):
    r, r32, b128 = c["TCC_EA0_RDREQ_sum"], c["TCC_EA0_RDREQ_32B_sum"], c["TCC_BUBBLE_sum"]
    return 32.0 * r32 + 128.0 * b128 + 64.0 * (r - r32 - b128)


def write_bytes(c):
    w, w64 = c["TCC_EA0_WRREQ_sum"], c["TCC_EA0_WRREQ_64B_sum"]
    return 64.0 * w64 + 32.0 * (w - w64)


def roofline(args):
    d = args.dir
    meta = json.load(open(os.path.join(d, "meta.json")))
    k = meta["kernel"]
    ms, calls = trace_ms(os.path.join(d, "trace"), k)
    rd, nd = counters(os.path.join(d, "rd"), k)
    wr, _ = counters(os.path.join(d, "wr"), k)
    crd, _ = counters(os.path.join(d, "rd"), "k_copy")
    cwr, _ = counters(os.path.join(d, "wr"), "k_copy")
    cb = meta["copy_bytes"]
    rf = cb / read_bytes(crd)
    wf = cb / write_bytes(cwr)
    rb = read_bytes(rd) * rf
    wb = write_bytes(wr) * wf
    hbm = rb + wb
    comp = meta["compulsory_bytes_per_launch"]
    out = [f"# HBM traffic and roofline per launch of {k}: workload {meta['workload']}, "
           f"{meta['precision']}" + (f", schedule k={meta['kernel_id']}" if meta["method"] == "event" else ""),
           f"# commands: {args.command}",
           f"# launches averaged: {nd} (counters), {calls} (trace)",
           f"# read bytes = 32 RDREQ_32B + 128 TCC_BUBBLE + 64 (RDREQ - RDREQ_32B - TCC_BUBBLE), "
           f"x {rf:.4f} so that the copy kernel reads its {cb} bytes",
           f"# write bytes = 64 WRREQ_64B + 32 (WRREQ - WRREQ_64B), x {wf:.4f} by the same copy",
           f"# the copy's reads are {cb / crd['TCC_EA0_RDREQ_sum']:.1f} bytes per RDREQ request "
           f"(one 128-byte line each, with no 32-byte or TCC_BUBBLE requests), so FETCH_SIZE's "
           f"64-byte weighting halves them; the lookup kernels' requests are of the same kinds",
           f"# copy kernel per launch: RDREQ {crd['TCC_EA0_RDREQ_sum']:.0f}, RDREQ_32B "
           f"{crd['TCC_EA0_RDREQ_32B_sum']:.0f}, TCC_BUBBLE {crd['TCC_BUBBLE_sum']:.0f}; WRREQ "
           f"{cwr['TCC_EA0_WRREQ_sum']:.0f}, WRREQ_64B {cwr['TCC_EA0_WRREQ_64B_sum']:.0f}",
           f"# {k} per launch: RDREQ {rd['TCC_EA0_RDREQ_sum']:.0f}, RDREQ_32B "
           f"{rd['TCC_EA0_RDREQ_32B_sum']:.0f}, TCC_BUBBLE {rd['TCC_BUBBLE_sum']:.0f}; WRREQ "
           f"{wr['TCC_EA0_WRREQ_sum']:.0f}, WRREQ_64B {wr['TCC_EA0_WRREQ_64B_sum']:.0f}",
           f"kernel_ms = {ms:.4f}",
           f"lookups_per_launch = {meta['lookups'] / (34 if meta['method'] == 'history' else 1):.0f}",
           f"read_bytes_per_launch = {rb:.6e}",
           f"write_bytes_per_launch = {wb:.6e}",
           f"hbm_bytes_per_launch = {hbm:.6e}",
           f"hbm_gbs = {hbm / (ms * 1e-3) / 1e9:.1f}",
           f"compulsory_bytes_per_launch = {comp:.6e}",
           f"roofline_frac = {comp / (ms * 1e-3) / HBM_PEAK:.4f}",
           f"traffic_over_compulsory = {hbm / comp:.3f}",
           f"l2_hit_rate = {wr['TCC_HIT_sum'] / (wr['TCC_HIT_sum'] + wr['TCC_MISS_sum']):.4f}"]
    if "balg_bytes_per_lookup" in meta:
        alg = meta["balg_bytes_per_lookup"] * meta["lookups"]
        out.append(f"algorithmic_bytes_per_launch = {alg:.6e}")
        out.append(f"gather_efficiency = {alg / hbm:.4f}")
    dram_dir = os.path.join(d, "dram")
    if os.path.isdir(dram_dir):
        dr, _ = counters(dram_dir, k)
        out.append(f"dram_share_of_read_requests = "
                   f"{dr['TCC_EA0_RDREQ_DRAM_sum'] / max(rd['TCC_EA0_RDREQ_sum'], 1):.4f}")
        out.append(f"l2_read_sectors_bytes = {32.0 * dr['TCC_READ_SECTORS_sum']:.6e}")
    tcp_dir = os.path.join(d, "tcp")
    if os.path.isdir(tcp_dir):
        tc, _ = counters(tcp_dir, k)
        xcd_cycles = tc["GRBM_GUI_ACTIVE"] / 8.0  # the counter sums the 8 XCDs
        out.append(f"l1_accesses_per_vmem_instruction = "
                   f"{tc['TCP_TOTAL_CACHE_ACCESSES_sum'] / max(tc['TA_FLAT_READ_WAVEFRONTS_sum'], 1):.1f}")
        out.append(f"ta_busy = {tc['TA_TA_BUSY_sum'] / (256.0 * max(xcd_cycles, 1)):.3f}")
    sq_dir = os.path.join(d, "sq")
    if os.path.isdir(sq_dir):
        sq, _ = counters(sq_dir, k)
        out.append(f"valu_insts = {sq['SQ_INSTS_VALU']:.6e}")
        out.append(f"smem_insts = {sq['SQ_INSTS_SMEM']:.6e}")
        out.append(f"vmem_read_insts = {sq['SQ_INSTS_VMEM_RD']:.6e}")
        out.append(f"waves = {sq['SQ_WAVES']:.0f}")
        out.append(f"wave_time_waiting = {sq['SQ_WAIT_ANY'] / max(sq['SQ_WAVE_CYCLES'], 1):.4f}")
        out.append(f"wave_time_in_valu = {sq['SQ_ACTIVE_INST_VALU'] / max(sq['SQ_WAVE_CYCLES'], 1):.4f}")
    print("\n".join(out))


def stats(args):
    rows = list(csv.DictReader(open(_one(os.path.join(args.dir, "**", "*kernel_stats.csv")))))
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
    t = sub.add_parser("roofline")
    t.add_argument("dir")
    t.add_argument("--command", default="")
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", default="")
    args = ap.parse_args()
    roofline(args) if args.cmd == "roofline" else stats(args)


if __name__ == "__main__":
    main()
