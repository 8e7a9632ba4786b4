"""Fixed workload for rocprofv3 runs of the lookup path.

Runs one workload a few times plus a plain device copy of known size (a
check of the byte counters: the copy reads and writes exactly COPY_BYTES),
and writes what the summary needs as JSON: the lookup kernel's name and HIP
event time, its compulsory bytes per launch (the distinct 128-byte lines its
lookups read, xsb_hip_lines_*) and, for event runs, the algorithmic bytes.

    rocprofv3 --pmc TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum --output-format csv -d OUT -o run -- \\
        python3 xsbench_amd/profile_driver.py --workload large_unionized --kernel 6 --meta OUT/meta.json
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import xsbench_amd as X  # noqa: E402

WORKLOADS = {
    "large_unionized": ("large", "unionized", "event", None),
    "large_hash": ("large", "hash", "event", None),
    "small_unionized": ("small", "unionized", "event", None),
    "xl45k_hash_history": ("XL", "hash", "history", 45000),
    # 512 batches, sizes log-uniform in [1e3, 2e6] (xsbench_make_batches seed 2024)
    "large_unionized_batched": ("large", "unionized", "batched", None),
}

COPY_BYTES = 1 << 30


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="large_unionized", choices=sorted(WORKLOADS))
    ap.add_argument("--kernel", type=int, default=6)
    ap.add_argument("--precision", default="fp64", choices=sorted(X.PRECISIONS))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--lookups", type=int, default=17_000_000)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--meta", default=None, help="write workload facts as JSON here")
    ap.add_argument("--plan", default="2048,16,2,3",
                    help="batched: tile,stage,sort,mapping (fixed, so no tuning runs are traced)")
    args = ap.parse_args()
    size, grid, method, gp = WORKLOADS[args.workload]
    inp = X.make_inputs(size, grid, "event", gridpoints=gp)
    sd = X.grid_init(inp)
    prec = X.PRECISIONS[args.precision]
    dev = X.Device(inp, sd, build_fp32=prec == X.XS_FP32, build_bf16=prec == X.XS_BF16)
    ms = []
    if method == "batched":
        sizes, seeds, thr = X.make_batches(2024, 512, 1000, 2_000_000)
        tile, stage, sort, mapping = (int(v) for v in args.plan.split(","))
    for _ in range(args.reps):
        if method == "event":
            r = dev.event(args.lookups, 0, kernel_id=args.kernel, precision=prec)
        elif method == "batched":
            r, _, _ = dev.batched(sizes, seeds, thr, precision=prec, mode=0, tile=tile,
                                  stage=stage, sort=sort, mapping=mapping)
        else:
            r = dev.history(args.particles, 34, kernel_id=args.kernel, precision=prec)
        ms.append(r["ms_lookup"])
        print(f"[profile] {args.workload} k={args.kernel} {args.precision}: {r['ms_total']:.3f} ms, "
              f"lookup kernels {r['ms_lookup']:.3f} ms, hash {r['hash']}", flush=True)
    copy_ms = dev.copy_probe(COPY_BYTES, reps=args.reps)
    print(f"[profile] copy of {COPY_BYTES} bytes: {copy_ms:.3f} ms "
          f"({2 * COPY_BYTES / copy_ms / 1e6:.0f} GB/s)", flush=True)
    if args.meta:
        meta = {"workload": args.workload, "kernel_id": args.kernel, "precision": args.precision,
                "copy_bytes": COPY_BYTES, "lookup_ms": min(ms), "method": method}
        if method == "event":
            meta["kernel"] = {0: "k_event_fused", 1: "k_lookup_list"}.get(args.kernel, "k_lookup_ids")
            meta["lookups"] = args.lookups
            meta["balg_bytes_per_lookup"] = dev.balg(args.lookups, precision=prec)
            comp = dev.lines_event(args.lookups, kernel_id=args.kernel, precision=prec)
        elif method == "batched":
            meta["kernel"] = "k_bsorted"
            meta["lookups"] = int(sizes.sum())
            comp = dev.lines_batched(sizes, seeds, thr, precision=prec)
        else:
            meta["kernel"] = "k_hround"
            meta["lookups"] = args.particles * 34
            comp = dev.lines_history(args.particles, 34, precision=prec)
        meta["compulsory_bytes_per_launch"] = comp["bytes"] / comp["launches"]
        meta["compulsory"] = comp
        with open(args.meta, "w") as f:
            json.dump(meta, f)
    dev.close()


if __name__ == "__main__":
    main()
