"""Fixed workload for hardware-counter runs (rocprofv3 --pmc).

Runs the lookup path a few times on one workload plus a plain device copy of
known size, so that per-dispatch counters of the lookup kernel can be scaled
by what the same counters read for the copy (bytes moved = 2 x copy size).

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -o run -- \
        python3 -m xsbench_amd.profile_driver --workload large_unionized --kernel 6
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import xsbench_amd as X  # noqa: E402

WORKLOADS = {
    "large_unionized": ("large", "unionized", "event", None),
    "large_hash": ("large", "hash", "event", None),
    "small_unionized": ("small", "unionized", "event", None),
    "xl45k_hash_history": ("XL", "hash", "history", 45000),
}

COPY_BYTES = 1 << 30


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="large_unionized", choices=sorted(WORKLOADS))
    ap.add_argument("--kernel", type=int, default=6)
    ap.add_argument("--precision", default="fp64", choices=sorted(X.PRECISIONS))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--lookups", type=int, default=17_000_000)
    ap.add_argument("--particles", type=int, default=100_000)
    args = ap.parse_args()
    size, grid, method, gp = WORKLOADS[args.workload]
    inp = X.make_inputs(size, grid, "event", gridpoints=gp)
    sd = X.grid_init(inp)
    prec = X.PRECISIONS[args.precision]
    dev = X.Device(inp, sd, build_fp32=prec == X.XS_FP32, build_bf16=prec == X.XS_BF16)
    for _ in range(args.reps):
        if method == "event":
            r = dev.event(args.lookups, 0, kernel_id=args.kernel, precision=prec)
        else:
            r = dev.history(args.particles, 34, kernel_id=args.kernel, precision=prec)
        print(f"[profile] {args.workload} k={args.kernel} {args.precision}: {r['ms_total']:.3f} ms, "
              f"lookup kernels {r['ms_lookup']:.3f} ms, hash {r['hash']}", flush=True)
    ms = dev.copy_probe(COPY_BYTES, reps=args.reps)
    print(f"[profile] copy of {COPY_BYTES} bytes: {ms:.3f} ms ({2 * COPY_BYTES / ms / 1e6:.0f} GB/s)",
          flush=True)
    dev.close()


if __name__ == "__main__":
    main()
