"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs only where the reference is mounted: it loads oracle/_ref/libxsbench_ref.so
(built from openmp-threading/*.c by oracle/ref.mk) and records, through the
reference's own functions:

* hashes.json   -- verification sums of run_event_based_simulation /
                   run_history_based_simulation / _optimization_1 per config
* digests.json  -- FNV-1a digests of every SimulationData array per config
* lookups_<cfg>.npz -- per-lookup (id, energy, mat, UEG index, macro_xs[5],
                   argmax) for the first lookups and a spread of later ids,
                   computed by fast_forward_LCG / LCG_random_double / pick_mat
                   / grid_search / calculate_macro_xs
* micro_<cfg>.npz -- calculate_micro_xs for every nuclide of 2 lookups per
                   material
* history_<cfg>.npz -- per-particle verification sums of the history loop
                   (Simulation.c:141-237 driven step by step through the
                   reference's functions)
* sampling.npz  -- LCG / fast-forward / pick_mat vectors

Usage:  python tests/golden/make_golden.py [--quick]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# (name, size, grid, gridpoints)
DATA_CFGS = [
    ("small_unionized", "small", "unionized", None),
    ("small_hash", "small", "hash", None),
    ("small_nuclide", "small", "nuclide", None),
    ("large_unionized", "large", "unionized", None),
    ("large_hash", "large", "hash", None),
    ("xl45k_hash", "XL", "hash", 45000),
]

# (key, data cfg, method, lookups, particles, driver)
RUN_CFGS = [
    ("event_small_unionized_17M", "small_unionized", "event", None, None, "event"),
    ("event_small_hash_17M", "small_hash", "event", None, None, "event"),
    ("event_small_nuclide_17M", "small_nuclide", "event", None, None, "event"),
    ("event_small_unionized_1M", "small_unionized", "event", 1000000, None, "event"),
    ("event_small_unionized_1M_opt1", "small_unionized", "event", 1000000, None, "opt1"),
    ("event_small_hash_100k", "small_hash", "event", 100000, None, "event"),
    ("event_small_nuclide_100k", "small_nuclide", "event", 100000, None, "event"),
    ("history_small_unionized_500k", "small_unionized", "history", None, None, "history"),
    ("history_small_hash_20k", "small_hash", "history", 34, 20000, "history"),
    ("event_large_unionized_17M", "large_unionized", "event", None, None, "event"),
    ("event_large_unionized_1M", "large_unionized", "event", 1000000, None, "event"),
    ("event_large_hash_17M", "large_hash", "event", None, None, "event"),
    ("history_large_unionized_500k", "large_unionized", "history", None, None, "history"),
    ("history_large_hash_20k", "large_hash", "history", 34, 20000, "history"),
    ("history_xl45k_hash_1M", "xl45k_hash", "history", 34, 1000000, "history"),
    ("history_xl45k_hash_20k", "xl45k_hash", "history", 34, 20000, "history"),
]

QUICK_SKIP = {"event_large_unionized_17M", "event_large_hash_17M", "history_large_unionized_500k",
              "history_xl45k_hash_1M", "event_small_nuclide_17M"}


def lookup_ids(total, n_first=1024, n_spread=1024, seed=7):
    rng = np.random.default_rng(seed)
    first = np.arange(min(n_first, total), dtype=np.int64)
    spread = np.sort(rng.choice(np.arange(n_first, total, dtype=np.int64), size=n_spread, replace=False))
    return np.concatenate([first, spread])


def ref_lookup(ref, inp, sd, E, mat):
    macro = (C.c_double * 5)()
    ref.calculate_macro_xs(E, mat, inp.n_isotopes, inp.n_gridpoints, sd.num_nucs, sd.concs,
                           sd.unionized_energy_array, sd.index_grid, sd.nuclide_grid, sd.mats, macro,
                           inp.grid_type, inp.hash_bins, sd.max_num_nucs)
    m = np.array(macro[:], dtype=np.float64)
    # argmax exactly as openmp-threading/Simulation.c:100-110
    mx, am = -1.0, 0
    for j in range(5):
        if m[j] > mx:
            mx, am = m[j], j
    return m, am


def dump_lookups(ref, inp, sd, name, total):
    ids = lookup_ids(total)
    n = len(ids)
    E = np.empty(n)
    mat = np.empty(n, np.int32)
    idx = np.full(n, -1, np.int64)
    macro = np.empty((n, 5))
    am = np.empty(n, np.int32)
    for k, i in enumerate(ids):
        seed = C.c_uint64(ref.fast_forward_LCG(O.STARTING_SEED, 2 * int(i)))
        E[k] = ref.LCG_random_double(C.byref(seed))
        mat[k] = ref.pick_mat(C.byref(seed))
        if inp.grid_type == O.UNIONIZED:
            idx[k] = ref.grid_search(inp.n_isotopes * inp.n_gridpoints, E[k], sd.unionized_energy_array)
        macro[k], am[k] = ref_lookup(ref, inp, sd, E[k], int(mat[k]))
    np.savez_compressed(os.path.join(OUT, f"lookups_{name}.npz"), ids=ids, energy=E, mat=mat,
                        ueg_idx=idx, macro=macro, argmax=am)


def dump_micro(ref, inp, sd, name):
    """calculate_micro_xs for every nuclide of the first 2 sampled lookups of each material."""
    picks = {m: [] for m in range(12)}
    i = 0
    while any(len(v) < 2 for v in picks.values()):
        seed = C.c_uint64(ref.fast_forward_LCG(O.STARTING_SEED, 2 * i))
        E = ref.LCG_random_double(C.byref(seed))
        m = ref.pick_mat(C.byref(seed))
        if len(picks[m]) < 2:
            picks[m].append((i, E))
        i += 1
    rec_id, rec_E, rec_mat, rec_off, blocks = [], [], [], [0], []
    for m in range(12):
        for (lid, E) in picks[m]:
            if inp.grid_type == O.UNIONIZED:
                idx = ref.grid_search(inp.n_isotopes * inp.n_gridpoints, E, sd.unionized_energy_array)
            elif inp.grid_type == O.HASH:
                du = 1.0 / inp.hash_bins
                idx = int(E / du)
            else:
                idx = -1
            nn = sd.num_nucs[m]
            blk = np.empty((nn, 5))
            xs = (C.c_double * 5)()
            for j in range(nn):
                nuc = sd.mats[m * sd.max_num_nucs + j]
                ref.calculate_micro_xs(E, nuc, inp.n_isotopes, inp.n_gridpoints, sd.unionized_energy_array,
                                       sd.index_grid, sd.nuclide_grid, idx, xs, inp.grid_type, inp.hash_bins)
                blk[j] = xs[:]
            rec_id.append(lid)
            rec_E.append(E)
            rec_mat.append(m)
            blocks.append(blk)
            rec_off.append(rec_off[-1] + nn)
    np.savez_compressed(os.path.join(OUT, f"micro_{name}.npz"), ids=np.array(rec_id, np.int64),
                        energy=np.array(rec_E), mat=np.array(rec_mat, np.int32),
                        offsets=np.array(rec_off, np.int64), micro=np.concatenate(blocks))


def dump_history(ref, inp, sd, name, n_particles):
    """Per-particle sums: openmp-threading/Simulation.c:162-234, step by step."""
    sums = np.empty(n_particles, np.uint64)
    seeds = np.empty(n_particles, np.uint64)
    for p in range(n_particles):
        off = (p * inp.lookups * 10) & 0xFFFFFFFF
        off = off - (1 << 32) if off >= (1 << 31) else off
        seed = C.c_uint64(ref.fast_forward_LCG(O.STARTING_SEED, off & 0xFFFFFFFFFFFFFFFF))
        E = ref.LCG_random_double(C.byref(seed))
        mat = ref.pick_mat(C.byref(seed))
        v = 0
        for _ in range(inp.lookups):
            m, am = ref_lookup(ref, inp, sd, E, mat)
            v += am + 1
            nf = int(np.sum(m > 1.0))
            if nf > 0:
                seed = C.c_uint64(ref.fast_forward_LCG(seed.value, nf))
            E = ref.LCG_random_double(C.byref(seed))
            mat = ref.pick_mat(C.byref(seed))
        sums[p] = v
        seeds[p] = seed.value
    np.savez_compressed(os.path.join(OUT, f"history_{name}.npz"), first=np.int64(0), sums=sums,
                        final_seed=seeds, lookups=np.int64(inp.lookups))


def dump_sampling(ref):
    rng = np.random.default_rng(11)
    seeds = rng.integers(0, 2**63, size=4096, dtype=np.uint64)
    ns = np.concatenate([np.arange(0, 64), rng.integers(0, 2**40, size=2048 - 64)]).astype(np.uint64)
    ff = np.array([ref.fast_forward_LCG(int(s), int(n)) for s, n in zip(seeds[:2048], ns)], np.uint64)
    lcg_vals, lcg_next, mats = [], [], []
    for s in seeds:
        x = C.c_uint64(int(s))
        lcg_vals.append(ref.LCG_random_double(C.byref(x)))
        lcg_next.append(x.value)
        y = C.c_uint64(int(s))
        mats.append(ref.pick_mat(C.byref(y)))
    np.savez_compressed(os.path.join(OUT, "sampling.npz"), seeds=seeds, ff_n=ns, ff_out=ff,
                        lcg_val=np.array(lcg_vals), lcg_next=np.array(lcg_next, np.uint64),
                        pick_mat=np.array(mats, np.int32))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--quick", action="store_true", help="skip the multi-minute reference runs")
    ap.add_argument("--only", default=None, help="comma list of data cfgs")
    args = ap.parse_args()
    O.build(reference=True)
    ref = O.Reference()
    orc = O.Oracle()
    hashes_path = os.path.join(OUT, "hashes.json")
    digests_path = os.path.join(OUT, "digests.json")
    hashes = json.load(open(hashes_path)) if os.path.exists(hashes_path) else {}
    digests = json.load(open(digests_path)) if os.path.exists(digests_path) else {}
    dump_sampling(ref)
    only = set(args.only.split(",")) if args.only else None
    for name, size, grid, gp in DATA_CFGS:
        if only and name not in only:
            continue
        inp = O.make_inputs(size, grid, "event", gridpoints=gp)
        t = time.time()
        sd = ref.grid_init_do_not_profile(inp, 1)
        print(f"[golden] {name}: reference grid init {time.time() - t:.1f}s", flush=True)
        digests[name] = O.sd_digests(orc.fnv1a, inp, sd)
        digests[name].update(n_isotopes=int(inp.n_isotopes), n_gridpoints=int(inp.n_gridpoints),
                             grid_type=int(inp.grid_type), hash_bins=int(inp.hash_bins))
        dump_lookups(ref, inp, sd, name, 17000000)
        if name in ("small_unionized", "small_hash", "large_unionized", "large_hash", "xl45k_hash"):
            dump_micro(ref, inp, sd, name)
        for key, dcfg, method, lk, pt, drv in RUN_CFGS:
            if dcfg != name or (args.quick and key in QUICK_SKIP):
                continue
            rin = O.make_inputs(size, grid, method, lookups=lk, particles=pt, gridpoints=gp)
            t = time.time()
            if drv == "event":
                v = ref.run_event_based_simulation(rin, sd, 1)
            elif drv == "opt1":
                rin.nthreads = os.cpu_count()
                v = ref.run_event_based_simulation_optimization_1(rin, sd, 1)
            else:
                v = ref.run_history_based_simulation(rin, sd, 1)
            dt = time.time() - t
            hashes[key] = {"vsum": int(v), "hash": int(v) % 999983, "data": name, "method": method,
                           "lookups": int(rin.lookups), "particles": int(rin.particles),
                           "driver": drv, "ref_seconds_8_threads": round(dt, 2)}
            print(f"[golden] {key}: hash {int(v) % 999983} in {dt:.1f}s", flush=True)
        hist_n = 64 if "xl45k" in name or "large" in name else 256
        hin = O.make_inputs(size, grid, "history", lookups=34, particles=hist_n, gridpoints=gp)
        dump_history(ref, hin, sd, name, hist_n)
        json.dump(hashes, open(hashes_path, "w"), indent=1, sort_keys=True)
        json.dump(digests, open(digests_path, "w"), indent=1, sort_keys=True)
        del sd  # reference arrays are never freed by the reference; the process ends soon
    print("[golden] done", flush=True)


if __name__ == "__main__":
    main()
