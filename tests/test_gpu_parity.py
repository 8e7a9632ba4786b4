"""GPU parity: the HIP path through the C ABI against the reference's golden
vectors, the published hashes and the oracle, bit for bit in FP64.

Every test is marked gpu; run on an MI355X with
    python -u -m pytest tests -m gpu -x -v --timeout 120 --timeout-method thread
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
import xsbench_amd as X
from conftest import DATA_CFGS, golden_npz

pytestmark = pytest.mark.gpu

KERNELS = [X.K_BASELINE, X.K_SPLIT, X.K_MAT_LAUNCH, X.K_FUEL_LAUNCH, X.K_MAT_SORT, X.K_FUEL_SORT,
           X.K_MAT_E_SORT]

_CACHE = {}


def data(name):
    """(inputs, host SD, device handle) per golden data configuration, built
    once per session; every device copy also holds the FP32 and BF16 tables."""
    if name not in _CACHE:
        size, grid, gp = DATA_CFGS[name]
        inp = X.make_inputs(size, grid, "event", gridpoints=gp)
        sd = X.grid_init(inp)
        print(f"[gpu-test] upload {name}", flush=True)
        dev = X.Device(inp, sd, build_fp32=True, build_bf16=True)
        _CACHE[name] = (inp, sd, dev)
    return _CACHE[name]


def as_oracle(inp, sd):
    """the oracle's view of the product's tables (same struct layout)"""
    oi = O.Inputs.from_buffer_copy(inp)
    oi._hm = inp._hm
    oi.HM = oi._hm
    return oi, O.SimulationData.from_buffer_copy(sd)


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


# ---------------------------------------------------------------------------
# per-lookup parity
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", list(DATA_CFGS))
@pytest.mark.parametrize("k", KERNELS)
def test_first_lookups_match_reference_vectors(name, k):
    inp, sd, dev = data(name)
    g = golden_npz(f"lookups_{name}.npz")
    n = 1024
    assert np.array_equal(g["ids"][:n], np.arange(n))
    res, out = dev.event(n, 0, kernel_id=k, detail=True)
    assert same_bits(out["energy"], g["energy"][:n])
    assert np.array_equal(out["mat"], g["mat"][:n])
    if inp.grid_type == X.UNIONIZED:
        assert np.array_equal(out["index"], g["ueg_idx"][:n])
    assert same_bits(out["macro"], g["macro"][:n])
    assert np.array_equal(out["argmax"], g["argmax"][:n])
    assert res["vsum"] == int(g["argmax"][:n].sum()) + n


@pytest.mark.parametrize("name", list(DATA_CFGS))
def test_spread_lookups_match_reference_vectors(name):
    inp, sd, dev = data(name)
    g = golden_npz(f"lookups_{name}.npz")
    ids = g["ids"][1024:]
    for j in range(0, len(ids), 16):
        for q, i in enumerate(ids[j:j + 16]):
            _, out = dev.event(1, int(i), kernel_id=X.K_BASELINE, detail=True)
            assert same_bits(out["macro"][0], g["macro"][1024 + j + q])
            assert out["argmax"][0] == g["argmax"][1024 + j + q]


@pytest.mark.parametrize("name", ["small_unionized", "large_unionized", "large_hash", "small_nuclide"])
@pytest.mark.parametrize("k", [X.K_BASELINE, X.K_SPLIT, X.K_MAT_E_SORT])
def test_window_matches_oracle(orc, name, k):
    inp, sd, dev = data(name)
    oi, osd = as_oracle(inp, sd)
    first, n = 5_000_003, 1 << 15
    _, out = dev.event(n, first, kernel_id=k, detail=True)
    E, m, idx, macro, am = orc.event_detail(oi, osd, np.arange(first, first + n))
    assert same_bits(out["energy"], E) and np.array_equal(out["mat"], m)
    assert same_bits(out["macro"], macro) and np.array_equal(out["argmax"], am)


def test_sorted_schedule_matches_reference_order_all_17m_lookups():
    """every lookup of the large unionized workload, macro XS bit for bit and
    argmax: the sorted schedule (wave-shared interval records, per-material
    index rows) against the reference-order kernel (points, the index grid's
    brackets)"""
    inp, sd, dev = data("large_unionized")
    n = 17_000_000
    _, ref = dev.event(n, 0, kernel_id=X.K_BASELINE, detail=True)
    _, got = dev.event(n, 0, kernel_id=X.K_MAT_E_SORT, detail=True)
    assert same_bits(got["macro"], ref["macro"])
    assert np.array_equal(got["argmax"], ref["argmax"])


# ---------------------------------------------------------------------------
# full-size hashes
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k", KERNELS)
def test_large_unionized_event_hash_every_schedule(k):
    inp, sd, dev = data("large_unionized")
    r = dev.event(17_000_000, 0, kernel_id=k)
    print(f"[gpu-test] large unionized k={k}: {r['ms_total']:.2f} ms", flush=True)
    assert r["hash"] == X.PUBLISHED[("event", "large")]
    if k in (X.K_MAT_SORT, X.K_MAT_E_SORT):
        assert sum(r["material_counts"]) == 17_000_000


@pytest.mark.parametrize("n", [1, 3, 8191, 8193, 100_003])
def test_sorted_schedule_odd_and_partial_tiles(orc, n):
    """odd lookup counts and partial sort tiles (pair buffers, histogram tails)
    on the material-energy schedule, per lookup against the oracle"""
    inp, sd, dev = data("small_unionized")
    oi, osd = as_oracle(inp, sd)
    first = 424_242
    E, m, idx, macro, am = orc.event_detail(oi, osd, np.arange(first, first + n))
    r, out = dev.event(n, first, kernel_id=X.K_MAT_E_SORT, detail=True)
    assert same_bits(out["macro"], macro) and np.array_equal(out["argmax"], am)
    assert r["vsum"] == int(am.sum()) + n
    assert sum(r["material_counts"]) == n


@pytest.mark.parametrize("k", [X.K_MAT_LAUNCH, X.K_FUEL_LAUNCH, X.K_MAT_SORT, X.K_FUEL_SORT,
                               X.K_MAT_E_SORT])
def test_sorted_schedule_material_counts(orc, k):
    """the per-material counts a sorted schedule reports (from the sort's
    digit starts) against the oracle's sampled materials"""
    inp, sd, dev = data("small_unionized")
    oi, osd = as_oracle(inp, sd)
    first, n = 777, 1 << 18
    _, m, _, _, _ = orc.event_detail(oi, osd, np.arange(first, first + n))
    want = np.bincount(m, minlength=12)
    got = np.array(dev.event(n, first, kernel_id=k)["material_counts"])
    if k in (X.K_FUEL_LAUNCH, X.K_FUEL_SORT):
        assert got[0] == want[0] and got[1:].sum() == 0
    else:
        assert np.array_equal(got, want)


@pytest.mark.parametrize("name,k", [("large_hash", X.K_BASELINE), ("large_hash", X.K_MAT_E_SORT),
                                    ("small_unionized", X.K_BASELINE), ("small_hash", X.K_MAT_E_SORT),
                                    ("small_nuclide", X.K_BASELINE), ("small_nuclide", X.K_MAT_E_SORT)])
def test_event_hash_grid_invariance(name, k):
    inp, sd, dev = data(name)
    r = dev.event(17_000_000, 0, kernel_id=k)
    size = DATA_CFGS[name][0]
    assert r["hash"] == X.PUBLISHED[("event", size)]


def test_small_one_million_lookups_golden(golden_hashes):
    inp, sd, dev = data("small_unionized")
    for k in (X.K_BASELINE, X.K_MAT_E_SORT):
        assert dev.event(1_000_000, 0, kernel_id=k)["vsum"] == golden_hashes["event_small_unionized_1M"]["vsum"]


# ---------------------------------------------------------------------------
# history mode
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "large_hash", "xl45k_hash"])
@pytest.mark.parametrize("kernel", [0, 6])
def test_history_particles_match_reference(name, kernel):
    inp, sd, dev = data(name)
    g = golden_npz(f"history_{name}.npz")
    n = len(g["sums"])
    r, out = dev.history(n, int(g["lookups"]), kernel_id=kernel, detail=True)
    assert np.array_equal(out["sums"], g["sums"])
    assert np.array_equal(out["final_seed"], g["final_seed"])


@pytest.mark.parametrize("kernel", [0, 6])
@pytest.mark.parametrize("name,size", [("small_unionized", "small"), ("large_unionized", "large")])
def test_history_published_hashes(name, size, kernel):
    inp, sd, dev = data(name)
    r = dev.history(500_000, 34, kernel_id=kernel)
    print(f"[gpu-test] history {name} k={kernel}: {r['ms_total']:.2f} ms", flush=True)
    assert r["hash"] == X.PUBLISHED[("history", size)]


@pytest.mark.parametrize("kernel", [0, 6])
def test_history_xl45k_golden(golden_hashes, kernel):
    inp, sd, dev = data("xl45k_hash")
    for key in ("history_xl45k_hash_20k", "history_xl45k_hash_1M"):
        h = golden_hashes[key]
        r = dev.history(h["particles"], h["lookups"], kernel_id=kernel)
        print(f"[gpu-test] {key} k={kernel}: {r['ms_total']:.2f} ms", flush=True)
        assert r["vsum"] == h["vsum"]


@pytest.mark.parametrize("kernel", [0, 6])
def test_history_xl45k_fp32_golden(golden_hashes, kernel):
    """the XL-45000 hash-grid history workload (1M particles x 34 lookups) with
    FP32 tables: the reference's own sum (its build in the survey container)"""
    inp, sd, dev = data("xl45k_hash")
    h = golden_hashes["history_xl45k_hash_1M"]
    r = dev.history(h["particles"], h["lookups"], kernel_id=kernel, precision=X.XS_FP32)
    print(f"[gpu-test] xl45k FP32 history 1M k={kernel}: {r['ms_total']:.2f} ms, "
          f"{r['n_fallback']} FP64 re-evaluations", flush=True)
    assert r["vsum"] == h["vsum"]


def test_history_rounds_graph_replays():
    """the captured chain of rounds replays with the same result, with other
    work between replays, and recaptures for a new particle count"""
    inp, sd, dev = data("small_hash")
    want = dev.history(70000, 34, kernel_id=0)["vsum"]
    for rep in range(3):
        assert dev.history(70000, 34, kernel_id=6)["vsum"] == want, rep
        dev.event(20000, kernel_id=X.K_MAT_E_SORT)
    assert dev.history(1000, 5, kernel_id=6)["vsum"] == dev.history(1000, 5, kernel_id=0)["vsum"]
    assert dev.history(1, 34, kernel_id=6)["vsum"] == dev.history(1, 34, kernel_id=0)["vsum"]


# ---------------------------------------------------------------------------
# grid construction on the device (GridInit.c:3-151) against the host build,
# itself byte-identical to the reference (digest fixtures)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "small_nuclide", "large_hash",
                                  "xl45k_hash", "large_unionized"])
def test_device_grid_construction_matches_host(name):
    inp, sd, _ = data(name)
    want = X.sd_arrays(inp, sd)
    print(f"[gpu-test] device grid build {name}", flush=True)
    with X.Device(inp, None) as dev:
        got = dev.download()
        assert np.array_equal(got["nuclide_grid"].reshape(-1, 6).view(np.uint64),
                              want["nuclide_grid"].view(np.uint64))
        if inp.grid_type == X.UNIONIZED:
            assert np.array_equal(got["unionized_energy_array"].view(np.uint64),
                                  want["unionized_energy_array"].view(np.uint64))
        if inp.grid_type != X.NUCLIDE:
            assert np.array_equal(got["index_grid"].reshape(-1), want["index_grid"])
        del got
        if name == "large_unionized":
            assert dev.event(17_000_000, kernel_id=X.K_MAT_E_SORT)["hash"] == 952131


def test_device_grid_construction_tiny_grids():
    """tiny grids (2 and 3 points per nuclide) on a hash and a unionized grid:
    the device tables equal the host's, including the sweep's rule at the
    last interval"""
    for grid, gp in (("unionized", 2), ("unionized", 3), ("hash", 3), ("unionized", 40)):
        inp = X.make_inputs("small", grid, "event", gridpoints=gp, hash_bins=17)
        sd = X.grid_init(inp)
        want = X.sd_arrays(inp, sd)
        with X.Device(inp, None) as dev:
            got = dev.download()
        assert np.array_equal(got["nuclide_grid"].reshape(-1, 6), want["nuclide_grid"])
        if grid == "unionized":
            assert np.array_equal(got["unionized_energy_array"], want["unionized_energy_array"])
        assert np.array_equal(got["index_grid"].reshape(-1), want["index_grid"]), (grid, gp)


@pytest.mark.parametrize("grid,gp", [("unionized", 2), ("unionized", 3), ("hash", 3),
                                     ("unionized", 40)])
def test_tiny_grid_lookups_match_oracle(orc, grid, gp):
    """lookups on 2-, 3- and 40-point grids (every bracket at the clamped
    last interval, records next to the zero record) on the sorted and the
    fused schedules, against the oracle"""
    inp = X.make_inputs("small", grid, "event", gridpoints=gp, hash_bins=17)
    sd = X.grid_init(inp)
    oi, osd = as_oracle(inp, sd)
    first, n = 1000, 1 << 14
    E, m, idx, macro, am = orc.event_detail(oi, osd, np.arange(first, first + n))
    with X.Device(inp, sd) as dev:
        for k in (X.K_BASELINE, X.K_MAT_E_SORT):
            _, out = dev.event(n, first, kernel_id=k, detail=True)
            assert same_bits(out["macro"], macro) and np.array_equal(out["argmax"], am), (grid, gp, k)


def test_zero_width_intervals_match_oracle(orc):
    """Intervals of zero width: one nuclide's every second energy is lowered onto
    the one before it after the grids were built, so the index grid still
    brackets lookups into them and the reference divides by zero (infinite
    sums). The FP64 schedules give the oracle's values, and with FP32 tables,
    whose interval records then hold a NaN reciprocal, those lookups are
    re-evaluated in FP64 and return the same sums."""
    inp = X.make_inputs("small", "unionized", "event")
    sd = X.grid_init(inp)
    g = np.ctypeslib.as_array(C.cast(sd.nuclide_grid, C.POINTER(C.c_double)),
                              (inp.n_isotopes, inp.n_gridpoints, 6))
    mats = np.ctypeslib.as_array(sd.mats, (12 * sd.max_num_nucs,)).reshape(12, -1)
    nuc = int(mats[2, 0])  # a nuclide of the most frequent material
    hi = g[nuc, 2::2, 0]
    g[nuc, 2::2, 0] = g[nuc, 1:-1:2, 0][: hi.shape[0]]
    oi, osd = as_oracle(inp, sd)
    first, n = 1000, 1 << 14
    E, m, idx, macro, am = orc.event_detail(oi, osd, np.arange(first, first + n))
    bad = ~np.isfinite(macro).all(axis=1)
    assert bad.sum() > 1000
    print(f"[gpu-test] zero-width intervals: {int(bad.sum())} of {n} lookups with infinite sums",
          flush=True)
    with X.Device(inp, sd, build_fp32=True) as dev:
        for k in (X.K_BASELINE, X.K_SPLIT, X.K_MAT_E_SORT):
            _, out = dev.event(n, first, kernel_id=k, detail=True)
            assert np.array_equal(out["macro"], macro, equal_nan=True), k
            assert same_bits(out["macro"][~bad], macro[~bad]) and np.array_equal(out["argmax"], am), k
            if k == X.K_BASELINE:  # the fused kernel is FP64 only
                continue
            _, mix = dev.event(n, first, kernel_id=k, precision=X.XS_FP32, detail=True)
            assert np.array_equal(mix["argmax"], am), k
            assert np.all(mix["fallback"][bad] == 1), k
            assert np.array_equal(mix["macro"][bad], macro[bad], equal_nan=True), k


def test_unsorted_rows_keep_the_bisection(orc):
    """A nuclide row that is not sorted (neighbouring energies swapped after
    the grids were built): upload finds the descents, and the hash grid's
    wave path then bisects as the reference does instead of counting the
    bin range's energies; every lookup matches the oracle bit for bit."""
    # ~14 points per bin (within the counted range), and enough lookups that
    # sorted waves share a bin (the wave path)
    inp = X.make_inputs("small", "hash", "event", hash_bins=800)
    sd = X.grid_init(inp)
    g = np.ctypeslib.as_array(C.cast(sd.nuclide_grid, C.POINTER(C.c_double)),
                              (inp.n_isotopes, inp.n_gridpoints, 6))
    mats = np.ctypeslib.as_array(sd.mats, (12 * sd.max_num_nucs,)).reshape(12, -1)
    for nuc in {int(mats[2, 0]), int(mats[0, 0])}:
        e = g[nuc, :, 0].copy()
        g[nuc, 1:-2:3, 0] = e[2:-1:3]
        g[nuc, 2:-1:3, 0] = e[1:-2:3]
    oi, osd = as_oracle(inp, sd)
    first, n = 777, 1 << 19
    E, m, idx, macro, am = orc.event_detail(oi, osd, np.arange(first, first + n))
    # history: the rounds (and the mixed per-lane loop) count short bin
    # ranges per lane on sorted rows; on these rows they must bisect too
    P = 70_000
    oh = O.Inputs.from_buffer_copy(oi)
    oh.lookups = 34  # lookups per particle in history mode
    sums, seeds = orc.history_detail(oh, osd, 0, P)
    with X.Device(inp, sd, build_fp32=True) as dev:
        for k in (X.K_BASELINE, X.K_MAT_E_SORT):
            _, out = dev.event(n, first, kernel_id=k, detail=True)
            assert same_bits(out["macro"], macro) and np.array_equal(out["argmax"], am), k
        for k in (0, 6):
            for prec in (X.XS_FP64, X.XS_FP32):
                _, h = dev.history(P, 34, kernel_id=k, precision=prec, detail=True)
                assert np.array_equal(h["sums"], sums), (k, prec)
                assert np.array_equal(h["final_seed"], seeds), (k, prec)


# ---------------------------------------------------------------------------
# per-material unionized index tables: uint16 brackets when n_gridpoints <=
# 65536 (the default), int32 otherwise or on request (BUILD_IDX32)
# ---------------------------------------------------------------------------
def _midx_bytes(dev, n_union, width):
    """bytes of the per-material index tables: rows padded to 16 bytes"""
    nn = np.ctypeslib.as_array(C.cast(dev.tables.num_nucs, C.POINTER(C.c_int)), (12,))
    per = 16 // width
    return sum(n_union * ((int(v) + per - 1) // per * per) * width for v in nn)


@pytest.mark.parametrize("name", ["small_unionized", "large_unionized"])
def test_index_tables_32bit_match_oracle(orc, name):
    inp, sd, dev16 = data(name)
    oi, osd = as_oracle(inp, sd)
    n_union = inp.n_isotopes * inp.n_gridpoints
    assert dev16.info()["bytes_midx"] == _midx_bytes(dev16, n_union, 2)
    want = X.PUBLISHED[("event", DATA_CFGS[name][0])]
    print(f"[gpu-test] 32-bit index tables {name}", flush=True)
    with X.Device(inp, sd, build_fp32=True, idx32=True) as dev:
        assert dev.info()["bytes_midx"] == _midx_bytes(dev, n_union, 4)
        first, n = 5_000_003, 1 << 15
        E, m, idx, macro, am = orc.event_detail(oi, osd, np.arange(first, first + n))
        for k in (X.K_BASELINE, X.K_MAT_E_SORT):
            _, out = dev.event(n, first, kernel_id=k, detail=True)
            assert same_bits(out["macro"], macro) and np.array_equal(out["argmax"], am)
        for k in (X.K_BASELINE, X.K_MAT_E_SORT):
            assert dev.event(17_000_000, kernel_id=k)["hash"] == want
        assert dev.event(17_000_000, kernel_id=X.K_MAT_E_SORT, precision=X.XS_FP32)["hash"] == want
        _, h16 = dev16.history(20000, 34, kernel_id=6, detail=True)
        _, h32 = dev.history(20000, 34, kernel_id=6, detail=True)
        assert np.array_equal(h16["sums"], h32["sums"])
        sizes, seeds, thr = orc.batch_plan(2024, 24, 1000, 60000)
        _, v16, _ = dev16.batched(sizes, seeds, thr, mode=0)
        _, v32, _ = dev.batched(sizes, seeds, thr, mode=0)
        assert np.array_equal(v16, v32)


def test_index_tables_width_boundary():
    """n_gridpoints 65536 (largest uint16 case, brackets up to 65534) and
    65537 (int32 chosen by itself): equal to forced int32 tables"""
    for gp, width in ((65536, 2), (65537, 4)):
        inp = X.make_inputs("small", "unionized", "event", gridpoints=gp)
        n_union = inp.n_isotopes * gp
        print(f"[gpu-test] index width boundary gridpoints={gp}", flush=True)
        with X.Device(inp, None) as a, X.Device(inp, None, idx32=True) as b:
            assert a.info()["bytes_midx"] == _midx_bytes(a, n_union, width)
            assert b.info()["bytes_midx"] == _midx_bytes(b, n_union, 4)
            assert a.info()["bytes_ivl"] == n_union * 96  # one 12-double record per point
            for k in (X.K_BASELINE, X.K_MAT_E_SORT):
                _, oa = a.event(1 << 16, 12345, kernel_id=k, detail=True)
                _, ob = b.event(1 << 16, 12345, kernel_id=k, detail=True)
                assert np.array_equal(oa["index"], ob["index"]) and oa["index"].max() > 0.99 * n_union
                assert same_bits(oa["macro"], ob["macro"])
            assert (a.event(1_000_000, kernel_id=X.K_MAT_E_SORT)["hash"]
                    == b.event(1_000_000, kernel_id=X.K_MAT_E_SORT)["hash"])


# ---------------------------------------------------------------------------
# mixed precision: identical decisions, bounded error
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [X.XS_FP32, X.XS_BF16])
@pytest.mark.parametrize("name", ["large_unionized", "large_hash"])
def test_mixed_precision_event_hash_identical(name, prec):
    inp, sd, dev = data(name)
    for k in (X.K_SPLIT, X.K_MAT_E_SORT):
        r = dev.event(17_000_000, 0, kernel_id=k, precision=prec)
        print(f"[gpu-test] {name} prec={prec} k={k}: {r['ms_total']:.2f} ms, "
              f"{r['n_fallback']} FP64 re-evaluations", flush=True)
        assert r["hash"] == X.PUBLISHED[("event", "large")]


@pytest.mark.parametrize("kernel", [X.K_SPLIT, X.K_MAT_E_SORT])
@pytest.mark.parametrize("prec,p9999", [(X.XS_FP32, 1e-6), (X.XS_BF16, 2e-2)])
@pytest.mark.parametrize("name", ["large_hash", "large_unionized"])
def test_mixed_precision_error_per_channel(name, prec, p9999, kernel):
    """Every lookup's error lies inside its own rigorous bound B (the bound
    that guards the argmax); the per-channel relative error is reported.
    The sorted schedule runs the wave-uniform paths (with FP32 tables on the
    unionized grid: the FP32 interval records)."""
    inp, sd, dev = data(name)
    n = 1 << 18
    _, ref = dev.event(n, 0, kernel_id=kernel, precision=X.XS_FP64, detail=True)
    _, mix = dev.event(n, 0, kernel_id=kernel, precision=prec, detail=True)
    assert np.array_equal(ref["argmax"], mix["argmax"])
    err = np.abs(mix["macro"] - ref["macro"])
    assert np.all(err <= mix["bound"][:, None])
    rel = err / np.abs(ref["macro"])
    print(f"[gpu-test] {name} prec={prec} k={kernel}: relative error per channel max {rel.max(axis=0)}, "
          f"mean {rel.mean(axis=0)}, 99.99% {np.quantile(rel, 0.9999, axis=0)}; "
          f"{int(mix['fallback'].sum())} of {n} re-evaluated in FP64", flush=True)
    assert np.all(np.quantile(rel, 0.9999, axis=0) < p9999)
    assert np.all(rel.mean(axis=0) < p9999 / 10)
    if prec == X.XS_FP32:  # certified per lookup and channel
        assert rel.max() <= 1e-6
        assert mix["fallback"].mean() < 0.01
    assert np.all(err[mix["fallback"] == 1] == 0)  # re-evaluated lookups return FP64 sums


@pytest.mark.parametrize("name", ["large_unionized", "large_hash"])
def test_fp32_relative_error_all_17m_lookups(name):
    """FP32 tables: every one of the workload's 17M lookups, in windows of 2M,
    within 1e-6 relative of the FP64 path (bit-identical to the reference)
    per channel, inside its own bound, with the reference's argmax"""
    inp, sd, dev = data(name)
    total, win = 17_000_000, 1 << 21
    worst, fb = np.zeros(5), 0
    for first in range(0, total, win):
        n = min(win, total - first)
        _, ref = dev.event(n, first, kernel_id=X.K_MAT_E_SORT, precision=X.XS_FP64, detail=True)
        _, mix = dev.event(n, first, kernel_id=X.K_MAT_E_SORT, precision=X.XS_FP32, detail=True)
        assert np.array_equal(ref["argmax"], mix["argmax"]), first
        err = np.abs(mix["macro"] - ref["macro"])
        assert np.all(err <= mix["bound"][:, None]), first
        rel = err / np.abs(ref["macro"])
        worst = np.maximum(worst, rel.max(axis=0))
        fb += int(mix["fallback"].sum())
    print(f"[gpu-test] {name} FP32, all 17M lookups: relative error per channel max {worst}, "
          f"{fb} re-evaluated in FP64", flush=True)
    assert worst.max() <= 1e-6


@pytest.mark.parametrize("prec", [X.XS_FP32, X.XS_BF16])
@pytest.mark.parametrize("kernel", [0, 6])
def test_mixed_precision_history_identical(prec, kernel):
    inp, sd, dev = data("large_hash")
    r64 = dev.history(20000, 34)
    r = dev.history(20000, 34, precision=prec, kernel_id=kernel)
    assert r["vsum"] == r64["vsum"]


# ---------------------------------------------------------------------------
# explicit-sample API on device memory, batches, counters, errors
# ---------------------------------------------------------------------------
def _hip_runtime():
    """the HIP runtime libxsb_hip.so itself links (already loaded)"""
    X.hip_lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


@pytest.mark.parametrize("name", ["small_hash", "small_unionized"])
def test_lookup_on_device_buffers(orc, name):
    """explicit samples in device memory; samples with a material outside
    [0, 12) are skipped (argmax -1), energies outside [0, 1) stay in bounds"""
    inp, sd, dev = data(name)
    oi, osd = as_oracle(inp, sd)
    hip = _hip_runtime()
    rng = np.random.default_rng(17)
    n = 50000
    E = rng.random(n)
    M = rng.integers(0, 12, n).astype(np.int32)
    bad = np.arange(8) * 997
    M[bad[:4]] = [-1, 12, 1 << 30, -(1 << 30)]
    E[bad[4:]] = [-0.5, np.nan, 1.5, 1.0]
    ptrs = [C.c_void_p() for _ in range(5)]
    sizes = [n * 8, n * 4, n * 40, n * 4, 8]
    for p_, b in zip(ptrs, sizes):
        assert hip.hipMalloc(C.byref(p_), b) == 0
    try:
        H2D, D2H = 1, 2
        assert hip.hipMemcpy(ptrs[0], E.ctypes.data, n * 8, H2D) == 0
        assert hip.hipMemcpy(ptrs[1], M.ctypes.data, n * 4, H2D) == 0
        assert hip.hipMemset(ptrs[4], 0, 8) == 0
        dev.lookup(n, ptrs[0], ptrs[1], X.XS_FP64, ptrs[2], ptrs[3], ptrs[4], None)
        assert hip.hipDeviceSynchronize() == 0
        macro = np.empty((n, 5))
        am = np.empty(n, np.int32)
        vs = np.zeros(1, np.uint64)
        assert hip.hipMemcpy(macro.ctypes.data, ptrs[2], n * 40, D2H) == 0
        assert hip.hipMemcpy(am.ctypes.data, ptrs[3], n * 4, D2H) == 0
        assert hip.hipMemcpy(vs.ctypes.data, ptrs[4], 8, D2H) == 0
    finally:
        for p_ in ptrs:
            hip.hipFree(p_)
    assert np.all(am[bad[:4]] == -1)
    ok = np.ones(n, bool)
    ok[bad] = False
    want_macro, want_am = orc.lookup_detail(oi, osd, E[ok], M[ok])
    assert same_bits(macro[ok], want_macro)
    assert np.array_equal(am[ok], want_am)
    assert np.all((am[bad[4:]] >= 0) & (am[bad[4:]] < 5))
    assert int(vs[0]) == int(am[am >= 0].sum()) + int((am >= 0).sum())


def _batch_truth(orc, oi, osd, sizes, seeds, thr):
    return [orc.run_batch(oi, osd, int(seeds[b]), thr[b], int(sizes[b])) for b in range(len(sizes))]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("prec", [X.XS_FP64, X.XS_FP32])
def test_batched_per_batch_hashes(orc, mode, prec):
    inp, sd, dev = data("small_unionized")
    oi, osd = as_oracle(inp, sd)
    sizes, seeds, thr = orc.batch_plan(2024, 24, 1000, 60000)
    plan = {}
    r, vs, done = dev.batched(sizes, seeds, thr, precision=prec, mode=mode, plan=plan)
    assert [int(v) for v in vs] == _batch_truth(orc, oi, osd, sizes, seeds, thr)
    assert r["vsum"] == int(vs.sum())
    assert np.all(done > 0)
    if mode == 0:
        assert plan["sort"] in (1, 2) and plan["stage"] in (8, 16) and plan["tile"] >= 256
        assert plan["graph"] == 1


_BATCH_TRUTH = {}


@pytest.mark.parametrize("prec", [X.XS_FP64, X.XS_FP32])
def test_batched_workload_large_against_oracle(orc, prec):
    """the batched workload itself: 512 batches, sizes log-uniform in [1e3,
    2e6] (122M lookups), on the large unionized grid with the autotuned plan;
    every batch's sum against the oracle's (openmp-threading/Simulation.c:
    326-375 per lookup, the batch's own stream and material CDF)"""
    inp, sd, dev = data("large_unionized")
    oi, osd = as_oracle(inp, sd)
    sizes, seeds, thr = X.make_batches(2024, 512, 1000, 2_000_000)
    if "want" not in _BATCH_TRUTH:
        print(f"[gpu-test] oracle over {int(sizes.sum()):,} batched lookups", flush=True)
        _BATCH_TRUTH["want"] = _batch_truth(orc, oi, osd, sizes, seeds, thr)
    want = _BATCH_TRUTH["want"]
    plan = {}
    r, vs, done = dev.batched(sizes, seeds, thr, precision=prec, mode=0, plan=plan)
    print(f"[gpu-test] batched large prec={prec}: {r['ms_total']:.2f} ms, plan {plan}, "
          f"completion p50 {np.percentile(done, 50):.2f} ms p99 {np.percentile(done, 99):.2f} ms",
          flush=True)
    assert [int(v) for v in vs] == want
    assert r["vsum"] == sum(want)
    # the default windows: the smallest batches first, so most complete early
    assert plan["windows"] == 2
    assert np.percentile(done, 50) < 0.5 * np.percentile(done, 99)
    # the replayed graph of the tuned plan, with the 16-bit index tables it uses
    r2, vs2, _ = dev.batched(sizes, seeds, thr, precision=prec, mode=0)
    assert [int(v) for v in vs2] == want
    # one window: every batch completes at the end, the same sums
    p1 = {}
    _, vs3, done3 = dev.batched(sizes, seeds, thr, precision=prec, mode=0, window=-1, plan=p1)
    assert p1["windows"] == 1 and [int(v) for v in vs3] == want


@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "small_nuclide"])
@pytest.mark.parametrize("tile,stage,sort,window,mapping", [(256, 8, 1, 0, 2), (2048, 16, 2, 0, 1),
                                                            (8192, 8, 2, 25000, 2),
                                                            (512, 16, 1, 1, 1)])
def test_batched_sorted_plans_identical(orc, name, tile, stage, sort, window, mapping):
    """every plan and sort window of the sorted pipeline gives the oracle's per-batch sums"""
    inp, sd, dev = data(name)
    oi, osd = as_oracle(inp, sd)
    sizes, seeds, thr = orc.batch_plan(77, 12, 100, 30000)
    sizes[3] = 0  # an empty batch
    plan = {}
    r, vs, done = dev.batched(sizes, seeds, thr, mode=0, tile=tile, stage=stage, sort=sort,
                              window=window, mapping=mapping, plan=plan)
    assert [int(v) for v in vs] == _batch_truth(orc, oi, osd, sizes, seeds, thr)
    assert plan["tile"] == tile and plan["sort"] == sort and plan["tuned"] == 0
    assert plan["mapping"] == mapping
    assert done[3] == 0 and np.all(np.delete(done, 3) > 0)
    if window:
        assert plan["windows"] > 1


def test_batched_graph_replays_match_eager(orc):
    """the captured pipeline replayed several times, with other work and
    synchronisations between replays, gives the eager result every time"""
    inp, sd, dev = data("small_unionized")
    oi, osd = as_oracle(inp, sd)
    sizes, seeds, thr = orc.batch_plan(5, 16, 1000, 40000)
    _, v_eager, _ = dev.batched(sizes, seeds, thr, mode=0, tile=1024, stage=16, sort=2, eager=True)
    assert [int(v) for v in v_eager] == _batch_truth(orc, oi, osd, sizes, seeds, thr)
    for rep in range(4):
        plan = {}
        _, vs, _ = dev.batched(sizes, seeds, thr, mode=0, tile=1024, stage=16, sort=2, plan=plan)
        assert plan["graph"] == 1
        assert np.array_equal(vs, v_eager), rep
        assert dev.event(30000, kernel_id=X.K_MAT_E_SORT)["hash"] > 0  # eager work in between
    # new batch contents, same shape: the graph replays on the new tables
    seeds2 = seeds + np.uint64(1)
    _, vs2, _ = dev.batched(sizes, seeds2, thr, mode=0, tile=1024, stage=16, sort=2)
    assert [int(v) for v in vs2] == _batch_truth(orc, oi, osd, sizes, seeds2, thr)


def test_batched_autotuner_caches_its_plan(orc):
    """the first call with open plan fields tunes; a second call of the same
    shape reuses the cached plan and the captured graph; fixing one field
    tunes the rest again"""
    inp, sd, dev = data("small_hash")
    oi, osd = as_oracle(inp, sd)
    sizes, seeds, thr = orc.batch_plan(31, 20, 500, 20000)
    want = _batch_truth(orc, oi, osd, sizes, seeds, thr)
    p1, p2, p3 = {}, {}, {}
    _, v1, _ = dev.batched(sizes, seeds, thr, plan=p1)
    _, v2, _ = dev.batched(sizes, seeds, thr, plan=p2)
    assert p1["tuned"] == 1 and p2["tuned"] == 0 and p2["graph"] == 1
    assert {k: p1[k] for k in ("tile", "stage", "sort", "mapping")} == \
           {k: p2[k] for k in ("tile", "stage", "sort", "mapping")}
    other = 1 if p1["mapping"] != 1 else 2
    _, v3, _ = dev.batched(sizes, seeds, thr, mapping=other, plan=p3)
    assert p3["mapping"] == other and p3["tuned"] == 1
    for v in (v1, v2, v3):
        assert [int(x) for x in v] == want


def test_batched_many_batches_global_sums(orc):
    """more batches than the kernel's LDS table: per-batch sums go to device atomics"""
    inp, sd, dev = data("small_hash")
    oi, osd = as_oracle(inp, sd)
    sizes, seeds, thr = orc.batch_plan(99, 1500, 1, 60)
    r, vs, done = dev.batched(sizes, seeds, thr, mode=0, tile=2048, stage=16, sort=2)
    assert [int(v) for v in vs] == _batch_truth(orc, oi, osd, sizes, seeds, thr)
    assert np.all(done[sizes > 0] > 0)


@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "small_nuclide"])
def test_balg_counter_matches_oracle(orc, name):
    inp, sd, dev = data(name)
    oi, osd = as_oracle(inp, sd)
    n = 20000
    assert dev.balg(n) == pytest.approx(orc.balg_event(oi, osd, 0, n), rel=1e-12)


def test_error_paths_on_device():
    inp = X.make_inputs("small", "hash", "event", gridpoints=50)
    sd = X.grid_init(inp)
    with X.Device(inp, sd) as dev:  # no compact tables
        with pytest.raises(X.XsbError) as e:
            dev.event(100, precision=X.XS_FP32)
        assert e.value.code == 5
        with pytest.raises(X.XsbError) as e:
            dev.event(100, kernel_id=11)
        assert e.value.code == 1
        with pytest.raises(X.XsbError) as e:
            dev.event(-5)
        assert e.value.code == 1
        assert dev.event(0)["vsum"] == 0
        assert dev.event(0, kernel_id=X.K_MAT_E_SORT)["vsum"] == 0
        with pytest.raises(X.XsbError) as e:
            dev.history(10, 34, kernel_id=3)
        assert e.value.code == 1
        assert dev.history(0, 34, kernel_id=6)["vsum"] == 0
        assert dev.history(10, 0, kernel_id=6)["vsum"] == 0
        with pytest.raises(X.XsbError) as e:
            dev.batched([10], [1], np.tile(np.linspace(0.1, 2.0, 12), (1, 1)), mode=3)
        assert e.value.code == 1


# ---------------------------------------------------------------------------
# the reference's host surface end to end
# ---------------------------------------------------------------------------
def test_host_drivers_return_published_hash():
    lib = X.host_lib()
    inp = X.make_inputs("small", "unionized", "event")
    sd = X.grid_init(inp)
    # unstaged SD: the driver uploads a temporary copy
    assert lib.run_event_based_simulation(inp, sd, 1) % 999983 == X.PUBLISHED[("event", "small")]
    lib.move_simulation_data_to_device(inp, 1, sd)
    inp.kernel_id = 6
    assert lib.run_event_based_simulation(inp, sd, 1) % 999983 == X.PUBLISHED[("event", "small")]
    assert lib.run_event_based_simulation_optimization_1(inp, sd, 1) % 999983 == X.PUBLISHED[("event", "small")]
    hin = X.make_inputs("small", "unionized", "history")
    assert lib.run_history_based_simulation(hin, sd, 1) % 999983 == X.PUBLISHED[("history", "small")]
    lib.release_device_memory(sd)


@pytest.mark.parametrize("args", [["-m", "event", "-s", "small"],
                                  ["-m", "event", "-s", "small", "-G", "hash", "-k", "6", "-P", "fp32"],
                                  ["-m", "history", "-s", "small", "-G", "nuclide"]])
def test_program_flow_valid_checksum(capfd, args):
    """openmp-threading/Main.c flow in-process (no exec from a GPU process)"""
    rc = X.run_main(args)
    out = capfd.readouterr().out
    assert rc == 0, out[-3000:]
    assert "(Valid)" in out


# ---------------------------------------------------------------------------
# the reference's per-lookup functions exported by libxsbench.so
# ---------------------------------------------------------------------------
def _host_views(inp, sd):
    a = X.sd_arrays(inp, sd)
    grid_ptr = C.cast(sd.nuclide_grid, C.c_void_p)
    ueg_ptr = C.cast(sd.unionized_energy_array, C.c_void_p) if sd.length_unionized_energy_array else None
    ig_ptr = C.cast(sd.index_grid, C.c_void_p) if sd.length_index_grid else None
    return a, grid_ptr, ueg_ptr, ig_ptr


def _grid_index(lib, inp, sd, p):
    """the grid-level index the reference computes for energy p (Simulation.c:340-352)"""
    if inp.grid_type == X.UNIONIZED:
        return lib.grid_search(sd.length_unionized_energy_array, p,
                               C.cast(sd.unionized_energy_array, C.c_void_p))
    if inp.grid_type == X.HASH:
        return int(p / (1.0 / inp.hash_bins))
    return 0


@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "large_unionized", "large_hash",
                                  "xl45k_hash"])
def test_exported_micro_xs_matches_reference_vectors(name):
    """calculate_micro_xs (host export, run on the GPU) against the reference's
    per-nuclide micro XS, bit for bit, on every nuclide of the golden lookups"""
    inp, sd, dev = data(name)
    lib = X.host_lib()
    lib.move_simulation_data_to_device(inp, 1, sd)
    try:
        g = golden_npz(f"micro_{name}.npz")
        a, grid_ptr, ueg_ptr, ig_ptr = _host_views(inp, sd)
        out = (C.c_double * 5)()
        for k in range(len(g["ids"])):
            p, mat = float(g["energy"][k]), int(g["mat"][k])
            want = g["micro"][g["offsets"][k]:g["offsets"][k + 1]]
            idx = _grid_index(lib, inp, sd, p)
            for j in range(int(a["num_nucs"][mat])):
                nuc = int(a["mats"][mat, j])
                lib.calculate_micro_xs(p, nuc, inp.n_isotopes, inp.n_gridpoints, ueg_ptr, ig_ptr,
                                       grid_ptr, idx, out, inp.grid_type, inp.hash_bins)
                assert same_bits(np.array(out[:]), want[j]), (k, j)
    finally:
        lib.release_device_memory(sd)


@pytest.mark.parametrize("name", list(DATA_CFGS))
def test_exported_macro_xs_matches_reference_vectors(name):
    """calculate_macro_xs (host export, run on the GPU) on the golden lookups,
    one call per lookup, and xsb_hip_macro_host on all of them at once"""
    inp, sd, dev = data(name)
    lib = X.host_lib()
    g = golden_npz(f"lookups_{name}.npz")
    a, grid_ptr, ueg_ptr, ig_ptr = _host_views(inp, sd)
    out = (C.c_double * 5)()
    for k in range(64):
        p, mat = float(g["energy"][k]), int(g["mat"][k])
        lib.calculate_macro_xs(p, mat, inp.n_isotopes, inp.n_gridpoints, sd.num_nucs, sd.concs,
                               ueg_ptr, ig_ptr, grid_ptr, sd.mats, out, inp.grid_type,
                               inp.hash_bins, sd.max_num_nucs)
        assert same_bits(np.array(out[:]), g["macro"][k]), k
    lib.release_device_memory(sd)  # the copy the first call uploaded
    assert same_bits(dev.macro_host(g["energy"], g["mat"]), g["macro"])
    # the FP32 tables through the same entry point: within the certified bound
    mix = dev.macro_host(g["energy"], g["mat"], precision=X.XS_FP32)
    assert np.all(np.abs(mix - g["macro"]) <= 1e-6 * np.abs(g["macro"]))


def test_micro_host_rejects_out_of_range_brackets():
    inp, sd, dev = data("small_hash")
    with pytest.raises(X.XsbError) as e:
        dev.micro_host([0.5], [0], lo=[0], hi=[inp.n_gridpoints])
    assert e.value.code == 1
    with pytest.raises(X.XsbError) as e:
        dev.micro_host([0.5], [inp.n_isotopes], lo=[0], hi=[1])
    assert e.value.code == 1
    with pytest.raises(X.XsbError) as e:
        dev.macro_host([0.5], [12])
    assert e.value.code == 1


def test_history_kernel_ids_and_auto_schedule(capfd):
    """history mode with -k 3 runs (the reference ignores -k there) with the
    published hash; XSBENCH_SCHEDULE_AUTO sends -k 0 event runs to the sorted
    schedule"""
    rc = X.run_main(["-m", "history", "-s", "small", "-k", "3"])
    out = capfd.readouterr().out
    assert rc == 0 and "(Valid)" in out, out[-3000:]
    opts = X.host_lib().xsbench_get_options().contents
    assert opts.last_kernel_id == 6  # 500,000 particles: the automatic choice takes rounds
    lib = X.host_lib()
    inp = X.make_inputs("small", "unionized", "event")
    sd = X.grid_init(inp)
    lib.move_simulation_data_to_device(inp, 1, sd)
    try:
        saved = opts.schedule
        opts.schedule = X.SCHEDULE_AUTO
        assert lib.run_event_based_simulation(inp, sd, 1) % 999983 == X.PUBLISHED[("event", "small")]
        assert opts.last_kernel_id == X.K_MAT_E_SORT
        opts.schedule = X.SCHEDULE_ASKED
        assert lib.run_event_based_simulation(inp, sd, 1) % 999983 == X.PUBLISHED[("event", "small")]
        assert opts.last_kernel_id == X.K_BASELINE
        opts.schedule = saved
    finally:
        lib.release_device_memory(sd)


def test_device_grid_build_many_nuclides_few_points():
    """many nuclides on a two-point grid: the per-nuclide maxima of the device
    build (n_iso doubles) exceed the build's sort histogram, so they need
    their own buffer; tables equal the host build"""
    inp = X.make_inputs("small", "hash", "event", gridpoints=2, hash_bins=64)
    inp.n_isotopes = 1000
    sd = X.grid_init(inp)
    with X.Device(inp, sd) as host_dev, X.Device(inp, None) as gen_dev:
        a, b = host_dev.download(), gen_dev.download()
        for key in ("nuclide_grid", "index_grid"):
            assert np.array_equal(a[key], b[key]), key
        assert gen_dev.event(50_000, kernel_id=X.K_MAT_E_SORT)["vsum"] == \
            host_dev.event(50_000, kernel_id=X.K_BASELINE)["vsum"]


@pytest.mark.parametrize("name", list(DATA_CFGS))
def test_table_bytes_predicts_the_upload(name):
    """xsb_hip_table_bytes (host arithmetic) gives the sizes the upload keeps;
    no whole unionized index grid stays on the device"""
    inp, sd, dev = data(name)
    want = X.table_bytes(inp, build_fp32=True, build_bf16=True)
    got = dev.info()
    for k, v in want.items():
        if k.startswith("bytes_"):
            assert got[k] == v, k
    if inp.grid_type == X.UNIONIZED:
        assert got["bytes_index"] == 0


def test_xl_unionized_fits_and_matches_the_hash_grid():
    """the XL problem (355 x 238,847 points) on the unionized grid, built on
    the device: about 186 GB of tables, int32 per-material index tables and no
    whole index grid; its verification sum equals the hash grid's on the same
    lookups (the reference's grid-type invariance, README.md:164).  Runs last:
    it first frees the session's other device copies."""
    for name in list(_CACHE):
        _CACHE.pop(name)[2].close()
    n = 2_000_000
    want = {}
    for grid in ("hash", "unionized"):
        inp = X.make_inputs("XL", grid, "event")
        print(f"[gpu-test] XL {grid}: building on the device", flush=True)
        with X.Device(inp, None) as dev:
            info = dev.info()
            if grid == "unionized":
                assert info["bytes_index"] == 0
                assert info == {**info, **{k: v for k, v in X.table_bytes(inp).items()
                                           if k.startswith("bytes_")}}
            r = dev.event(n, 7, kernel_id=X.K_MAT_E_SORT)
            r0 = dev.event(n // 8, 7, kernel_id=X.K_BASELINE)
            print(f"[gpu-test] XL {grid}: {sum(v for k, v in info.items() if k.startswith('bytes_')) / 1e9:.1f} GB "
                  f"of tables, {r['ms_total']:.2f} ms per {n} lookups", flush=True)
            want[grid] = (r["vsum"], r0["vsum"])
    assert want["hash"] == want["unionized"]
