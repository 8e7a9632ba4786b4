"""Pins the oracle (oracle/xsb_oracle.c) before anything is checked against it.

Each function of the oracle is checked against the reference's own outputs
(golden vectors written by tests/golden/make_golden.py from the reference
build, and the hashes published in openmp-threading/io.c:85-106) and against
what the mathematics fixes (brute force, closed forms, toy tables whose
answers are exact), chosen so that a dropped term, a swapped operand or an
off-by-one fails at least one check.
"""
import bisect
import ctypes as C
import fractions
import math

import numpy as np
import pytest

import oracle as O
from conftest import DATA_CFGS, golden_npz

A_LCG = 2806196910506780709
M63 = 1 << 63


def f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# sampling
# ---------------------------------------------------------------------------
def test_lcg_matches_reference_and_closed_form(orc):
    g = golden_npz("sampling.npz")
    for s, val, nxt in zip(g["seeds"][:512], g["lcg_val"][:512], g["lcg_next"][:512]):
        x = C.c_uint64(int(s))
        v = orc.lcg_random_double(C.byref(x))
        assert x.value == int(nxt) and v == val
        # closed form: s' = (a s + 1) mod 2^63, value = s' / 2^63 (Python's
        # int -> float conversion is correctly rounded, the scaling exact)
        s2 = (A_LCG * int(s) + 1) % M63
        assert s2 == int(nxt) and float(s2) / 2.0**63 == val


def _ff_matrix(seed, n):
    """skip-ahead by exponentiating the affine map x -> a x + 1 (mod 2^63)"""
    def mul(P, Q):  # P, Q = (a, c) meaning x -> a x + c
        return ((P[0] * Q[0]) % M63, (P[0] * Q[1] + P[1]) % M63)
    R, B = (1, 0), (A_LCG, 1)
    while n:
        if n & 1:
            R = mul(B, R)
        B = mul(B, B)
        n >>= 1
    return (R[0] * seed + R[1]) % M63


def test_fast_forward_matches_reference_bruteforce_and_matrix_power(orc):
    g = golden_npz("sampling.npz")
    for s, n, out in zip(g["seeds"][:2048], g["ff_n"], g["ff_out"]):
        assert orc.fast_forward_lcg(int(s), int(n)) == int(out)
        assert _ff_matrix(int(s), int(n)) == int(out)
    rng = np.random.default_rng(3)
    for s in rng.integers(0, M63, size=8, dtype=np.uint64):
        x = C.c_uint64(int(s))
        for n in range(0, 300):
            assert orc.fast_forward_lcg(int(s), n) == x.value
            orc.lcg_random_double(C.byref(x))


def test_pick_mat_matches_reference_and_distribution(orc):
    g = golden_npz("sampling.npz")
    for s, m in zip(g["seeds"], g["pick_mat"]):
        x = C.c_uint64(int(s))
        assert orc.pick_mat(C.byref(x)) == m
    # thresholds: dist[i] + dist[i-1] + ... + dist[1] in that order
    dist = [0.140, 0.052, 0.275, 0.134, 0.154, 0.064, 0.066, 0.055, 0.008, 0.015, 0.025, 0.013]
    thr = (C.c_double * 12)()
    orc.pick_mat_thresholds(thr)
    for i in range(12):
        run = 0.0
        for j in range(i, 0, -1):
            run += dist[j]
        assert thr[i] == run
    assert all(thr[i] < thr[i + 1] for i in range(11))
    # frequencies: material i has probability dist[i] for i >= 1, fuel the rest
    x = C.c_uint64(12345)
    counts = np.zeros(12)
    n = 200000
    for _ in range(n):
        counts[orc.pick_mat(C.byref(x))] += 1
    p = np.array(dist[1:])
    assert abs(counts[0] / n - (1 - p.sum())) < 0.004
    assert np.all(np.abs(counts[1:] / n - p) < 0.004)


# ---------------------------------------------------------------------------
# searches: brute-force clamped upper_bound, with ties and out-of-range keys
# ---------------------------------------------------------------------------
def test_grid_search_is_clamped_upper_bound(orc):
    rng = np.random.default_rng(5)
    for n in (2, 3, 4, 7, 16, 33):
        for _ in range(20):
            A = np.sort(rng.integers(0, 6, size=n).astype(np.float64) / 4.0)
            for q in np.concatenate([A, [-1.0, 10.0, 0.1, 0.6, 1.3]]):
                ub = bisect.bisect_right(list(A), q)
                want = min(max(ub - 1, 0), n - 2)
                assert orc.grid_search(n, q, f64p(A)) == want


def test_grid_search_nuclide_respects_range(orc):
    rng = np.random.default_rng(6)
    for n in (3, 5, 12):
        for _ in range(30):
            E = np.sort(rng.integers(0, 5, size=n).astype(np.float64) / 4.0)
            pts = (O.NuclideGridPoint * n)()
            for k in range(n):
                pts[k].energy = E[k]
            lo = int(rng.integers(0, n - 1))
            hi = int(rng.integers(lo + 1, n))
            for q in np.concatenate([E, [-1.0, 5.0, 0.3]]):
                ub = bisect.bisect_right(list(E[lo:hi + 1]), q) + lo
                want = min(max(ub - 1, lo), hi - 1) if hi - lo > 1 else lo
                assert orc.grid_search_nuclide(n, q, pts, lo, hi) == want


# ---------------------------------------------------------------------------
# micro / macro XS on toy tables with exact answers
# ---------------------------------------------------------------------------
def _toy(points_per_nuc, rows):
    """nuclide grid from rows of (E, 5 xs) per nuclide"""
    n_iso = len(rows)
    arr = (O.NuclideGridPoint * (n_iso * points_per_nuc))()
    for i, r in enumerate(rows):
        for k, (e, *xs) in enumerate(r):
            p = arr[i * points_per_nuc + k]
            p.energy = e
            (p.total_xs, p.elastic_xs, p.absorbtion_xs, p.fission_xs, p.nu_fission_xs) = xs
    return arr


def _micro(orc, arr, n_iso, n_gp, p, nuc, grid_type, egrid=None, index=None, idx=-1, bins=1):
    out = (C.c_double * 5)()
    eg = egrid.ctypes.data_as(C.POINTER(C.c_double)) if egrid is not None else None
    ig = index.ctypes.data_as(C.POINTER(C.c_int)) if index is not None else None
    orc.calculate_micro_xs(p, nuc, n_iso, n_gp, eg, ig, arr, idx, out, grid_type, bins)
    return list(out)


def test_micro_xs_interpolates_and_extrapolates_exactly(orc):
    rows = [[(0.25, 1, 2, 3, 4, 5), (0.75, 3, 4, 5, 6, 7), (0.875, 11, 12, 13, 14, 15)]]
    arr = _toy(3, rows)
    # nuclide grid: midpoint, both ends, extrapolation above the last point
    assert _micro(orc, arr, 1, 3, 0.5, 0, O.NUCLIDE) == [2, 3, 4, 5, 6]
    assert _micro(orc, arr, 1, 3, 0.25, 0, O.NUCLIDE) == [1, 2, 3, 4, 5]
    assert _micro(orc, arr, 1, 3, 0.75, 0, O.NUCLIDE) == [3, 4, 5, 6, 7]
    # above the last point: bracket (1, 2), f = (0.875 - 1) / 0.125 = -1
    assert _micro(orc, arr, 1, 3, 1.0, 0, O.NUCLIDE) == [19, 20, 21, 22, 23]
    # below the first point: bracket (0, 1), f = (0.75 - 0) / 0.5 = 1.5
    assert _micro(orc, arr, 1, 3, 0.0, 0, O.NUCLIDE) == [0, 1, 2, 3, 4]
    # unionized: the index grid names the lower point directly; an index of
    # n-1 is moved down one point (Simulation.c:268-271)
    idx = np.array([0, 1, 2], np.int32)
    assert _micro(orc, arr, 1, 3, 0.8, 0, O.UNIONIZED, index=idx, idx=1) == \
        _micro(orc, arr, 1, 3, 0.8, 0, O.UNIONIZED, index=idx, idx=2)


def test_micro_xs_hash_low_edge_branch(orc):
    """Simulation.c:291-292: p <= E[u_low] selects point 0, not u_low."""
    rows = [[(0.125, 1, 1, 1, 1, 1), (0.25, 2, 2, 2, 2, 2), (0.5, 4, 4, 4, 4, 4),
             (0.875, 8, 8, 8, 8, 8)]]
    arr = _toy(4, rows)
    # two bins: bin 0 -> lower bound 0, bin 1 (energy 0.5) -> lower bound 2
    index = np.array([0, 2], np.int32)
    # p == E[u_low] = 0.5 in bin 1: bracket (0, 1) is used, f = (0.25-0.5)/0.125 = -2
    got = _micro(orc, arr, 1, 4, 0.5, 0, O.HASH, index=index, idx=1, bins=2)
    assert got == [4.0] * 5
    # just above: bracket (2, 3), f = (0.875 - 0.625) / 0.375
    got = _micro(orc, arr, 1, 4, 0.625, 0, O.HASH, index=index, idx=1, bins=2)
    assert got == pytest.approx([8 - (0.25 / 0.375) * 4] * 5, rel=1e-15)


def test_macro_xs_sums_in_nuclide_order(orc):
    """conc * micro added nuclide by nuclide: 1e16 + 1 - 1e16 is 0, not 1."""
    rows = [[(0.0, 1, 1, 1, 1, 1), (1.0, 1, 1, 1, 1, 1)]] * 3
    arr = _toy(2, rows)
    num_nucs = np.array([3] + [1] * 11, np.int32)
    mats = np.zeros((12, 3), np.int32)
    mats[0] = [0, 1, 2]
    concs = np.zeros((12, 3))
    concs[0] = [1e16, 1.0, -1e16]
    out = (C.c_double * 5)()
    ig = np.array([0, 0, 0], np.int32)
    orc.calculate_macro_xs(0.5, 0, 3, 2, num_nucs.ctypes.data_as(C.POINTER(C.c_int)),
                           concs.ctypes.data_as(C.POINTER(C.c_double)), None,
                           ig.ctypes.data_as(C.POINTER(C.c_int)), arr,
                           mats.ctypes.data_as(C.POINTER(C.c_int)), out, O.NUCLIDE, 1, 3)
    assert list(out) == [0.0] * 5
    concs[0] = [1.0, 1e16, -1e16]
    orc.calculate_macro_xs(0.5, 0, 3, 2, num_nucs.ctypes.data_as(C.POINTER(C.c_int)),
                           concs.ctypes.data_as(C.POINTER(C.c_double)), None,
                           ig.ctypes.data_as(C.POINTER(C.c_int)), arr,
                           mats.ctypes.data_as(C.POINTER(C.c_int)), out, O.NUCLIDE, 1, 3)
    assert list(out) == [0.0] * 5
    concs[0] = [1e16, -1e16, 1.0]
    orc.calculate_macro_xs(0.5, 0, 3, 2, num_nucs.ctypes.data_as(C.POINTER(C.c_int)),
                           concs.ctypes.data_as(C.POINTER(C.c_double)), None,
                           ig.ctypes.data_as(C.POINTER(C.c_int)), arr,
                           mats.ctypes.data_as(C.POINTER(C.c_int)), out, O.NUCLIDE, 1, 3)
    assert list(out) == [1.0] * 5


def test_argmax_ties_and_floor():
    """max starts at -1.0 and only a strictly larger value moves it."""
    lib = O.Oracle().lib
    f = lib.or_argmax5
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_double)]
    for vals, want in [([1, 3, 3, 2, 0], 1), ([-2, -3, -5, -1.5, -7], 0), ([-2] * 5, 0),
                       ([-2, -0.5, -5, -0.25, -7], 3),
                       ([0, 0, 0, 0, 0], 0), ([0, 0, 0, 0, 1e-300], 4)]:
        arr = (C.c_double * 5)(*vals)
        assert f(arr) == want


# ---------------------------------------------------------------------------
# data construction against the reference's arrays
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "small_nuclide", "large_hash",
                                  "xl45k_hash"])
def test_grid_init_digests_match_reference(orc, golden_digests, name):
    size, grid, gp = DATA_CFGS[name]
    inp = O.make_inputs(size, grid, "event", gridpoints=gp)
    sd = orc.grid_init(inp)
    try:
        d = orc.digests(inp, sd)
        g = {k: v for k, v in golden_digests[name].items() if k in d}
        assert d == g
    finally:
        orc.free_sd(C.byref(sd))


def test_index_grid_equals_bruteforce_bracket(orc):
    """unionized index grid == clamp(upper_bound(E_i, U[e]) - 1, 0, n - 2)"""
    inp = O.make_inputs("small", "unionized", "event", gridpoints=40)
    sd = orc.grid_init(inp)
    try:
        n_iso, n_gp = inp.n_isotopes, inp.n_gridpoints
        g = np.ctypeslib.as_array(C.cast(sd.nuclide_grid, C.POINTER(C.c_double)), shape=(n_iso * n_gp, 6))
        E = g[:, 0].reshape(n_iso, n_gp)
        U = np.ctypeslib.as_array(sd.unionized_energy_array, shape=(n_iso * n_gp,))
        ig = np.ctypeslib.as_array(sd.index_grid, shape=(n_iso * n_gp, n_iso))
        assert np.all(np.diff(U) >= 0) and np.array_equal(np.sort(E.ravel()), U)
        for i in range(n_iso):
            want = np.clip(np.searchsorted(E[i], U, side="right") - 1, 0, n_gp - 2)
            assert np.array_equal(ig[:, i], want)
    finally:
        orc.free_sd(C.byref(sd))


# ---------------------------------------------------------------------------
# per-lookup golden vectors (reference's calculate_macro_xs etc.)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "small_nuclide", "large_hash",
                                  "xl45k_hash"])
def test_lookups_match_reference_vectors(orc, name):
    size, grid, gp = DATA_CFGS[name]
    inp = O.make_inputs(size, grid, "event", gridpoints=gp)
    sd = orc.grid_init(inp)
    try:
        g = golden_npz(f"lookups_{name}.npz")
        E, m, idx, macro, am = orc.event_detail(inp, sd, g["ids"])
        assert np.array_equal(E, g["energy"]) and np.array_equal(m, g["mat"])
        if grid == "unionized":
            assert np.array_equal(idx, g["ueg_idx"])
        assert np.array_equal(macro.view(np.uint64), g["macro"].view(np.uint64))
        assert np.array_equal(am, g["argmax"])
        if name in ("small_hash", "large_hash", "xl45k_hash"):
            mg = golden_npz(f"micro_{name}.npz")
            for k in range(len(mg["ids"])):
                got = orc.micro_detail(inp, sd, float(mg["energy"][k]), int(mg["mat"][k]))
                want = mg["micro"][mg["offsets"][k]:mg["offsets"][k + 1]]
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    finally:
        orc.free_sd(C.byref(sd))


@pytest.mark.parametrize("name", ["small_unionized", "small_hash", "large_hash"])
def test_history_particles_match_reference(orc, name):
    size, grid, gp = DATA_CFGS[name]
    g = golden_npz(f"history_{name}.npz")
    n = len(g["sums"])
    inp = O.make_inputs(size, grid, "history", lookups=int(g["lookups"]), particles=n, gridpoints=gp)
    sd = orc.grid_init(inp)
    try:
        sums, seeds = orc.history_detail(inp, sd, 0, n)
        assert np.array_equal(sums, g["sums"]) and np.array_equal(seeds, g["final_seed"])
    finally:
        orc.free_sd(C.byref(sd))


# ---------------------------------------------------------------------------
# end-to-end hashes: published (io.c:85-106) and reference-computed
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("grid", ["unionized", "hash", "nuclide"])
def test_published_small_event_hash_every_grid(orc, grid):
    inp = O.make_inputs("small", grid, "event")
    sd = orc.grid_init(inp)
    try:
        v = orc.run_event(C.byref(inp), C.byref(sd), 0)
        assert v % 999983 == O.PUBLISHED[("event", "small")]
    finally:
        orc.free_sd(C.byref(sd))


def test_published_small_history_and_golden_small_configs(orc, golden_hashes):
    inp = O.make_inputs("small", "unionized", "history")
    sd = orc.grid_init(inp)
    try:
        assert orc.run_history(C.byref(inp), C.byref(sd), 0) % 999983 == O.PUBLISHED[("history", "small")]
        e1 = O.make_inputs("small", "unionized", "event", lookups=1000000)
        want = golden_hashes["event_small_unionized_1M"]["vsum"]
        assert orc.run_event(C.byref(e1), C.byref(sd), 0) == want
        assert orc.run_event_sorted(C.byref(e1), C.byref(sd), 0) == want
    finally:
        orc.free_sd(C.byref(sd))


@pytest.mark.slow
def test_published_large_hashes(orc):
    for grid in ("hash", "unionized"):
        inp = O.make_inputs("large", grid, "event")
        sd = orc.grid_init(inp)
        try:
            assert orc.run_event(C.byref(inp), C.byref(sd), 0) % 999983 == O.PUBLISHED[("event", "large")]
            if grid == "hash":
                h = O.make_inputs("large", grid, "history")
                assert orc.run_history(C.byref(h), C.byref(sd), 0) % 999983 == O.PUBLISHED[("history", "large")]
        finally:
            orc.free_sd(C.byref(sd))


def test_golden_history_hashes_at_20k(orc, golden_hashes):
    for key in ("history_small_hash_20k", "history_large_hash_20k", "history_xl45k_hash_20k"):
        h = golden_hashes[key]
        size, grid, gp = DATA_CFGS[h["data"]]
        inp = O.make_inputs(size, grid, "history", lookups=h["lookups"], particles=h["particles"],
                            gridpoints=gp)
        sd = orc.grid_init(inp)
        try:
            assert orc.run_history(C.byref(inp), C.byref(sd), 0) == h["vsum"]
        finally:
            orc.free_sd(C.byref(sd))


# ---------------------------------------------------------------------------
# algorithmic-byte counter: distinct words read, counted independently
# ---------------------------------------------------------------------------
def _count_words(inp, sd, arrays, i):
    """distinct (array, index) words the reference reads for lookup i, by
    instrumenting a Python walk of Simulation.c:241-375"""
    seed = O.Oracle().fast_forward_lcg(1070, 2 * i)
    s = C.c_uint64(seed)
    orc = O.Oracle()
    p = orc.lcg_random_double(C.byref(s))
    mat = orc.pick_mat(C.byref(s))
    n_iso, n_gp = inp.n_isotopes, inp.n_gridpoints
    words = set()
    E = arrays["E"]
    words.add(("num_nucs", mat))

    def bisect_walk(get, lo, hi, tag):
        while hi - lo > 1:
            mid = lo + (hi - lo) // 2
            words.add((tag, mid))
            if get(mid) > p:
                hi = mid
            else:
                lo = mid
        return lo

    idx = -1
    if inp.grid_type == O.UNIONIZED:
        U = arrays["U"]
        idx = bisect_walk(lambda k: U[k], 0, n_iso * n_gp - 1, "U")
    elif inp.grid_type == O.HASH:
        idx = int(p / (1.0 / inp.hash_bins))
    for j in range(sd.num_nucs[mat]):
        words.add(("mats", mat, j))
        words.add(("concs", mat, j))
        nuc = sd.mats[mat * sd.max_num_nucs + j]
        row = E[nuc]
        if inp.grid_type == O.UNIONIZED:
            words.add(("ig", idx, nuc))
            lower = arrays["IG"][idx * n_iso + nuc]
        elif inp.grid_type == O.NUCLIDE:
            lower = bisect_walk(lambda k: row[k], 0, n_gp - 1, ("E", nuc))
        else:
            IG = arrays["IG"]
            u_low = IG[idx * n_iso + nuc]
            words.add(("ig", idx, nuc))
            if idx == inp.hash_bins - 1:
                u_high = n_gp - 1
            else:
                u_high = IG[(idx + 1) * n_iso + nuc] + 1
                words.add(("ig", idx + 1, nuc))
            words.add((("E", nuc), u_low))
            words.add((("E", nuc), u_high))
            if p <= row[u_low]:
                lower = 0
            elif p >= row[u_high]:
                lower = n_gp - 1
            else:
                lower = bisect_walk(lambda k: row[k], u_low, u_high, ("E", nuc))
        if lower == n_gp - 1:
            lower -= 1
        for k in (lower, lower + 1):
            words.add((("E", nuc), k))
            words.add((("X", nuc), k))
    size = {"num_nucs": 4, "mats": 4, "concs": 8, "U": 8, "ig": 4}
    total = 0
    for w in words:
        tag = w[0]
        if isinstance(tag, tuple):
            total += 8 if tag[0] == "E" else 40
        else:
            total += size[tag]
    return total


@pytest.mark.parametrize("grid", ["unionized", "hash", "nuclide"])
def test_balg_counter_matches_word_walk(orc, grid):
    inp = O.make_inputs("small", grid, "event", gridpoints=200)
    sd = orc.grid_init(inp)
    try:
        n_iso, n_gp = inp.n_isotopes, inp.n_gridpoints
        g = np.ctypeslib.as_array(C.cast(sd.nuclide_grid, C.POINTER(C.c_double)), shape=(n_iso, n_gp, 6))
        arrays = {"E": g[:, :, 0]}
        if grid == "unionized":
            arrays["U"] = np.ctypeslib.as_array(sd.unionized_energy_array, shape=(n_iso * n_gp,))
        if grid != "nuclide":
            arrays["IG"] = np.ctypeslib.as_array(sd.index_grid, shape=(sd.length_index_grid,))
        ids = list(range(40))
        want = np.mean([_count_words(inp, sd, arrays, i) for i in ids])
        got = orc.balg_event(inp, sd, 0, len(ids))
        assert got == pytest.approx(want, rel=0, abs=1e-9)
    finally:
        orc.free_sd(C.byref(sd))


# ---------------------------------------------------------------------------
# planted mistakes: the golden micro vectors reject plausible slips
# ---------------------------------------------------------------------------
def test_golden_micro_vectors_reject_planted_mistakes(orc):
    inp = O.make_inputs("small", "unionized", "event")
    sd = orc.grid_init(inp)
    try:
        n_iso, n_gp = inp.n_isotopes, inp.n_gridpoints
        g = np.ctypeslib.as_array(C.cast(sd.nuclide_grid, C.POINTER(C.c_double)), shape=(n_iso, n_gp, 6))
        U = np.ctypeslib.as_array(sd.unionized_energy_array, shape=(n_iso * n_gp,))
        IG = np.ctypeslib.as_array(sd.index_grid, shape=(n_iso * n_gp, n_iso))
        mg = golden_npz("micro_small_unionized.npz")

        def lerp(p, nuc, variant):
            idx = min(max(int(np.searchsorted(U, p, side="right")) - 1, 0), len(U) - 2)
            k = int(IG[idx, nuc])
            if variant == "off_by_one":
                k = min(k + 1, n_gp - 2)
            lo, hi = g[nuc, k], g[nuc, k + 1]
            if variant == "swapped":
                lo, hi = hi, lo
            f = (hi[0] - p) / (hi[0] - lo[0])
            if variant == "sign":
                return hi[1:] + f * (hi[1:] - lo[1:])
            if variant == "fma":
                return np.array([float(fractions.Fraction(h) - fractions.Fraction(f) * fractions.Fraction(h - l))
                                 for h, l in zip(hi[1:], lo[1:])])
            return hi[1:] - f * (hi[1:] - lo[1:])

        caught = {v: False for v in ("off_by_one", "swapped", "sign", "fma")}
        for k in range(len(mg["ids"])):
            p, mat = float(mg["energy"][k]), int(mg["mat"][k])
            want = mg["micro"][mg["offsets"][k]:mg["offsets"][k + 1]]
            for j in range(sd.num_nucs[mat]):
                nuc = sd.mats[mat * sd.max_num_nucs + j]
                assert np.array_equal(lerp(p, nuc, "exact"), want[j])
                for v in caught:
                    if not np.array_equal(lerp(p, nuc, v), want[j]):
                        caught[v] = True
        assert all(caught.values()), caught
    finally:
        orc.free_sd(C.byref(sd))


# ---------------------------------------------------------------------------
# batched workload (no reference counterpart): two routes through the oracle
# ---------------------------------------------------------------------------
def test_batch_plan_and_batch_hash_by_brute_force(orc):
    sizes, seeds, thr = orc.batch_plan(99, 64, 1000, 2000000)
    assert sizes.min() >= 1000 and sizes.max() <= 2000000
    # log-uniform: the median of log-sizes sits near the middle
    assert abs(np.median(np.log(sizes)) - 0.5 * (math.log(1000) + math.log(2e6))) < 1.0
    assert np.all(np.diff(thr[:, :11], axis=1) > 0) and np.all(thr[:, 11] >= 1.0)
    inp = O.make_inputs("small", "unionized", "event")
    sd = orc.grid_init(inp)
    try:
        for b in range(3):
            n = 3000
            E = np.empty(n)
            M = np.empty(n, np.int32)
            for i in range(n):
                E[i], M[i] = orc.batch_sample(int(seeds[b]), thr[b], i)
            _, am = orc.lookup_detail(inp, sd, E, M)
            assert orc.run_batch(inp, sd, int(seeds[b]), thr[b], n) == int(am.sum()) + n
            # material frequencies follow the batch's own CDF
            cdf = np.concatenate([[0], thr[b, :11], [1.0]])
            freq = np.bincount(M, minlength=12) / n
            assert np.all(np.abs(freq - np.diff(cdf)) < 0.035)
    finally:
        orc.free_sd(C.byref(sd))
