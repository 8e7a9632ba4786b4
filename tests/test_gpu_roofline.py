"""The compulsory-byte counter (xsb_hip_lines_*) against an independent walk
of the same reads in numpy: for lookups taken from the oracle, the distinct
128-byte lines of each device table that the lookups read, computed from
the host tables and the device layout (per-material index rows, interval
records, the unionized search's bracket words and probes; hash index words,
bin ranges and bracket points)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import xsbench_amd as X

pytestmark = pytest.mark.gpu

LINE = 128
REG = {name: i for i, name in enumerate(X.LINE_REGIONS)}


def _lines(intervals):
    """distinct line ids of [start, end] byte intervals (inclusive)"""
    s = set()
    for a, b in intervals:
        s.update(range(a // LINE, b // LINE + 1))
    return len(s) * LINE


def _setup(size, grid, n, first=1_000_003):
    inp = X.make_inputs(size, grid, "event")
    sd = X.grid_init(inp)
    oi = O.Inputs.from_buffer_copy(inp)
    oi._hm = inp._hm
    oi.HM = oi._hm
    osd = O.SimulationData.from_buffer_copy(sd)
    E, m, idx, macro, am = O.Oracle().event_detail(oi, osd, np.arange(first, first + n))
    return inp, sd, E, m, idx


def _bracket_hash(g, nuc, n_gp, IG, idx, bins, p):
    """Simulation.c:273-302: the hash grid's lower point"""
    u_low = int(IG[idx, nuc])
    u_high = n_gp - 1 if idx == bins - 1 else int(IG[idx + 1, nuc]) + 1
    E = g[nuc, :, 0]
    if p <= E[u_low]:
        lo = 0
    elif p >= E[u_high]:
        lo = n_gp - 1
    else:
        a, b = u_low, u_high
        while b - a > 1:
            mid = a + (b - a) // 2
            if E[mid] > p:
                b = mid
            else:
                a = mid
        lo = a
    return u_low, u_high, (lo - 1 if lo == n_gp - 1 else lo)


def _walk_unionized(inp, sd, E, m, rec_bytes):
    """per-region line bytes of the unionized lookups (E, m): search bracket
    words and probes, per-material uint16 index rows, interval records"""
    a = X.sd_arrays(inp, sd)
    n_iso, n_gp = inp.n_isotopes, inp.n_gridpoints
    U = a["unionized_energy_array"]
    IG = a["index_grid"].reshape(-1, n_iso)
    nn, mats = a["num_nucs"], a["mats"]
    pitch = [(int(v) + 7) // 8 * 8 for v in nn]  # rows padded to 8 entries
    off = np.concatenate([[0], np.cumsum([len(U) * p for p in pitch])])
    acc = np.searchsorted(U, np.arange(65537) / 65536.0, side="right")
    ueg, accw, rows, recs = [], [], [], []
    for p, mat in zip(E, m):
        b = int(p * 65536)
        accw.append((4 * b, 4 * b + 7))
        lo, hi = int(acc[b]) - 1, int(acc[b + 1])
        while hi - lo > 1:
            mid = lo + (hi - lo) // 2
            ueg.append((8 * mid, 8 * mid + 7))
            if U[mid] > p:
                hi = mid
            else:
                lo = mid
        i = min(max(hi - 1, 0), len(U) - 2)
        start = 2 * (int(off[mat]) + i * pitch[mat])
        rows.append((start, start + 2 * int(nn[mat]) - 1))
        for j in range(int(nn[mat])):
            nuc = int(mats[mat, j])
            k = int(IG[i, nuc])
            k = k - 1 if k == n_gp - 1 else k
            r = (nuc * n_gp + k) * rec_bytes
            recs.append((r, r + rec_bytes - 1))
    return {"ueg": _lines(ueg), "acc": _lines(accw), "midx": _lines(rows), "rec": _lines(recs)}


def _walk_hash(inp, sd, E, m):
    """per-region line bytes of the hash-grid lookups (E, m): the bin's two
    index words, the energies of its range and the two bracket points"""
    a = X.sd_arrays(inp, sd)
    n_iso, n_gp, bins = inp.n_isotopes, inp.n_gridpoints, inp.hash_bins
    g = a["nuclide_grid"].reshape(n_iso, n_gp, 6)
    IG = a["index_grid"].reshape(-1, n_iso)
    pitch = (n_iso + 3) // 4 * 4
    nn, mats = a["num_nucs"], a["mats"]
    words, pts = [], []
    for p, mat in zip(E, m):
        b = min(int(p / (1.0 / bins)), bins - 1)
        for j in range(int(nn[mat])):
            nuc = int(mats[mat, j])
            w = 4 * (b * pitch + nuc)
            words.append((w, w + 3))
            if b != bins - 1:
                words.append((w + 4 * pitch, w + 4 * pitch + 3))
            u_low, u_high, lo = _bracket_hash(g, nuc, n_gp, IG, b, bins, p)
            base = nuc * n_gp * 48
            pts.append((base + 48 * u_low, base + 48 * u_high + 7))
            pts.append((base + 48 * lo, base + 48 * lo + 95))
    return {"index": _lines(words), "grid": _lines(pts)}


@pytest.mark.parametrize("prec", [X.XS_FP64, X.XS_FP32])
def test_line_count_unionized_matches_walk(prec):
    n, first = 20_000, 1_000_003
    inp, sd, E, m, idx = _setup("small", "unionized", n, first)
    w = _walk_unionized(inp, sd, E, m, 96 if prec == X.XS_FP64 else 64)
    with X.Device(inp, sd, build_fp32=True) as dev:
        got = dev.lines_event(n, first, kernel_id=X.K_MAT_E_SORT, precision=prec)
    rb = got["region_bytes"]
    assert rb[REG["ueg"]] == w["ueg"]
    assert rb[REG["acc"]] == w["acc"]
    assert rb[REG["midx"]] == w["midx"]
    assert rb[REG["ivl" if prec == X.XS_FP64 else "ivl32"]] == w["rec"]
    assert got["bytes_tables"] == w["ueg"] + w["acc"] + w["midx"] + w["rec"]
    assert got["bytes_schedule"] == (n * 8 + 127) // 128 * 128
    assert got["bytes"] == got["bytes_tables"] + got["bytes_schedule"] + got["bytes_small"]


def test_line_count_hash_matches_walk():
    n, first = 20_000, 1_000_003
    inp, sd, E, m, idx = _setup("small", "hash", n, first)
    w = _walk_hash(inp, sd, E, m)
    with X.Device(inp, sd) as dev:
        got = dev.lines_event(n, first, kernel_id=X.K_MAT_E_SORT)
    rb = got["region_bytes"]
    assert rb[REG["index"]] == w["index"]
    assert rb[REG["grid"]] == w["grid"]
    assert got["bytes_tables"] == w["index"] + w["grid"]


def _walk_hash_history(inp, sd, E, m):
    """per-region line bytes of the history rounds' hash lookups (E, m): the
    bin's 96-byte window record (index words and ten energies), eight more
    energies of the nuclide's row where the bin spans further, or its
    bisection over the row, and the two bracket points"""
    a = X.sd_arrays(inp, sd)
    n_iso, n_gp, bins = inp.n_isotopes, inp.n_gridpoints, inp.hash_bins
    g = a["nuclide_grid"].reshape(n_iso, n_gp, 6)
    IG = a["index_grid"].reshape(-1, n_iso)
    pitch = (n_iso + 3) // 4 * 4
    e_pitch = (n_gp + 16 + 7) // 8 * 8
    nn, mats = a["num_nucs"], a["mats"]
    win, row, pts = [], [], []
    for p, mat in zip(E, m):
        b = min(int(p / (1.0 / bins)), bins - 1)
        for j in range(int(nn[mat])):
            nuc = int(mats[mat, j])
            w = 96 * (b * pitch + nuc)
            win.append((w, w + 95))
            u_low, u_high, lo = _bracket_hash(g, nuc, n_gp, IG, b, bins, p)
            base, top = nuc * e_pitch * 8, u_high - (u_low & ~1)
            En = g[nuc, :, 0]
            if top > 17:
                row.append((base + 8 * u_high, base + 8 * u_high + 7))
                if En[u_low] < p < En[u_high]:
                    x, y = u_low, u_high
                    while y - x > 1:
                        mid = x + (y - x) // 2
                        row.append((base + 8 * mid, base + 8 * mid + 7))
                        x, y = (x, mid) if En[mid] > p else (mid, y)
            elif top > 9:
                s0 = (u_low & ~1) + 10
                row.append((base + 8 * s0, base + 8 * s0 + 63))
            pb = nuc * n_gp * 48
            pts.append((pb + 48 * lo, pb + 48 * lo + 95))
    return {"hwin": _lines(win), "esoa": _lines(row), "grid": _lines(pts)}


def test_line_count_history_matches_walk(orc):
    """round r's lines are those of lookup r of every particle: the chains
    replayed on the host (the reference's LCG, pick_mat and fast-forward,
    the oracle's macro XS for the channels above 1.0)"""
    P, R = 3000, 2
    inp = X.make_inputs("small", "hash", "event")
    sd = X.grid_init(inp)
    oi = O.Inputs.from_buffer_copy(inp)
    oi._hm = inp._hm
    oi.HM = oi._hm
    osd = O.SimulationData.from_buffer_copy(sd)
    lib = X.host_lib()
    # particle q starts R * 10 * q draws into the stream (Simulation.c:167; < 2^31 here)
    seeds = [C.c_uint64(lib.fast_forward_LCG(X.STARTING_SEED, q * R * 10)) for q in range(P)]
    want = 0
    for r in range(R):
        E = np.array([lib.LCG_random_double(C.byref(s)) for s in seeds])
        m = np.array([lib.pick_mat(C.byref(s)) for s in seeds], np.int32)
        w = _walk_hash_history(inp, sd, E, m)
        want += w["hwin"] + w["esoa"] + w["grid"]
        macro, _ = orc.lookup_detail(oi, osd, E, m)
        for q, row in enumerate(macro):
            nf = int((row > 1.0).sum())
            if nf:
                seeds[q] = C.c_uint64(lib.fast_forward_LCG(seeds[q].value, nf))
    with X.Device(inp, sd) as dev:
        got = dev.lines_history(P, R)
    assert got["launches"] == R
    assert got["bytes_tables"] == want


def test_line_count_batched_matches_walk():
    inp = X.make_inputs("small", "unionized", "event")
    sd = X.grid_init(inp)
    sizes, seeds, thr = X.make_batches(7, 6, 100, 3000)
    lib = X.host_lib()
    E, m = [], []
    e, mm = C.c_double(), C.c_int()
    for b in range(len(sizes)):
        t = np.ascontiguousarray(thr[b])
        for i in range(int(sizes[b])):
            lib.xsbench_batch_sample(int(seeds[b]), t.ctypes.data, i, C.byref(e), C.byref(mm))
            E.append(e.value)
            m.append(mm.value)
    w = _walk_unionized(inp, sd, E, m, 96)
    with X.Device(inp, sd) as dev:
        got = dev.lines_batched(sizes, seeds, thr)
    assert got["bytes_tables"] == w["ueg"] + w["acc"] + w["midx"] + w["rec"]
