"""ctypes loader for the oracle (test infrastructure only).

Two native libraries sit behind this module:

* ``oracle/libxsb_oracle.so`` -- this project's CPU restatement of the
  reference lookup path (``xsb_oracle.c``), always built by ``make -C oracle``;
* ``oracle/_ref/libxsbench_ref.so`` -- the reference's own openmp-threading
  sources compiled by ``oracle/ref.mk`` (only where the reference is mounted;
  the built file travels with the tree).

Only tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may use this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libxsb_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libxsbench_ref.so")
REF_EXE = os.path.join(HERE, "_ref", "XSBench")
REF_SRC = "/root/reference/openmp-threading"

UNIONIZED, NUCLIDE, HASH = 0, 1, 2
HISTORY_BASED, EVENT_BASED = 1, 2
STARTING_SEED = 1070

# Published verification hashes, openmp-threading/io.c:85-106
PUBLISHED = {
    ("event", "small"): 945990,
    ("event", "large"): 952131,
    ("history", "small"): 941535,
    ("history", "large"): 954318,
}


class NuclideGridPoint(C.Structure):
    """openmp-threading/XSbench_header.h:54-61"""
    _fields_ = [(n, C.c_double) for n in (
        "energy", "total_xs", "elastic_xs", "absorbtion_xs", "fission_xs", "nu_fission_xs")]


class Inputs(C.Structure):
    """openmp-threading/XSbench_header.h:63-75"""
    _fields_ = [
        ("nthreads", C.c_int), ("n_isotopes", C.c_long), ("n_gridpoints", C.c_long),
        ("lookups", C.c_int), ("HM", C.c_char_p), ("grid_type", C.c_int),
        ("hash_bins", C.c_int), ("particles", C.c_int), ("simulation_method", C.c_int),
        ("binary_mode", C.c_int), ("kernel_id", C.c_int),
    ]


class SimulationData(C.Structure):
    """openmp-threading/XSbench_header.h:77-102 (non-AML layout)"""
    _fields_ = [
        ("num_nucs", C.POINTER(C.c_int)), ("concs", C.POINTER(C.c_double)),
        ("mats", C.POINTER(C.c_int)), ("unionized_energy_array", C.POINTER(C.c_double)),
        ("index_grid", C.POINTER(C.c_int)), ("nuclide_grid", C.POINTER(NuclideGridPoint)),
        ("length_num_nucs", C.c_int), ("length_concs", C.c_int), ("length_mats", C.c_int),
        ("length_unionized_energy_array", C.c_int), ("length_index_grid", C.c_long),
        ("length_nuclide_grid", C.c_int), ("max_num_nucs", C.c_int),
        ("p_energy_samples", C.POINTER(C.c_double)), ("length_p_energy_samples", C.c_int),
        ("mat_samples", C.POINTER(C.c_int)), ("length_mat_samples", C.c_int),
    ]


def make_inputs(size="small", grid="unionized", method="event", lookups=None, particles=None,
                gridpoints=None, hash_bins=10000, nthreads=None):
    """Inputs as read_CLI (openmp-threading/io.c:227-454) would build them for
    ``-s size -G grid -m method [-l lookups] [-p particles] [-g gridpoints]``,
    with -l/-p given before -m (so event mode keeps an explicit -l)."""
    gt = {"unionized": UNIONIZED, "nuclide": NUCLIDE, "hash": HASH}[grid]
    n_iso = 68 if size.lower() == "small" else 355
    n_gp = 11303
    if gridpoints is not None:
        n_gp = int(gridpoints)
    elif size.lower() == "xl":
        n_gp = 238847
    elif size.lower() == "xxl":
        n_gp = int(238847 * 2.1)
    if method == "event":
        lk = int(lookups) if lookups is not None else (
            34 * (int(particles) if particles is not None else 500000))
        pt = 0 if lookups is None and particles is None else (
            int(particles) if particles is not None else 500000)
        sm = EVENT_BASED
    else:
        lk = int(lookups) if lookups is not None else 34
        pt = int(particles) if particles is not None else 500000
        sm = HISTORY_BASED
    inp = Inputs()
    inp.nthreads = nthreads or os.cpu_count()
    inp.n_isotopes = n_iso
    inp.n_gridpoints = n_gp
    inp.lookups = lk
    inp._hm = size.encode()
    inp.HM = inp._hm
    inp.grid_type = gt
    inp.hash_bins = hash_bins
    inp.particles = pt
    inp.simulation_method = sm
    inp.binary_mode = 0
    inp.kernel_id = 0
    return inp


_P = C.POINTER


def _bind(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


class Oracle:
    """This project's restatement (libxsb_oracle.so)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path)
        u64p, i32p, i64p, f64p = _P(C.c_uint64), _P(C.c_int), _P(C.c_long), _P(C.c_double)
        PI, PS = _P(Inputs), _P(SimulationData)
        self.lcg_random_double = _bind(L, "or_lcg_random_double", C.c_double, [u64p])
        self.fast_forward_lcg = _bind(L, "or_fast_forward_lcg", C.c_uint64, [C.c_uint64, C.c_uint64])
        self.pick_mat = _bind(L, "or_pick_mat", C.c_int, [u64p])
        self.pick_mat_thresholds = _bind(L, "or_pick_mat_thresholds", None, [f64p])
        self.grid_search = _bind(L, "or_grid_search", C.c_long, [C.c_long, C.c_double, f64p])
        self.grid_search_nuclide = _bind(L, "or_grid_search_nuclide", C.c_long,
                                         [C.c_long, C.c_double, _P(NuclideGridPoint), C.c_long, C.c_long])
        self.calculate_micro_xs = _bind(L, "or_calculate_micro_xs", None, [
            C.c_double, C.c_int, C.c_long, C.c_long, f64p, i32p, _P(NuclideGridPoint), C.c_long,
            f64p, C.c_int, C.c_int])
        self.calculate_macro_xs = _bind(L, "or_calculate_macro_xs", None, [
            C.c_double, C.c_int, C.c_long, C.c_long, i32p, f64p, f64p, i32p, _P(NuclideGridPoint),
            i32p, f64p, C.c_int, C.c_int, C.c_int])
        self.grid_init_raw = _bind(L, "or_grid_init", C.c_int, [PI, PS])
        self.free_sd = _bind(L, "or_free_simulation_data", None, [PS])
        self.fnv1a = _bind(L, "or_fnv1a", C.c_uint64, [C.c_void_p, C.c_size_t, C.c_uint64])
        self.run_event = _bind(L, "or_run_event", C.c_uint64, [PI, PS, C.c_int])
        self.run_event_range = _bind(L, "or_run_event_range", C.c_uint64,
                                     [PI, PS, C.c_long, C.c_long, C.c_int])
        self.run_history = _bind(L, "or_run_history", C.c_uint64, [PI, PS, C.c_int])
        self.run_event_sorted = _bind(L, "or_run_event_sorted", C.c_uint64, [PI, PS, C.c_int])
        self.event_detail_raw = _bind(L, "or_event_detail", None,
                                      [PI, PS, C.c_long, i64p, f64p, i32p, i64p, f64p, i32p])
        self.lookup_detail_raw = _bind(L, "or_lookup_detail", None,
                                       [PI, PS, C.c_long, f64p, i32p, f64p, i32p])
        self.micro_detail_raw = _bind(L, "or_micro_detail", None, [PI, PS, C.c_double, C.c_int, f64p])
        self.history_detail_raw = _bind(L, "or_history_detail", None,
                                        [PI, PS, C.c_long, C.c_long, u64p, u64p])
        self.batch_plan_raw = _bind(L, "or_batch_plan", None,
                                    [C.c_uint64, C.c_int, C.c_long, C.c_long, i64p, u64p, f64p])
        self.batch_sample_raw = _bind(L, "or_batch_sample", None,
                                      [C.c_uint64, f64p, C.c_long, f64p, i32p])
        self.run_batch_raw = _bind(L, "or_run_batch", C.c_uint64,
                                   [PI, PS, C.c_uint64, f64p, C.c_long, C.c_int])
        self.balg_event_raw = _bind(L, "or_balg_event", C.c_double,
                                    [PI, PS, C.c_long, C.c_long, _P(C.c_int * 5), C.c_int])

    # -- convenience wrappers --------------------------------------------------
    def grid_init(self, inp):
        sd = SimulationData()
        rc = self.grid_init_raw(C.byref(inp), C.byref(sd))
        if rc:
            raise MemoryError("or_grid_init failed")
        return sd

    def event_detail(self, inp, sd, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        n = len(ids)
        E = np.empty(n, np.float64)
        m = np.empty(n, np.int32)
        idx = np.empty(n, np.int64)
        macro = np.empty((n, 5), np.float64)
        am = np.empty(n, np.int32)
        self.event_detail_raw(C.byref(inp), C.byref(sd), n, _ptr(ids, C.c_long), _ptr(E, C.c_double),
                              _ptr(m, C.c_int), _ptr(idx, C.c_long), _ptr(macro, C.c_double),
                              _ptr(am, C.c_int))
        return E, m, idx, macro, am

    def lookup_detail(self, inp, sd, energy, mat):
        energy = np.ascontiguousarray(energy, np.float64)
        mat = np.ascontiguousarray(mat, np.int32)
        n = len(energy)
        macro = np.empty((n, 5), np.float64)
        am = np.empty(n, np.int32)
        self.lookup_detail_raw(C.byref(inp), C.byref(sd), n, _ptr(energy, C.c_double),
                               _ptr(mat, C.c_int), _ptr(macro, C.c_double), _ptr(am, C.c_int))
        return macro, am

    def micro_detail(self, inp, sd, energy, mat):
        n = sd.num_nucs[mat]
        out = np.empty((n, 5), np.float64)
        self.micro_detail_raw(C.byref(inp), C.byref(sd), energy, mat, _ptr(out, C.c_double))
        return out

    def history_detail(self, inp, sd, first, n):
        v = np.empty(n, np.uint64)
        s = np.empty(n, np.uint64)
        self.history_detail_raw(C.byref(inp), C.byref(sd), first, n, _ptr(v, C.c_uint64),
                                _ptr(s, C.c_uint64))
        return v, s

    def batch_plan(self, seed, n_batches, lo, hi):
        sizes = np.empty(n_batches, np.int64)
        seeds = np.empty(n_batches, np.uint64)
        thr = np.empty((n_batches, 12), np.float64)
        self.batch_plan_raw(seed, n_batches, lo, hi, _ptr(sizes, C.c_long), _ptr(seeds, C.c_uint64),
                            _ptr(thr, C.c_double))
        return sizes, seeds, thr

    def batch_sample(self, batch_seed, thr, i):
        thr = np.ascontiguousarray(thr, np.float64)
        e = C.c_double()
        m = C.c_int()
        self.batch_sample_raw(batch_seed, _ptr(thr, C.c_double), i, C.byref(e), C.byref(m))
        return e.value, m.value

    def run_batch(self, inp, sd, batch_seed, thr, size, nthreads=0):
        thr = np.ascontiguousarray(thr, np.float64)
        return self.run_batch_raw(C.byref(inp), C.byref(sd), batch_seed, _ptr(thr, C.c_double),
                                  size, nthreads)

    def balg_event(self, inp, sd, first, count, fmt=(8, 40, 4, 8, 4), nthreads=0):
        f = (C.c_int * 5)(*fmt)
        return self.balg_event_raw(C.byref(inp), C.byref(sd), first, count, C.byref(f), nthreads)

    def digests(self, inp, sd):
        """FNV-1a digests of every SimulationData array (valid entries only)."""
        return sd_digests(self.fnv1a, inp, sd)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def sd_digests(fnv, inp, sd):
    out = {}
    n_pts = inp.n_isotopes * inp.n_gridpoints
    out["nuclide_grid"] = fnv(C.cast(sd.nuclide_grid, C.c_void_p), n_pts * 48, 0)
    if sd.length_unionized_energy_array > 0:
        out["unionized_energy_array"] = fnv(C.cast(sd.unionized_energy_array, C.c_void_p),
                                            sd.length_unionized_energy_array * 8, 0)
    if sd.length_index_grid > 0:
        out["index_grid"] = fnv(C.cast(sd.index_grid, C.c_void_p), sd.length_index_grid * 4, 0)
    out["num_nucs"] = fnv(C.cast(sd.num_nucs, C.c_void_p), 12 * 4, 0)
    h_m = 0
    h_c = 0
    for m in range(12):
        n = sd.num_nucs[m]
        base = m * sd.max_num_nucs
        h_m = fnv(C.cast(C.byref(sd.mats.contents, base * 4), C.c_void_p), n * 4, h_m)
        h_c = fnv(C.cast(C.byref(sd.concs.contents, base * 8), C.c_void_p), n * 8, h_c)
    out["mats"] = h_m
    out["concs"] = h_c
    out["max_num_nucs"] = sd.max_num_nucs
    return {k: int(v) for k, v in out.items()}


class Reference:
    """The reference's own openmp-threading code (oracle/_ref/libxsbench_ref.so)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref` where the reference is mounted")
        L = self.lib = C.CDLL(path)
        u64p, i32p, f64p = _P(C.c_uint64), _P(C.c_int), _P(C.c_double)
        self.grid_init_do_not_profile = _bind(L, "grid_init_do_not_profile", SimulationData, [Inputs, C.c_int])
        self.run_event_based_simulation = _bind(L, "run_event_based_simulation", C.c_ulonglong,
                                                [Inputs, SimulationData, C.c_int])
        self.run_history_based_simulation = _bind(L, "run_history_based_simulation", C.c_ulonglong,
                                                  [Inputs, SimulationData, C.c_int])
        self.run_event_based_simulation_optimization_1 = _bind(
            L, "run_event_based_simulation_optimization_1", C.c_ulonglong, [Inputs, SimulationData, C.c_int])
        self.calculate_macro_xs = _bind(L, "calculate_macro_xs", None, [
            C.c_double, C.c_int, C.c_long, C.c_long, i32p, f64p, f64p, i32p, _P(NuclideGridPoint), i32p,
            f64p, C.c_int, C.c_int, C.c_int])
        self.calculate_micro_xs = _bind(L, "calculate_micro_xs", None, [
            C.c_double, C.c_int, C.c_long, C.c_long, f64p, i32p, _P(NuclideGridPoint), C.c_long, f64p,
            C.c_int, C.c_int])
        self.grid_search = _bind(L, "grid_search", C.c_long, [C.c_long, C.c_double, f64p])
        self.grid_search_nuclide = _bind(L, "grid_search_nuclide", C.c_long,
                                         [C.c_long, C.c_double, _P(NuclideGridPoint), C.c_long, C.c_long])
        self.pick_mat = _bind(L, "pick_mat", C.c_int, [u64p])
        self.LCG_random_double = _bind(L, "LCG_random_double", C.c_double, [u64p])
        self.fast_forward_LCG = _bind(L, "fast_forward_LCG", C.c_uint64, [C.c_uint64, C.c_uint64])


def build(reference=True):
    """Compile the oracle (and the reference where its sources are mounted)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if reference and os.path.isdir(REF_SRC):
        subprocess.check_call(["make", "-s", "-C", HERE, "-f", "ref.mk"])
