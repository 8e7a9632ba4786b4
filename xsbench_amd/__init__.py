"""MI355X (gfx950) implementation of XSBench's macroscopic cross-section lookup.

Python mirror of the two C interfaces of this package:

* ``lib/libxsbench.so`` -- the reference's host surface (include/xsbench.h):
  ``Inputs`` / ``SimulationData``, ``grid_init_do_not_profile``,
  ``run_event_based_simulation`` / ``run_history_based_simulation`` ...
* ``lib/libxsb_hip.so`` -- the C ABI of the HIP kernels (include/xsb_hip.h).

Both are loaded with ctypes.  There is no Python or CPU implementation of the
lookup path here: when the HIP library is missing, importing the native layer
raises ``NativeLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG, "lib")
HIP_SO = os.path.join(LIB_DIR, "libxsb_hip.so")
HOST_SO = os.path.join(LIB_DIR, "libxsbench.so")
EXE = os.path.join(PKG, "bin", "XSBench")

UNIONIZED, NUCLIDE, HASH = 0, 1, 2
HISTORY_BASED, EVENT_BASED = 1, 2
STARTING_SEED = 1070
GRID_TYPES = {"unionized": UNIONIZED, "nuclide": NUCLIDE, "hash": HASH}

XS_FP64, XS_FP32, XS_BF16 = 0, 1, 2
PRECISIONS = {"fp64": XS_FP64, "fp32": XS_FP32, "bf16": XS_BF16}
BUILD_FP32, BUILD_BF16, BUILD_GRIDS, BUILD_IDX32 = 0x1, 0x2, 0x4, 0x8

# lookup schedules = the reference's -k kernel ids (cuda/Main.cu:62-83)
K_BASELINE, K_SPLIT, K_MAT_LAUNCH, K_FUEL_LAUNCH, K_MAT_SORT, K_FUEL_SORT, K_MAT_E_SORT = range(7)
K_AUTO = -1
NOALLOC_MAX_LOOKUPS = 1 << 21  # XSB_NOALLOC_MAX_LOOKUPS
# xsbench_options.schedule (include/xsbench.h)
SCHEDULE_ASKED, SCHEDULE_AUTO = 0, 1

ERRORS = {0: "success", 1: "invalid argument", 2: "no usable HIP device", 3: "device out of memory",
          4: "HIP runtime error", 5: "unsupported option"}

# Published verification hashes, openmp-threading/io.c:85-106
PUBLISHED = {("event", "small"): 945990, ("event", "large"): 952131,
             ("history", "small"): 941535, ("history", "large"): 954318}


class NativeLibraryMissing(RuntimeError):
    pass


class XsbError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {message}")
        self.code = code


# ---------------------------------------------------------------------------
# struct layouts
# ---------------------------------------------------------------------------
class NuclideGridPoint(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "energy", "total_xs", "elastic_xs", "absorbtion_xs", "fission_xs", "nu_fission_xs")]


class Inputs(C.Structure):
    _fields_ = [
        ("nthreads", C.c_int), ("n_isotopes", C.c_long), ("n_gridpoints", C.c_long),
        ("lookups", C.c_int), ("HM", C.c_char_p), ("grid_type", C.c_int),
        ("hash_bins", C.c_int), ("particles", C.c_int), ("simulation_method", C.c_int),
        ("binary_mode", C.c_int), ("kernel_id", C.c_int),
    ]


class SimulationData(C.Structure):
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


class XsbTables(C.Structure):
    _fields_ = [
        ("n_isotopes", C.c_int64), ("n_gridpoints", C.c_int64), ("grid_type", C.c_int32),
        ("hash_bins", C.c_int32), ("num_nucs", C.c_void_p), ("mats", C.c_void_p),
        ("concs", C.c_void_p), ("max_num_nucs", C.c_int32), ("reserved0", C.c_int32),
        ("unionized_energy_array", C.c_void_p), ("length_unionized_energy_array", C.c_int64),
        ("index_grid", C.c_void_p), ("length_index_grid", C.c_int64),
        ("nuclide_grid", C.c_void_p), ("length_nuclide_grid", C.c_int64),
    ]


class XsbEventArgs(C.Structure):
    _fields_ = [
        ("lookups", C.c_int64), ("first_lookup", C.c_int64), ("seed", C.c_uint64),
        ("kernel_id", C.c_int32), ("xs_precision", C.c_int32),
        ("out_energy", C.c_void_p), ("out_mat", C.c_void_p), ("out_index", C.c_void_p),
        ("out_macro", C.c_void_p), ("out_argmax", C.c_void_p), ("out_bound", C.c_void_p),
        ("out_fallback", C.c_void_p),
    ]


class XsbHistoryArgs(C.Structure):
    _fields_ = [
        ("particles", C.c_int64), ("lookups_per_particle", C.c_int32), ("kernel_id", C.c_int32),
        ("xs_precision", C.c_int32), ("reserved0", C.c_int32), ("seed", C.c_uint64),
        ("out_particle_sum", C.c_void_p), ("out_final_seed", C.c_void_p),
    ]


class XsbBatchPlan(C.Structure):
    _fields_ = [
        ("tile", C.c_int32), ("stage", C.c_int32), ("sort", C.c_int32), ("windows", C.c_int32),
        ("tuned", C.c_int32), ("graph", C.c_int32), ("mapping", C.c_int32),
        ("reserved0", C.c_int32), ("ms_tuning", C.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class XsbBatchedArgs(C.Structure):
    _fields_ = [
        ("n_batches", C.c_int32), ("xs_precision", C.c_int32), ("sizes", C.c_void_p),
        ("seeds", C.c_void_p), ("thresholds", C.c_void_p), ("mode", C.c_int32),
        ("tile", C.c_int32), ("stage", C.c_int32), ("sort", C.c_int32), ("window", C.c_int64),
        ("eager", C.c_int32), ("mapping", C.c_int32),
        ("out_vsum", C.c_void_p), ("out_done_ms", C.c_void_p), ("out_plan", C.c_void_p),
    ]


class XsbResult(C.Structure):
    _fields_ = [
        ("vsum", C.c_uint64), ("ms_total", C.c_double), ("ms_lookup", C.c_double),
        ("ms_sample", C.c_double), ("n_fallback", C.c_int64),
        ("material_counts", C.c_int64 * 12), ("kernel_id", C.c_int32), ("reserved0", C.c_int32),
    ]

    def as_dict(self):
        return {"vsum": int(self.vsum), "hash": int(self.vsum) % 999983, "ms_total": self.ms_total,
                "ms_lookup": self.ms_lookup, "ms_sample": self.ms_sample,
                "n_fallback": int(self.n_fallback), "kernel_id": int(self.kernel_id),
                "material_counts": [int(x) for x in self.material_counts]}


# tables of xsb_lines.region_bytes, in order (XSB_LINE_REGIONS)
LINE_REGIONS = ("ueg", "acc", "midx", "ivl", "ivl32", "index", "grid", "xs32", "xs16", "hwin",
                "esoa")
N_LINE_REGIONS = len(LINE_REGIONS)


class XsbLines(C.Structure):
    _fields_ = [("bytes", C.c_int64), ("bytes_tables", C.c_int64), ("bytes_schedule", C.c_int64),
                ("bytes_small", C.c_int64), ("launches", C.c_int64),
                ("region_bytes", C.c_int64 * len(LINE_REGIONS))]

    def as_dict(self):
        d = {f: int(getattr(self, f)) for f, _ in self._fields_[:5]}
        d["region_bytes"] = [int(x) for x in self.region_bytes]
        return d


class XsbenchOptions(C.Structure):
    """xsbench_options (include/xsbench.h): options of the host surface"""
    _fields_ = [
        ("xs_precision", C.c_int), ("device", C.c_int), ("verbose", C.c_int),
        ("last_kernel_ms", C.c_double), ("last_lookup_ms", C.c_double),
        ("last_fallbacks", C.c_long), ("last_error", C.c_int), ("schedule", C.c_int),
        ("last_kernel_id", C.c_int),
    ]


class XsbInfo(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("grid_type", C.c_int32), ("n_isotopes", C.c_int64),
        ("n_gridpoints", C.c_int64), ("n_union", C.c_int64), ("bytes_fp64", C.c_int64),
        ("bytes_fp32", C.c_int64), ("bytes_bf16", C.c_int64), ("bytes_index", C.c_int64),
        ("bytes_midx", C.c_int64), ("bytes_union", C.c_int64), ("bytes_ivl", C.c_int64),
        ("cu_count", C.c_int32),
        ("xcd_count", C.c_int32),
        ("arch", C.c_char * 32),
    ]


# C-ABI entry points declared in include/xsb_hip.h (checked by the tests)
HIP_SYMBOLS = ("xsb_hip_upload", "xsb_hip_release", "xsb_hip_event", "xsb_hip_history",
               "xsb_hip_lookup", "xsb_hip_batched", "xsb_hip_balg", "xsb_hip_copy_probe",
               "xsb_hip_info", "xsb_hip_strerror", "xsb_hip_last_error", "xsb_hip_download",
               "xsb_hip_macro_host", "xsb_hip_micro_host", "xsb_hip_event_noalloc",
               "xsb_hip_lines_event", "xsb_hip_lines_history", "xsb_hip_lines_batched",
               "xsb_hip_table_bytes")
HOST_SYMBOLS = ("read_CLI", "print_inputs", "print_results", "grid_init_do_not_profile",
                "load_num_nucs", "load_mats", "load_concs", "run_event_based_simulation",
                "run_history_based_simulation", "run_event_based_simulation_optimization_1",
                "move_simulation_data_to_device", "release_device_memory", "LCG_random_double",
                "fast_forward_LCG", "binary_write", "binary_read", "estimate_mem_usage",
                "xsbench_get_options", "xsbench_last_init_parallel", "xsbench_main",
                "xsbench_make_batches", "xsbench_batch_sample", "run_batched_simulation",
                "calculate_micro_xs", "calculate_macro_xs", "grid_search", "grid_search_nuclide",
                "pick_mat")


# ---------------------------------------------------------------------------
# loading
# ---------------------------------------------------------------------------
_hip = None
_host = None


def build(jobs=8):
    """Compile libxsb_hip.so (hipcc, gfx950), libxsbench.so and bin/XSBench."""
    subprocess.check_call(["make", "-s", "-C", PKG, f"-j{jobs}"])


def _bind(lib, name, restype, argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def hip_lib():
    """The HIP C-ABI library; raises NativeLibraryMissing when it is not built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_SO):
            raise NativeLibraryMissing(f"{HIP_SO} is missing: run xsbench_amd.build()")
        L = C.CDLL(HIP_SO, mode=C.RTLD_GLOBAL)
        vp = C.c_void_p
        _bind(L, "xsb_hip_upload", C.c_int, [C.c_int, C.POINTER(XsbTables), C.c_uint32, C.POINTER(vp)])
        _bind(L, "xsb_hip_release", C.c_int, [vp])
        _bind(L, "xsb_hip_event", C.c_int, [vp, C.POINTER(XsbEventArgs), C.POINTER(XsbResult)])
        _bind(L, "xsb_hip_history", C.c_int, [vp, C.POINTER(XsbHistoryArgs), C.POINTER(XsbResult)])
        _bind(L, "xsb_hip_event_noalloc", C.c_int,
              [vp, C.POINTER(XsbEventArgs), C.POINTER(XsbResult)])
        _bind(L, "xsb_hip_lines_event", C.c_int, [vp, C.POINTER(XsbEventArgs), C.POINTER(XsbLines)])
        _bind(L, "xsb_hip_lines_history", C.c_int,
              [vp, C.POINTER(XsbHistoryArgs), C.POINTER(XsbLines)])
        _bind(L, "xsb_hip_lines_batched", C.c_int,
              [vp, C.POINTER(XsbBatchedArgs), C.POINTER(XsbLines)])
        _bind(L, "xsb_hip_lookup", C.c_int, [vp, C.c_int64, vp, vp, C.c_int32, vp, vp, vp, vp])
        _bind(L, "xsb_hip_batched", C.c_int, [vp, C.POINTER(XsbBatchedArgs), C.POINTER(XsbResult)])
        _bind(L, "xsb_hip_balg", C.c_int, [vp, C.POINTER(XsbEventArgs), C.POINTER(C.c_double)])
        _bind(L, "xsb_hip_copy_probe", C.c_int, [vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)])
        _bind(L, "xsb_hip_info", C.c_int, [vp, C.POINTER(XsbInfo)])
        _bind(L, "xsb_hip_table_bytes", C.c_int, [C.POINTER(XsbTables), C.c_uint32,
                                                 C.POINTER(XsbInfo)])
        _bind(L, "xsb_hip_download", C.c_int, [vp, vp, vp, vp])
        _bind(L, "xsb_hip_macro_host", C.c_int, [vp, C.c_int64, vp, vp, C.c_int32, vp])
        _bind(L, "xsb_hip_micro_host", C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp])
        _bind(L, "xsb_hip_strerror", C.c_char_p, [C.c_int])
        _bind(L, "xsb_hip_last_error", C.c_char_p, [])
        _hip = L
    return _hip


def host_lib():
    """The host C surface (libxsbench.so, links libxsb_hip.so)."""
    global _host
    if _host is None:
        hip_lib()
        if not os.path.exists(HOST_SO):
            raise NativeLibraryMissing(f"{HOST_SO} is missing: run xsbench_amd.build()")
        L = C.CDLL(HOST_SO)
        _bind(L, "grid_init_do_not_profile", SimulationData, [Inputs, C.c_int])
        _bind(L, "run_event_based_simulation", C.c_ulonglong, [Inputs, SimulationData, C.c_int])
        _bind(L, "run_history_based_simulation", C.c_ulonglong, [Inputs, SimulationData, C.c_int])
        _bind(L, "run_event_based_simulation_optimization_1", C.c_ulonglong,
              [Inputs, SimulationData, C.c_int])
        _bind(L, "move_simulation_data_to_device", SimulationData, [Inputs, C.c_int, SimulationData])
        _bind(L, "release_device_memory", None, [SimulationData])
        _bind(L, "LCG_random_double", C.c_double, [C.POINTER(C.c_uint64)])
        _bind(L, "fast_forward_LCG", C.c_uint64, [C.c_uint64, C.c_uint64])
        _bind(L, "xsbench_last_init_parallel", C.c_int, [])
        _bind(L, "xsbench_get_options", C.POINTER(XsbenchOptions), [])
        _bind(L, "estimate_mem_usage", C.c_size_t, [Inputs])
        _bind(L, "xsbench_main", C.c_int, [C.c_int, C.POINTER(C.c_char_p)])
        _bind(L, "grid_search", C.c_long, [C.c_long, C.c_double, C.c_void_p])
        _bind(L, "grid_search_nuclide", C.c_long, [C.c_long, C.c_double, C.c_void_p, C.c_long,
                                                   C.c_long])
        _bind(L, "pick_mat", C.c_int, [C.POINTER(C.c_uint64)])
        _bind(L, "calculate_micro_xs", None,
              [C.c_double, C.c_int, C.c_long, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
               C.c_long, C.c_void_p, C.c_int, C.c_int])
        _bind(L, "calculate_macro_xs", None,
              [C.c_double, C.c_int, C.c_long, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int])
        _bind(L, "load_num_nucs", C.POINTER(C.c_int), [C.c_long])
        _bind(L, "load_mats", C.POINTER(C.c_int), [C.POINTER(C.c_int), C.c_long, C.POINTER(C.c_int)])
        _bind(L, "load_concs", C.POINTER(C.c_double), [C.POINTER(C.c_int), C.c_int])
        _bind(L, "xsbench_make_batches", None, [C.c_uint64, C.c_int, C.c_long, C.c_long, C.c_void_p,
                                                C.c_void_p, C.c_void_p])
        _bind(L, "xsbench_batch_sample", None, [C.c_uint64, C.c_void_p, C.c_long,
                                                C.POINTER(C.c_double), C.POINTER(C.c_int)])
        _bind(L, "run_batched_simulation", C.c_ulonglong,
              [Inputs, SimulationData, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
               C.c_void_p])
        _host = L
    return _host


def _check(rc):
    if rc != 0:
        raise XsbError(rc, hip_lib().xsb_hip_last_error().decode(errors="replace"))


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def make_batches(seed, n_batches, lo, hi):
    """The batched workload (xsbench_make_batches): per-batch sizes (log-uniform
    in [lo, hi]), stream seeds and [n_batches, 12] cumulative material thresholds."""
    sizes = np.empty(n_batches, np.int64)
    seeds = np.empty(n_batches, np.uint64)
    thr = np.empty((n_batches, 12), np.float64)
    host_lib().xsbench_make_batches(seed, n_batches, lo, hi, sizes.ctypes.data, seeds.ctypes.data,
                                    thr.ctypes.data)
    return sizes, seeds, thr


# ---------------------------------------------------------------------------
# Inputs (openmp-threading/io.c:227-454 semantics)
# ---------------------------------------------------------------------------
def make_inputs(size="large", grid="unionized", method="event", lookups=None, particles=None,
                gridpoints=None, hash_bins=10000, kernel_id=0, nthreads=None):
    """Inputs as read_CLI builds them for ``-s size -G grid -m method [-l] [-p] [-g]``
    (with -l/-p given before -m, so an explicit event lookup count is kept)."""
    n_iso = 68 if size.lower() == "small" else 355
    if gridpoints is not None:
        n_gp = int(gridpoints)
    elif size.lower() == "xl":
        n_gp = 238847
    elif size.lower() == "xxl":
        n_gp = int(238847 * 2.1)
    else:
        n_gp = 11303
    if method == "event":
        lk = int(lookups) if lookups is not None else 34 * (int(particles) if particles is not None else 500000)
        pt = 0 if (lookups is None and particles is None) else (int(particles) if particles is not None else 500000)
        sm = EVENT_BASED
    else:
        lk = int(lookups) if lookups is not None else 34
        pt = int(particles) if particles is not None else 500000
        sm = HISTORY_BASED
    inp = Inputs()
    inp.nthreads = nthreads or (os.cpu_count() or 1)
    inp.n_isotopes = n_iso
    inp.n_gridpoints = n_gp
    inp.lookups = lk
    inp._hm = size.encode()
    inp.HM = inp._hm
    inp.grid_type = GRID_TYPES[grid]
    inp.hash_bins = hash_bins
    inp.particles = pt
    inp.simulation_method = sm
    inp.binary_mode = 0
    inp.kernel_id = kernel_id
    return inp


def run_main(args):
    """The reference program flow (Main.c) in this process: returns its exit code."""
    argv = [b"XSBench"] + [a.encode() for a in args]
    arr = (C.c_char_p * (len(argv) + 1))(*argv, None)
    return host_lib().xsbench_main(len(argv), arr)


def grid_init(inp, mype=1):
    """Host construction of every table (GridInit.c), through libxsbench.so."""
    return host_lib().grid_init_do_not_profile(inp, mype)


def tables_from_sd(inp, sd):
    t = XsbTables()
    t.n_isotopes = inp.n_isotopes
    t.n_gridpoints = inp.n_gridpoints
    t.grid_type = inp.grid_type
    t.hash_bins = inp.hash_bins
    t.num_nucs = C.cast(sd.num_nucs, C.c_void_p)
    t.mats = C.cast(sd.mats, C.c_void_p)
    t.concs = C.cast(sd.concs, C.c_void_p)
    t.max_num_nucs = sd.max_num_nucs
    if inp.grid_type == UNIONIZED:
        t.unionized_energy_array = C.cast(sd.unionized_energy_array, C.c_void_p)
        t.length_unionized_energy_array = sd.length_unionized_energy_array
    if inp.grid_type != NUCLIDE:
        t.index_grid = C.cast(sd.index_grid, C.c_void_p)
        t.length_index_grid = sd.length_index_grid
    t.nuclide_grid = C.cast(sd.nuclide_grid, C.c_void_p)
    t.length_nuclide_grid = sd.length_nuclide_grid
    return t


def table_bytes(inp, build_fp32=False, build_bf16=False, idx32=False):
    """the resident table sizes an upload of inp's problem would take
    (xsb_hip_table_bytes: host arithmetic, no device)"""
    t = material_tables(inp)
    flags = (BUILD_FP32 if build_fp32 else 0) | (BUILD_BF16 if build_bf16 else 0)
    flags |= BUILD_IDX32 if idx32 else 0
    i = XsbInfo()
    _check(hip_lib().xsb_hip_table_bytes(C.byref(t), flags, C.byref(i)))
    return {f: getattr(i, f) for f, _ in XsbInfo._fields_}


def material_tables(inp):
    """xsb_tables holding only the sizes and the material tables (Materials.c:7-116),
    for a device that generates its grids itself (BUILD_GRIDS)."""
    L = host_lib()
    t = XsbTables()
    t.n_isotopes = inp.n_isotopes
    t.n_gridpoints = inp.n_gridpoints
    t.grid_type = inp.grid_type
    t.hash_bins = inp.hash_bins
    num_nucs = L.load_num_nucs(inp.n_isotopes)
    max_n = C.c_int()
    mats = L.load_mats(num_nucs, inp.n_isotopes, C.byref(max_n))
    concs = L.load_concs(num_nucs, max_n.value)
    t.num_nucs = C.cast(num_nucs, C.c_void_p)
    t.mats = C.cast(mats, C.c_void_p)
    t.concs = C.cast(concs, C.c_void_p)
    t.max_num_nucs = max_n.value
    return t


class Device:
    """A device copy of the tables (one xsb_hip handle, one HIP stream).
    With sd=None the grids are generated on the device (xsb_hip_upload with
    BUILD_GRIDS): only the material tables come from the host."""

    def __init__(self, inp, sd, device=0, build_fp32=False, build_bf16=False, idx32=False):
        self.lib = hip_lib()
        self.inp = inp
        self.tables = tables_from_sd(inp, sd) if sd is not None else material_tables(inp)
        flags = (BUILD_FP32 if build_fp32 else 0) | (BUILD_BF16 if build_bf16 else 0)
        flags |= BUILD_IDX32 if idx32 else 0
        if sd is None:
            flags |= BUILD_GRIDS
        h = C.c_void_p()
        _check(self.lib.xsb_hip_upload(device, C.byref(self.tables), flags, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self.lib.xsb_hip_release(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def download(self):
        """The device tables as host arrays in the reference's layout:
        nuclide grid [n_iso, n_gp, 6], unionized energies, index grid [rows, n_iso]."""
        inp = self.inp
        n = inp.n_isotopes * inp.n_gridpoints
        grid = np.empty((inp.n_isotopes, inp.n_gridpoints, 6), np.float64)
        ueg = np.empty(n if inp.grid_type == UNIONIZED else 0, np.float64)
        rows = n if inp.grid_type == UNIONIZED else (inp.hash_bins if inp.grid_type == HASH else 0)
        index = np.empty((rows, inp.n_isotopes), np.int32)
        _check(self.lib.xsb_hip_download(self.handle, grid.ctypes.data,
                                         ueg.ctypes.data if len(ueg) else None,
                                         index.ctypes.data if rows else None))
        return {"nuclide_grid": grid, "unionized_energy_array": ueg, "index_grid": index}

    def info(self):
        i = XsbInfo()
        _check(self.lib.xsb_hip_info(self.handle, C.byref(i)))
        return {f: getattr(i, f) for f, _ in XsbInfo._fields_}

    def event(self, lookups, first=0, kernel_id=K_BASELINE, precision=XS_FP64, detail=False,
              seed=STARTING_SEED):
        """run_event_based_simulation on the GPU.  With detail=True also returns
        per-lookup energy, material, grid index, macro_xs[5] and argmax."""
        a = XsbEventArgs()
        a.lookups = lookups
        a.first_lookup = first
        a.seed = seed
        a.kernel_id = kernel_id
        a.xs_precision = precision
        out = None
        if detail:
            out = {"energy": np.empty(lookups, np.float64), "mat": np.empty(lookups, np.int32),
                   "index": np.empty(lookups, np.int64), "macro": np.empty((lookups, 5), np.float64),
                   "argmax": np.empty(lookups, np.int32), "bound": np.empty(lookups, np.float64),
                   "fallback": np.empty(lookups, np.int32)}
            a.out_bound = out["bound"].ctypes.data
            a.out_fallback = out["fallback"].ctypes.data
            a.out_energy = out["energy"].ctypes.data
            a.out_mat = out["mat"].ctypes.data
            a.out_index = out["index"].ctypes.data
            a.out_macro = out["macro"].ctypes.data
            a.out_argmax = out["argmax"].ctypes.data
        r = XsbResult()
        _check(self.lib.xsb_hip_event(self.handle, C.byref(a), C.byref(r)))
        return (r.as_dict(), out) if detail else r.as_dict()

    def event_noalloc(self, lookups, first=0, kernel_id=K_BASELINE, precision=XS_FP64,
                      seed=STARTING_SEED):
        """xsb_hip_event_noalloc: the event run on the library's fixed work
        memory (at most NOALLOC_MAX_LOOKUPS lookups, no per-lookup outputs)."""
        a = XsbEventArgs()
        a.lookups = lookups
        a.first_lookup = first
        a.seed = seed
        a.kernel_id = kernel_id
        a.xs_precision = precision
        r = XsbResult()
        _check(self.lib.xsb_hip_event_noalloc(self.handle, C.byref(a), C.byref(r)))
        return r.as_dict()

    def history(self, particles, lookups=34, kernel_id=0, precision=XS_FP64, detail=False,
                seed=STARTING_SEED):
        a = XsbHistoryArgs()
        a.particles = particles
        a.lookups_per_particle = lookups
        a.kernel_id = kernel_id
        a.xs_precision = precision
        a.seed = seed
        out = None
        if detail:
            out = {"sums": np.empty(particles, np.uint64), "final_seed": np.empty(particles, np.uint64)}
            a.out_particle_sum = out["sums"].ctypes.data
            a.out_final_seed = out["final_seed"].ctypes.data
        r = XsbResult()
        _check(self.lib.xsb_hip_history(self.handle, C.byref(a), C.byref(r)))
        return (r.as_dict(), out) if detail else r.as_dict()

    def batched(self, sizes, seeds, thresholds, precision=XS_FP64, mode=0, tile=0, stage=0, sort=0,
                window=0, eager=False, mapping=0, plan=None):
        """xsb_hip_batched; returns (result dict, per-batch vsum, per-batch completion ms).
        Pass a dict as `plan` to receive the plan that ran (tile, stage, sort, ...)."""
        sizes = np.ascontiguousarray(sizes, np.int64)
        seeds = np.ascontiguousarray(seeds, np.uint64)
        thr = np.ascontiguousarray(thresholds, np.float64)
        nb = len(sizes)
        vs = np.empty(nb, np.uint64)
        done = np.empty(nb, np.float64)
        a = XsbBatchedArgs()
        a.n_batches = nb
        a.xs_precision = precision
        a.sizes = sizes.ctypes.data
        a.seeds = seeds.ctypes.data
        a.thresholds = thr.ctypes.data
        a.mode = mode
        a.tile = tile
        a.stage = stage
        a.sort = sort
        a.window = window
        a.eager = 1 if eager else 0
        a.mapping = mapping
        a.out_vsum = vs.ctypes.data
        a.out_done_ms = done.ctypes.data
        pl = XsbBatchPlan()
        a.out_plan = C.addressof(pl)
        r = XsbResult()
        _check(self.lib.xsb_hip_batched(self.handle, C.byref(a), C.byref(r)))
        if plan is not None:
            plan.update(pl.as_dict())
        return r.as_dict(), vs, done

    def balg(self, lookups, first=0, precision=XS_FP64, seed=STARTING_SEED):
        a = XsbEventArgs()
        a.lookups = lookups
        a.first_lookup = first
        a.seed = seed
        a.xs_precision = precision
        b = C.c_double()
        _check(self.lib.xsb_hip_balg(self.handle, C.byref(a), C.byref(b)))
        return b.value

    def lines_event(self, lookups, first=0, kernel_id=K_MAT_E_SORT, precision=XS_FP64,
                    seed=STARTING_SEED):
        """compulsory bytes of an event run (xsb_hip_lines_event)"""
        a = XsbEventArgs()
        a.lookups = lookups
        a.first_lookup = first
        a.seed = seed
        a.kernel_id = kernel_id
        a.xs_precision = precision
        o = XsbLines()
        _check(self.lib.xsb_hip_lines_event(self.handle, C.byref(a), C.byref(o)))
        return o.as_dict()

    def lines_history(self, particles, lookups=34, precision=XS_FP64, seed=STARTING_SEED):
        """compulsory bytes of a history run, over its rounds (xsb_hip_lines_history)"""
        a = XsbHistoryArgs()
        a.particles = particles
        a.lookups_per_particle = lookups
        a.kernel_id = K_MAT_E_SORT
        a.xs_precision = precision
        a.seed = seed
        o = XsbLines()
        _check(self.lib.xsb_hip_lines_history(self.handle, C.byref(a), C.byref(o)))
        return o.as_dict()

    def lines_batched(self, sizes, seeds, thresholds, precision=XS_FP64):
        """compulsory bytes of a batched run (xsb_hip_lines_batched)"""
        sizes = np.ascontiguousarray(sizes, np.int64)
        seeds = np.ascontiguousarray(seeds, np.uint64)
        thr = np.ascontiguousarray(thresholds, np.float64)
        a = XsbBatchedArgs()
        a.n_batches = len(sizes)
        a.xs_precision = precision
        a.sizes = sizes.ctypes.data
        a.seeds = seeds.ctypes.data
        a.thresholds = thr.ctypes.data
        o = XsbLines()
        _check(self.lib.xsb_hip_lines_batched(self.handle, C.byref(a), C.byref(o)))
        return o.as_dict()

    def copy_probe(self, nbytes, reps=5):
        ms = C.c_double()
        _check(self.lib.xsb_hip_copy_probe(self.handle, nbytes, reps, C.byref(ms)))
        return ms.value

    def macro_host(self, energy, mat, precision=XS_FP64):
        """calculate_macro_xs over host samples (xsb_hip_macro_host): [n, 5] sums."""
        e = np.ascontiguousarray(energy, np.float64)
        m = np.ascontiguousarray(mat, np.int32)
        out = np.empty((len(e), 5), np.float64)
        _check(self.lib.xsb_hip_macro_host(self.handle, len(e), e.ctypes.data, m.ctypes.data,
                                           precision, out.ctypes.data))
        return out

    def micro_host(self, energy, nuc, lo=None, hi=None):
        """calculate_micro_xs over host samples (xsb_hip_micro_host): [n, 5] XS."""
        e = np.ascontiguousarray(energy, np.float64)
        k = np.ascontiguousarray(nuc, np.int32)
        lo = None if lo is None else np.ascontiguousarray(lo, np.int32)
        hi = None if hi is None else np.ascontiguousarray(hi, np.int32)
        out = np.empty((len(e), 5), np.float64)
        _check(self.lib.xsb_hip_micro_host(self.handle, len(e), e.ctypes.data, k.ctypes.data,
                                           None if lo is None else lo.ctypes.data,
                                           None if hi is None else hi.ctypes.data,
                                           out.ctypes.data))
        return out

    def lookup(self, n, d_energy, d_mat, precision=XS_FP64, d_macro=None, d_argmax=None,
               d_vsum=None, stream=None):
        """calculate_macro_xs over n samples already in device memory (raw pointers)."""
        _check(self.lib.xsb_hip_lookup(self.handle, n, d_energy, d_mat, precision, d_macro,
                                       d_argmax, d_vsum, stream))


def sd_arrays(inp, sd):
    """numpy views of the host SimulationData arrays (no copies)."""
    n_pts = inp.n_isotopes * inp.n_gridpoints
    out = {
        "nuclide_grid": np.ctypeslib.as_array(C.cast(sd.nuclide_grid, C.POINTER(C.c_double)),
                                              shape=(n_pts, 6)),
        "num_nucs": np.ctypeslib.as_array(sd.num_nucs, shape=(12,)),
        "mats": np.ctypeslib.as_array(sd.mats, shape=(12, sd.max_num_nucs)),
        "concs": np.ctypeslib.as_array(sd.concs, shape=(12, sd.max_num_nucs)),
    }
    if sd.length_unionized_energy_array > 0:
        out["unionized_energy_array"] = np.ctypeslib.as_array(
            sd.unionized_energy_array, shape=(sd.length_unionized_energy_array,))
    if sd.length_index_grid > 0:
        out["index_grid"] = np.ctypeslib.as_array(sd.index_grid, shape=(sd.length_index_grid,))
    return out
