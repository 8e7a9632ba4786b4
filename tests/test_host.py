"""Host surface of the product (libxsbench.so, bin/XSBench) on the CPU:
table construction against the reference's arrays, the CLI, the data file,
and the C-ABI library's exports and argument checks (no device work)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import xsbench_amd as X
from conftest import DATA_CFGS, ROOT, golden_npz


@pytest.mark.parametrize("name", list(DATA_CFGS))
def test_product_grid_init_is_byte_identical_to_reference(orc, golden_digests, name):
    size, grid, gp = DATA_CFGS[name]
    inp = X.make_inputs(size, grid, "event", gridpoints=gp)
    sd = X.grid_init(inp)
    d = O.sd_digests(orc.fnv1a, inp, sd)
    g = {k: v for k, v in golden_digests[name].items() if k in d}
    assert d == g
    if grid == "unionized":
        # no duplicate energies at these sizes: the parallel sweep ran
        assert X.host_lib().xsbench_last_init_parallel() == 1


def test_product_index_grid_serial_path_on_duplicate_energies(orc):
    """With duplicate unionized energies the reference's serial sweep must
    run; force duplicates by giving two nuclides one shared energy grid."""
    lib = X.host_lib()
    # build tiny tables by hand through the C functions the product exports
    inp = X.make_inputs("small", "unionized", "event", gridpoints=8)
    sd = X.grid_init(inp)
    assert lib.xsbench_last_init_parallel() == 1  # baseline: no duplicates
    # the oracle's serial sweep and the product agree on the tiny grid too
    o_inp = O.make_inputs("small", "unionized", "event", gridpoints=8)
    osd = orc.grid_init(o_inp)
    try:
        a = X.sd_arrays(inp, sd)["index_grid"]
        b = np.ctypeslib.as_array(osd.index_grid, shape=(osd.length_index_grid,))
        assert np.array_equal(a, b)
    finally:
        orc.free_sd(C.byref(osd))


def test_host_lcg_and_fast_forward_match_reference_vectors():
    lib = X.host_lib()
    g = golden_npz("sampling.npz")
    for s, val, nxt in zip(g["seeds"][:256], g["lcg_val"][:256], g["lcg_next"][:256]):
        x = C.c_uint64(int(s))
        assert lib.LCG_random_double(C.byref(x)) == val and x.value == int(nxt)
    for s, n, out in zip(g["seeds"][:512], g["ff_n"][:512], g["ff_out"][:512]):
        assert lib.fast_forward_LCG(int(s), int(n)) == int(out)


def test_batch_generator_matches_oracle(orc):
    """the product's batched-workload generator and per-batch sampling stream
    equal the oracle's definitions"""
    sizes, seeds, thr = X.make_batches(2024, 512, 1000, 2_000_000)
    s2, se2, t2 = orc.batch_plan(2024, 512, 1000, 2_000_000)
    assert np.array_equal(sizes, s2) and np.array_equal(seeds, se2) and np.array_equal(thr, t2)
    assert sizes.min() >= 1000 and sizes.max() <= 2_000_000
    assert np.all(np.diff(thr[:, :11], axis=1) > 0) and np.all(thr[:, 11] == 2.0)
    # log-uniform: the median size is near sqrt(1e3 * 2e6)
    assert 2.5e4 < np.median(sizes) < 8e4
    lib = X.host_lib()
    e, m = C.c_double(), C.c_int()
    rng = np.random.default_rng(3)
    for b in rng.integers(0, 512, 20):
        t = np.ascontiguousarray(thr[b])
        for i in (0, 1, int(sizes[b]) - 1, int(rng.integers(0, sizes[b]))):
            lib.xsbench_batch_sample(int(seeds[b]), t.ctypes.data, i, C.byref(e), C.byref(m))
            assert (e.value, m.value) == orc.batch_sample(int(seeds[b]), t, i)


def test_materials_tables():
    inp = X.make_inputs("large", "hash", "event", gridpoints=4)
    sd = X.grid_init(inp)
    a = X.sd_arrays(inp, sd)
    assert list(a["num_nucs"]) == [321, 5, 4, 4, 27, 21, 21, 21, 21, 21, 9, 9]
    assert sd.max_num_nucs == 321
    assert list(a["mats"][0, 34:321]) == list(range(68, 355))
    assert list(a["mats"][2, :4]) == [24, 41, 4, 5]


# ---------------------------------------------------------------------------
# executable: CLI errors (openmp-threading/io.c:208-225) need no GPU
# ---------------------------------------------------------------------------
def _run_exe(args, cwd=None, timeout=120):
    return subprocess.run([X.EXE] + args, capture_output=True, text=True, cwd=cwd, timeout=timeout)


@pytest.mark.parametrize("args", [["-m", "sideways"], ["-G", "cubic"], ["-s", "huge"], ["-l"],
                                  ["-q"], ["-l", "0"], ["-P", "fp8"]])
def test_cli_errors_exit_4(args):
    r = _run_exe(args)
    assert r.returncode == 4 and "Usage" in r.stdout


def test_cli_unknown_kernel_exits_1():
    r = _run_exe(["-k", "9", "-s", "small", "-m", "event", "-l", "10"])
    assert r.returncode == 1 and "No kernel ID 9" in r.stdout


# ---------------------------------------------------------------------------
# XS_data.dat: versioned write/read round trip, reference raw layout, checks
# ---------------------------------------------------------------------------
def _host_io():
    lib = X.host_lib()
    lib.binary_write.restype = None
    lib.binary_write.argtypes = [X.Inputs, X.SimulationData]
    lib.binary_read.restype = X.SimulationData
    lib.binary_read.argtypes = [X.Inputs]
    return lib


def _child(code, cwd):
    """run host IO in a child process (binary_read exits on bad files)"""
    return subprocess.run(["python3", "-c", code], cwd=cwd, capture_output=True, text=True,
                          timeout=300, env=dict(os.environ, PYTHONPATH=ROOT + ":" + os.path.join(ROOT, "tests")))


IO_SNIPPET = """
import ctypes as C, sys, oracle as O, xsbench_amd as X
from test_host import _host_io
lib = _host_io()
inp = X.make_inputs('small', '{grid}', 'event', gridpoints={gp})
{body}
"""


@pytest.mark.parametrize("grid", ["unionized", "hash", "nuclide"])
def test_binary_file_round_trip(tmp_path, grid):
    body = """
sd = X.grid_init(inp)
lib.binary_write(inp, sd)
sd2 = lib.binary_read(inp)
orc = O.Oracle()
assert O.sd_digests(orc.fnv1a, inp, sd) == O.sd_digests(orc.fnv1a, inp, sd2)
print('ok')
"""
    r = _child(IO_SNIPPET.format(grid=grid, gp=50, body=body), str(tmp_path))
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def test_binary_file_reads_reference_raw_layout(tmp_path):
    """openmp-threading/io.c:456-474 layout: raw struct, then six arrays."""
    body = """
sd = X.grid_init(inp)
a = X.sd_arrays(inp, sd)
with open('XS_data.dat', 'wb') as f:
    f.write(bytes(sd))
    f.write(a['num_nucs'].tobytes()); f.write(a['concs'].tobytes()); f.write(a['mats'].tobytes())
    f.write(a['nuclide_grid'].tobytes()); f.write(a['index_grid'].tobytes())
    f.write(a['unionized_energy_array'].tobytes())
sd2 = lib.binary_read(inp)
orc = O.Oracle()
assert O.sd_digests(orc.fnv1a, inp, sd) == O.sd_digests(orc.fnv1a, inp, sd2)
print('ok')
"""
    r = _child(IO_SNIPPET.format(grid="unionized", gp=50, body=body), str(tmp_path))
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("damage", ["flip", "truncate", "mismatch"])
def test_binary_file_rejects_damaged_or_mismatched(tmp_path, damage):
    body = """
sd = X.grid_init(inp)
lib.binary_write(inp, sd)
import os
if '{damage}' == 'flip':
    with open('XS_data.dat', 'r+b') as f:
        f.seek(4000); b = f.read(1); f.seek(4000); f.write(bytes([b[0] ^ 1]))
elif '{damage}' == 'truncate':
    os.truncate('XS_data.dat', os.path.getsize('XS_data.dat') - 100)
else:
    inp.hash_bins = 77; inp.grid_type = 2
sd2 = lib.binary_read(inp)
print('read')
""".replace("{damage}", damage)
    r = _child(IO_SNIPPET.format(grid="unionized", gp=50, body=body), str(tmp_path))
    assert r.returncode == 1 and "read" not in r.stdout and "not usable" in r.stderr


# ---------------------------------------------------------------------------
# C ABI: exports and argument validation (no device work)
# ---------------------------------------------------------------------------
def _declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


def test_hip_library_exports_every_declared_symbol():
    names = _declared("xsb_hip.h", "xsb_hip_")
    assert set(names) == set(X.HIP_SYMBOLS)
    lib = C.CDLL(X.HIP_SO)
    for n in names:
        assert hasattr(lib, n), n


_ABI_STRUCTS = {"xsb_tables": "XsbTables", "xsb_event_args": "XsbEventArgs",
                "xsb_history_args": "XsbHistoryArgs", "xsb_batch_plan": "XsbBatchPlan",
                "xsb_batched_args": "XsbBatchedArgs", "xsb_result": "XsbResult",
                "xsb_info": "XsbInfo", "xsbench_options": "XsbenchOptions", "xsb_lines": "XsbLines"}
_ABI_CONSTS = {"XSB_XS_FP64": "XS_FP64", "XSB_XS_FP32": "XS_FP32", "XSB_XS_BF16": "XS_BF16",
               "XSB_BUILD_FP32": "BUILD_FP32", "XSB_BUILD_BF16": "BUILD_BF16",
               "XSB_BUILD_GRIDS": "BUILD_GRIDS", "XSB_BUILD_IDX32": "BUILD_IDX32",
               "XSB_UNIONIZED": "UNIONIZED", "XSB_NUCLIDE": "NUCLIDE", "XSB_HASH": "HASH",
               "XSB_K_BASELINE": "K_BASELINE", "XSB_K_MAT_E_SORT": "K_MAT_E_SORT",
               "XSB_K_AUTO": "K_AUTO", "XSBENCH_SCHEDULE_ASKED": "SCHEDULE_ASKED",
               "XSBENCH_SCHEDULE_AUTO": "SCHEDULE_AUTO",
               "XSB_NOALLOC_MAX_LOOKUPS": "NOALLOC_MAX_LOOKUPS", "XSB_LINE_REGIONS": "N_LINE_REGIONS"}


def test_abi_structs_and_constants_match_the_python_mirror(tmp_path):
    """sizes and field offsets of every xsb_hip.h struct, and its constants,
    as the C compiler sees them, against the ctypes mirror"""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "xsb_hip.h"',
             '#include "xsbench.h"', "int main(void) {"]
    for cs, ps in _ABI_STRUCTS.items():
        lines.append(f'printf("{cs} size %zu\\n", sizeof({cs}));')
        for f, _ in getattr(X, ps)._fields_:
            lines.append(f'printf("{cs} {f} %zu\\n", offsetof({cs}, {f}));')
    for c in _ABI_CONSTS:
        lines.append(f'printf("const {c} %lld\\n", (long long)({c}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        a, b, v = line.split()
        if a == "const":
            assert int(v) == getattr(X, _ABI_CONSTS[b]), line
        elif b == "size":
            assert int(v) == C.sizeof(getattr(X, _ABI_STRUCTS[a])), line
        else:
            assert int(v) == getattr(getattr(X, _ABI_STRUCTS[a]), b).offset, line
        seen += 1
    assert seen == len(_ABI_CONSTS) + sum(1 + len(getattr(X, p)._fields_)
                                          for p in _ABI_STRUCTS.values())


def test_host_library_exports_every_declared_function():
    text = open(os.path.join(ROOT, "include", "xsbench.h")).read()
    names = set(re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, re.M))
    lib = C.CDLL(X.HOST_SO)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    for n in X.HOST_SYMBOLS:
        assert hasattr(lib, n)


def test_upload_rejects_bad_tables_before_touching_a_device():
    lib = X.hip_lib()
    h = C.c_void_p()
    t = X.XsbTables()
    assert lib.xsb_hip_upload(0, None, 0, C.byref(h)) == 1
    assert lib.xsb_hip_upload(0, C.byref(t), 0, C.byref(h)) == 1  # n_isotopes = 0
    inp = X.make_inputs("small", "unionized", "event", gridpoints=20)
    sd = X.grid_init(inp)
    t = X.tables_from_sd(inp, sd)
    t.length_index_grid -= 1
    assert lib.xsb_hip_upload(0, C.byref(t), 0, C.byref(h)) == 1
    assert b"index grid length" in lib.xsb_hip_last_error()
    t = X.tables_from_sd(inp, sd)
    t.grid_type = 7
    assert lib.xsb_hip_upload(0, C.byref(t), 0, C.byref(h)) == 1
    assert not h.value
    # NULL handles are refused, never dereferenced
    r = X.XsbResult()
    a = X.XsbEventArgs()
    assert lib.xsb_hip_event(None, C.byref(a), C.byref(r)) == 1
    assert lib.xsb_hip_release(None) == 0
    assert lib.xsb_hip_strerror(5) == b"unsupported option"


def test_no_device_is_reported_not_crashed():
    """On a machine without a GPU a valid upload fails with XSB_ERR_NO_DEVICE."""
    code = """
import ctypes as C, xsbench_amd as X
inp = X.make_inputs('small', 'hash', 'event', gridpoints=20)
sd = X.grid_init(inp)
t = X.tables_from_sd(inp, sd)
h = C.c_void_p()
print('rc', X.hip_lib().xsb_hip_upload(0, C.byref(t), 0, C.byref(h)))
"""
    env = dict(os.environ, PYTHONPATH=ROOT, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and "rc 2" in r.stdout, r.stdout + r.stderr


# ---------------------------------------------------------------------------
# the restart rule of the device's chunked index-grid construction
# ---------------------------------------------------------------------------
def _serial_sweep(E, U):
    """openmp-threading/GridInit.c:102-119, one nuclide row of E per column"""
    n_iso, n_gp = E.shape
    idx = np.zeros(n_iso, np.int64)
    high = E[:, 1].copy()
    out = np.zeros((len(U), n_iso), np.int64)
    for e, u in enumerate(U):
        for i in range(n_iso):
            if not (u < high[i]) and idx[i] != n_gp - 2:
                idx[i] += 1
                high[i] = E[i, idx[i] + 1]
            out[e, i] = idx[i]
    return out


def _chunked_sweep(E, U, R):
    """the device's split: each chunk of R rows restarts from the state at the
    end of the run of equal energies before it, min(#{k>=1: E_k <= x}, n_gp-2)"""
    n_iso, n_gp = E.shape
    out = np.full((len(U), n_iso), -1, np.int64)
    for e0 in range(0, len(U), R):
        e1 = min(e0 + R, len(U))
        for i in range(n_iso):
            r = e0
            while r > 0 and U[r - 1] == U[r]:
                r -= 1
            idx = 0 if r == 0 else min(int(np.searchsorted(E[i, 1:], U[r - 1], side="right")),
                                       n_gp - 2)
            high = E[i, idx + 1]
            for e in range(r, e1):
                if not (U[e] < high) and idx != n_gp - 2:
                    idx += 1
                    high = E[i, idx + 1]
                if e >= e0:
                    out[e, i] = idx
    return out


def test_chunked_index_sweep_equals_serial_sweep():
    """duplicates within and across nuclides, runs straddling chunk edges,
    two-point nuclides"""
    rng = np.random.default_rng(5)
    for _ in range(200):
        n_iso, n_gp = int(rng.integers(1, 5)), int(rng.integers(2, 9))
        E = np.sort(rng.integers(0, int(rng.integers(2, 12)), size=(n_iso, n_gp)) / 16.0, axis=1)
        U = np.sort(E.ravel())
        R = int(rng.integers(1, 6))
        assert np.array_equal(_chunked_sweep(E, U, R), _serial_sweep(E, U)), (E, R)


# ---------------------------------------------------------------------------
# the reference's own program against this library
# ---------------------------------------------------------------------------
REF_OMP = "/root/reference/openmp-threading"
LIBDIR = os.path.join(ROOT, "xsbench_amd", "lib")


def _ref_prototypes():
    """names of the functions openmp-threading/XSbench_header.h:104-150 declares"""
    text = open(os.path.join(REF_OMP, "XSbench_header.h")).read()
    body = text[text.index("// io.c"):]
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", body, re.M)))


@pytest.mark.skipif(not os.path.isdir(REF_OMP), reason="reference sources not mounted")
def test_every_reference_header_function_links_from_the_library(tmp_path):
    """a translation unit compiled against the reference's own header takes
    the address of every function it declares; it links against
    libxsbench.so alone (no reference object), so each one is defined there
    with C linkage"""
    names = _ref_prototypes()
    for n in ("calculate_micro_xs", "calculate_macro_xs", "grid_search", "grid_search_nuclide",
              "pick_mat", "run_event_based_simulation", "grid_init_do_not_profile"):
        assert n in names
    src = tmp_path / "all.c"
    src.write_text('#include "XSbench_header.h"\n'
                   "void *all_functions[] = {" + ", ".join(f"(void *){n}" for n in names) + "};\n"
                   "int main(void) { return all_functions[0] == 0; }\n")
    exe = tmp_path / "all"
    subprocess.run(["gcc", "-std=gnu99", "-I", REF_OMP, str(src), "-o", str(exe), "-L", LIBDIR,
                    "-lxsbench", "-lxsb_hip", f"-Wl,-rpath,{LIBDIR}", "-lm"], check=True)
    assert subprocess.run([str(exe)], timeout=60).returncode == 0


@pytest.mark.skipif(not os.path.isdir(REF_OMP), reason="reference sources not mounted")
def test_reference_main_links_and_reaches_the_device(tmp_path):
    """openmp-threading/Main.c, unchanged, built against libxsbench.so: it
    runs the reference's flow (CLI, grid init, the event driver) and, with no
    GPU, stops cleanly at the upload with the library's message"""
    exe = tmp_path / "XSBench_ref_main"
    subprocess.run(["gcc", "-std=gnu99", "-O2", "-fopenmp", "-DOPENMP", "-I", REF_OMP,
                    os.path.join(REF_OMP, "Main.c"), "-o", str(exe), "-L", LIBDIR, "-lxsbench",
                    "-lxsb_hip", f"-Wl,-rpath,{LIBDIR}", "-lm"], check=True)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    for k in ("0", "1"):
        r = subprocess.run([str(exe), "-s", "small", "-m", "event", "-l", "1000", "-k", k,
                            "-t", "2"], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 1, r.stdout + r.stderr
        assert "Intializing material data" in r.stdout or "Initializing" in r.stdout, r.stdout
        assert "no usable HIP device" in r.stderr, r.stderr


def test_host_search_and_pick_mat_helpers(orc):
    """grid_search / grid_search_nuclide / pick_mat exported with the
    reference's signatures: pick_mat against the reference's own outputs,
    the searches against a brute-force clamped upper bound (ties, keys out
    of range) and against the oracle"""
    import bisect
    lib = X.host_lib()
    g = golden_npz("sampling.npz")
    for s, m in zip(g["seeds"], g["pick_mat"]):
        x = C.c_uint64(int(s))
        y = C.c_uint64(int(s))
        assert lib.pick_mat(C.byref(x)) == m
        orc.pick_mat(C.byref(y))
        assert x.value == y.value
    rng = np.random.default_rng(11)
    for n in (2, 3, 5, 16, 40):
        for _ in range(25):
            A = np.sort(rng.integers(0, 7, size=n).astype(np.float64) / 4.0)
            pts = (O.NuclideGridPoint * n)()
            for k in range(n):
                pts[k].energy = A[k]
            lo = int(rng.integers(0, n - 1))
            hi = int(rng.integers(lo + 1, n))
            for q in np.concatenate([A, [-1.0, 9.0, 0.3, 0.8]]):
                want = min(max(bisect.bisect_right(list(A), q) - 1, 0), n - 2)
                assert lib.grid_search(n, q, A.ctypes.data) == want
                ub = bisect.bisect_right(list(A[lo:hi + 1]), q) + lo
                want_n = min(max(ub - 1, lo), hi - 1) if hi - lo > 1 else lo
                got = lib.grid_search_nuclide(n, q, C.addressof(pts), lo, hi)
                assert got == want_n == orc.grid_search_nuclide(n, q, pts, lo, hi)


def test_table_sizes_fit_one_gpu():
    """the resident tables of an upload, from xsb_hip_table_bytes (host
    arithmetic, the sizes the upload records): no whole unionized index grid
    is kept, so the XL problem on the unionized grid (355 x 238,847 points,
    int32 per-material index tables) fits one 288 GB MI355X with FP32 and
    BF16 tables too; the large problem's per-material tables are 16-bit"""
    xl = X.table_bytes(X.make_inputs("XL", "unionized", "event"), build_fp32=True, build_bf16=True)
    total = sum(v for k, v in xl.items() if k.startswith("bytes_"))
    assert xl["bytes_index"] == 0
    assert total < 250e9, total  # leaves room for the staging and sort buffers
    n = 355 * 238847
    # int32 entries (n_gp > 65536): rows of 324, 8, 4, 4, 28, 24 x 5, 12, 12 entries
    assert xl["bytes_midx"] == n * 4 * (324 + 8 + 4 + 4 + 28 + 24 * 5 + 12 + 12)
    lg = X.table_bytes(X.make_inputs("large", "unionized", "event"))
    # uint16 entries: rows of 328, 8, 8, 8, 32, 24 x 5, 16, 16 entries
    assert lg["bytes_index"] == 0 and lg["bytes_midx"] == 355 * 11303 * 2 * (328 + 8 * 3 + 32 + 24 * 5 + 16 * 2)
    hs = X.table_bytes(X.make_inputs("large", "hash", "event"))
    # the int32 index grid and the 96-byte window records, rows of 356 entries;
    # the energy rows, 11303 + 16 rounded up to 11320 doubles
    assert hs["bytes_index"] == 10000 * 356 * (4 + 96) + 355 * 11320 * 8 and hs["bytes_midx"] == 0
