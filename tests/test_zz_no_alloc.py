"""The allocation-free event entry point (xsb_hip_event_noalloc) against
xsb_hip_event through the C ABI: the same results on every schedule and
precision, two calls in a row with the second smaller, two handles sharing
the GPU's fixed work array one after the other, and the limits it enforces."""
import ctypes as C

import numpy as np
import pytest

import xsbench_amd as X

pytestmark = pytest.mark.gpu

_DEV = {}


def device(size):
    if size not in _DEV:
        inp = X.make_inputs(size, "unionized", "event")
        sd = X.grid_init(inp)
        print(f"[gpu-test] upload {size} unionized", flush=True)
        _DEV[size] = (inp, sd, X.Device(inp, sd, build_fp32=True))
    return _DEV[size][2]


def _same(got, want):
    for key in ("vsum", "n_fallback", "material_counts", "kernel_id"):
        assert got[key] == want[key], key


@pytest.mark.parametrize("size", ["small", "large"])
@pytest.mark.parametrize("k,prec", [(X.K_MAT_E_SORT, X.XS_FP64), (X.K_MAT_E_SORT, X.XS_FP32),
                                    (X.K_MAT_SORT, X.XS_FP64), (X.K_FUEL_SORT, X.XS_FP64),
                                    (X.K_SPLIT, X.XS_FP64), (X.K_BASELINE, X.XS_FP64)])
def test_noalloc_matches_event_two_calls(size, k, prec):
    dev = device(size)
    calls = [(X.NOALLOC_MAX_LOOKUPS, 1_234_567), (100_003, 7)]  # the second smaller
    want = [dev.event(n, first, kernel_id=k, precision=prec) for n, first in calls]
    got = [dev.event_noalloc(n, first, kernel_id=k, precision=prec) for n, first in calls]
    print(f"[gpu-test] noalloc {size} k={k} prec={prec}: {got[0]['ms_total']:.3f} ms "
          f"(event {want[0]['ms_total']:.3f} ms)", flush=True)
    for g, w in zip(got, want):
        _same(g, w)
    if k in (X.K_MAT_SORT, X.K_MAT_E_SORT):
        assert [sum(g["material_counts"]) for g in got] == [n for n, _ in calls]


def test_noalloc_two_handles_share_the_array():
    """two handles on one GPU use the same array, one call after the other"""
    a, b = device("small"), device("large")
    for rep in range(2):
        for dev, n in ((a, 300_000), (b, 50_000), (a, 2_000)):
            _same(dev.event_noalloc(n, rep, kernel_id=X.K_MAT_E_SORT),
                  dev.event(n, rep, kernel_id=X.K_MAT_E_SORT))


def test_noalloc_limits():
    dev = device("small")
    with pytest.raises(X.XsbError) as e:
        dev.event_noalloc(X.NOALLOC_MAX_LOOKUPS + 1, kernel_id=X.K_MAT_E_SORT)
    assert e.value.code == 1 and "maximum" in str(e.value)
    a = X.XsbEventArgs()
    a.lookups = 10
    a.seed = X.STARTING_SEED
    a.kernel_id = X.K_MAT_E_SORT
    macro = np.empty((10, 5))
    a.out_macro = macro.ctypes.data
    r = X.XsbResult()
    assert dev.lib.xsb_hip_event_noalloc(dev.handle, C.byref(a), C.byref(r)) == 1
    assert dev.event_noalloc(0, kernel_id=X.K_MAT_E_SORT)["vsum"] == 0
