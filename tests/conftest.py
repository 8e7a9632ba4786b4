import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")
    config.addinivalue_line("markers", "slow: CPU test taking more than ~10 s")


def _ensure_built():
    """Build what is missing (never rebuilds what exists)."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "libxsb_oracle.so")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"])
    pkg = os.path.join(ROOT, "xsbench_amd")
    if not (os.path.exists(os.path.join(pkg, "lib", "libxsb_hip.so"))
            and os.path.exists(os.path.join(pkg, "lib", "libxsbench.so"))):
        subprocess.check_call(["make", "-s", "-C", pkg, "-j8"])


_ensure_built()


@pytest.fixture(scope="session")
def golden_hashes():
    with open(os.path.join(GOLDEN, "hashes.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Oracle()


def golden_npz(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


# (size, grid, gridpoints) of each golden data configuration
DATA_CFGS = {
    "small_unionized": ("small", "unionized", None),
    "small_hash": ("small", "hash", None),
    "small_nuclide": ("small", "nuclide", None),
    "large_unionized": ("large", "unionized", None),
    "large_hash": ("large", "hash", None),
    "xl45k_hash": ("XL", "hash", 45000),
}
