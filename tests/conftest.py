import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C-ABI)")


def golden(name):
    """Load tests/golden/<name>.npz (written by tests/golden/make_golden.py)."""
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def solve_fixtures():
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*_solve.npz")))
    return [n for n in names if "p_ix" in np.load(os.path.join(GOLDEN, n + ".npz")).files]


def probe_fixtures():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*_probe.npz")))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle
