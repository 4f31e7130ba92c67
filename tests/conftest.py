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


def output_tokens(text):
    """The numbers (with a decimal point) of the result blocks of an ADPRES
    .out file: from the radial power distribution (AsmPow, src/mod_io.f90)
    to the CPU-time breakdown, i.e. the power, axial, flux and T-H summary
    tables.  Integers (row, plane and position labels) are left out: the
    position of a maximum flips between mirror nodes on a printed tie."""
    import re
    lines = text.splitlines()
    try:
        i0 = next(i for i, ln in enumerate(lines) if "Radial Power Distribution" in ln)
    except StopIteration:
        return []
    i1 = next((i for i, ln in enumerate(lines) if "CPU time breakdown" in ln and i > i0), len(lines))
    num = re.compile(r"[-+]?\d+\.\d+(?:E[-+]\d+)?")
    return [t for ln in lines[i0:i1] for t in num.findall(ln)]


def printed_unit(tok):
    """One unit of the last printed digit of a number token."""
    mant, _, exp = tok.partition("E")
    return 10.0 ** ((int(exp) if exp else 0) - len(mant.split(".")[1]))


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


def lu_pivot_problem():
    """The 18x18x2 two-group core with one fuel node's group-1 data set so
    that the diagonal of its PNM node system (get_a2matvec's matrix
    B E + 3 I, src/mod_nodal.f90:769-805, E = 2/7) vanishes at k = 1:
    0.25 h^2 / D (sigr - nuf / k) = -10.5 with h = 10 cm, D = 1.4,
    sigr = 0.05, nuf = 0.638.  The reference's LU_solve then stops with
    "DIAGONAL ELEMENTS CLOSE TO ZERO" (src/mod_nodal.f90:856-862).
    Returns (problem, node index, (i, j, k) of the node, 1-based)."""
    import dataclasses
    from adpres_amd.configs import generate
    p = generate("core2d_small", 1).problem
    n0 = int(np.nonzero((p.ix == 9) & (p.iy == 9) & (p.iz == 1))[0][0])
    D, sigr, nuf, sigf = (np.array(a, copy=True) for a in (p.D, p.sigr, p.nuf, p.sigf))
    D[0, n0], sigr[0, n0], nuf[0, n0] = 1.4, 0.05, 0.638
    sigf[0, n0] = 0.638 / 2.43
    q = dataclasses.replace(p, D=D, sigr=sigr, nuf=nuf, sigf=sigf, sigtr=1.0 / (3.0 * D))
    return q, n0, (9, 9, 1)
