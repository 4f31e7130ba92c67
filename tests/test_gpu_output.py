"""The radial and axial power maps and the radial flux map on the device
(adg_power_maps, adg_flux_map: the numbers of io::AsmPow, io::AxiPow and
io::AsmFlux, src/mod_io.f90:3267-3644) through the C-ABI: bit for bit the
oracle's maps of the same node powers and flux, and the reference's printed
maps digit for digit."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from adpres_amd.engine import Engine
from adpres_amd.problem import deck_layout, problem_from_dict

pytestmark = pytest.mark.gpu

MAP_DECKS = [("iaea3d", "IAEA3Ds"), ("iaea2d", "IAEA2D"), ("pnm", "PNM"), ("biblis", "BIBLIS"),
             ("koeberg", "KOEBERG"), ("dvp", "DVP"), ("fdm", "FDM")]


@pytest.mark.parametrize("fixture,deck", MAP_DECKS)
def test_gpu_power_maps(fixture, deck):
    from oracle.oracle import power_maps
    from test_oracle import printed_power_maps
    d = golden(fixture + "_solve")
    p = problem_from_dict(d)
    lay = deck_layout(os.path.join(GOLDEN, "decks", deck))
    print(f"{deck} layout {lay[:3]}", flush=True)
    e = Engine(p)
    e.set_state(d["f0"], d["fs0"], float(d["Ke"]), d["dn"])
    pw = e.powdis()
    fasm, pmax, faxi, amax = e.power_maps(*lay)
    rc, ofasm, opmax, ofaxi, oamax = power_maps(p, *lay, pw)
    assert rc == 0
    assert np.array_equal(fasm, ofasm) and pmax == opmax
    assert np.array_equal(faxi, ofaxi) and amax == oamax
    # and what the reference prints for the deck
    rad, axi, amaxv = printed_power_maps(fasm, pmax, faxi)
    toks = [str(t) for t in golden("outputs_sweep")[deck.replace("-", "_")]]
    assert toks[:len(rad)] == rad
    nz = lay[2]
    ax = toks[len(rad):len(rad) + 2 * nz + 1]
    assert ax[0:2 * nz:2] == axi and ax[2 * nz] == amaxv
    # the flux map of the same state
    from oracle.oracle import flux_map
    from test_oracle import deck_prints_flux, printed_flux_map
    nx, ny, _, xdiv, ydiv, _ = lay
    flx, negf = e.flux_map(nx, ny, xdiv, ydiv, 1.0)
    rc, oflx, onegf = flux_map(p, nx, ny, xdiv, ydiv, e.state()["f0"], 1.0)
    assert rc == 0 and np.array_equal(flx, oflx) and negf == onegf == 0
    if deck_prints_flux(os.path.join(GOLDEN, "decks", deck)):
        fl = printed_flux_map(flx)
        i0 = len(rad) + 2 * nz + 1
        assert toks[i0:i0 + len(fl)] == fl
    e.close()
