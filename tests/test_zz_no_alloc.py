"""adg_power_maps_noalloc (the output maps without any device allocation)
through the C-ABI: bit for bit adg_power_maps on two calls in a row, the
second on a smaller core (the scratch keeps the first call's contents), and
a clear ADG_E_ARG for a layout beyond the scratch's limit."""

import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from adpres_amd.engine import ADG_E_ARG, ADGError, Engine
from adpres_amd.problem import deck_layout, problem_from_dict

pytestmark = pytest.mark.gpu


def _maps(fixture, deck):
    d = golden(fixture + "_solve")
    p = problem_from_dict(d)
    lay = deck_layout(os.path.join(GOLDEN, "decks", deck))
    e = Engine(p)
    e.set_state(d["f0"], d["fs0"], float(d["Ke"]), d["dn"])
    ref = e.power_maps(*lay)
    got = e.power_maps(*lay, noalloc=True)
    e.close()
    return ref, got


def test_gpu_power_maps_noalloc_matches_power_maps():
    # first a 3D staggered core, then a smaller 2D one on the same scratch
    for fixture, deck in (("iaea3d", "IAEA3Ds"), ("iaea2d", "IAEA2D")):
        print(f"noalloc maps: {deck}", flush=True)
        (fasm, pmax, faxi, amax), (gfasm, gpmax, gfaxi, gamax) = _maps(fixture, deck)
        assert np.array_equal(fasm, gfasm) and pmax == gpmax
        assert np.array_equal(faxi, gfaxi) and amax == gamax
        assert np.all(np.isfinite(gfasm)) and gfasm.max() > 1.0


def test_gpu_power_maps_noalloc_rejects_layouts_beyond_its_scratch():
    from adpres_amd.configs import _mk_problem, _stack, _two_group_material
    # 60 x 60 assemblies of 6 x 6 nodes: nxx*nyy + nx*ny + nz = 133,202 doubles > 131,072
    rng = np.random.default_rng(5)
    mats = [_two_group_material(rng, True)]
    nx = 60
    geo = dict(xsize=[12.0] * nx, xdiv=[6] * nx, ysize=[12.0] * nx, ydiv=[6] * nx, zsize=[10.0, 10.0],
               zdiv=[1, 1], planars=[np.ones((nx, nx), np.int32)], zpln=[1, 1])
    p, _ = _mk_problem("big2d", 2, _stack(mats), geo, (1, 1, 1, 1, 2, 2), [[1.0, 0.0]])
    print(f"noalloc maps: rejection on {p.nxx}x{p.nyy}x{p.nzz}", flush=True)
    e = Engine(p)
    with pytest.raises(ADGError) as ei:
        e.power_maps(nx, nx, 2, [6] * nx, [6] * nx, [1, 1], noalloc=True)
    assert ei.value.status == ADG_E_ARG and "exceeds" in str(ei.value)
    # the allocating entry point takes the same layout
    e.reset()
    e.outer(nout=2, raise_on_error=False)
    fasm, _, faxi, _ = e.power_maps(nx, nx, 2, [6] * nx, [6] * nx, [1, 1])
    assert fasm.shape == (nx, nx) and np.all(np.isfinite(faxi))
    e.close()
