"""The persistent BiCGSTAB kernel (one launch per group solve, state on chip)
against the flat two-kernels-per-iteration form of the same iteration
(cmfd::bicg, src/mod_cmfd.f90:1203-1243) through the C-ABI, and its timeout
status (forced from the host with a one-poll wait limit)."""

import os

import numpy as np
import pytest

from conftest import golden, probe_fixtures
from adpres_amd.configs import generate
from adpres_amd.engine import ADG_E_DEVICE_SYNC, ADG_OK, Engine
from adpres_amd.problem import problem_from_dict

pytestmark = pytest.mark.gpu


def engine(p, precision="fp64", **env):
    """A context created with the given ADG_* environment (read at creation)."""
    old = {k: os.environ.get(k) for k in env}
    os.environ.update({k: str(v) for k, v in env.items()})
    try:
        return Engine(p, precision)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


CORES = [("core2d_small", 1), ("core3d_2g", 1), ("core3d_8g", 2)]


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("name,seed", CORES)
def test_gpu_persistent_bicg_matches_flat(name, seed, precision):
    """FP64: within 1e-11 of the flat form (the same operations; only the
    summation order of the reductions differs).  FP32: 20 iterations carry
    any rounding difference to 1e-5-1e-4 between two FP32 runs, so the FP32
    persistent form is held to the flat FP64 result: no further from it
    than twice the flat FP32 form's distance, plus 1e-6."""
    p = generate(name, seed).problem
    print(f"persist vs flat bicg: {name} {precision}", flush=True)
    ep = engine(p, precision)
    ef = engine(p, precision, ADG_PERSIST=0)
    e64 = engine(p, "fp64", ADG_PERSIST=0) if precision != "fp64" else None
    info = ep.solver_info()
    assert info["persistent"] and not ef.solver_info()["persistent"], info
    rng = np.random.default_rng(3)
    for g in range(p.ng):
        b = rng.uniform(0.5, 1.5, p.nnod)
        x0 = rng.uniform(0.5, 1.5, p.nnod)
        for nin in (1, 2, 20):
            xp = ep.bicg(g, nin, b, x0)
            xf = ef.bicg(g, nin, b, x0)
            if e64 is None:
                err = np.max(np.abs(xp - xf)) / np.max(np.abs(xf))
                assert err <= 1e-11, (g, nin, err)
            else:
                x64 = e64.bicg(g, nin, b, x0)
                scale = np.max(np.abs(x64))
                ep_err = np.max(np.abs(xp - x64)) / scale
                ef_err = np.max(np.abs(xf - x64)) / scale
                assert ep_err <= 2.0 * ef_err + 1e-6, (g, nin, ep_err, ef_err)
    for e in (ep, ef, e64):
        if e is not None:
            e.close()


@pytest.mark.parametrize("name", probe_fixtures())
def test_gpu_persistent_bicg_probe_matches_reference(name):
    """One bicg call per group on the reference's probe states: the
    persistent kernel within 1e-9 of the reference's own FP64 bicg."""
    d = golden(name)
    p = problem_from_dict(d)
    e = Engine(p)
    assert e.solver_info()["persistent"]
    e.set_state(d["f0"], np.zeros(p.nnod), float(d["Ke"]), d["dn_in"])
    for g in range(p.ng):
        x = e.bicg(g, int(d["nin"]), d["b"][g], d["f0"][g])
        ref = d["x_out"][g]
        assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref)), g
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_gpu_persistent_solve_matches_flat_solve(precision):
    """A whole SANM solve (outer iterations, nodal updates, many persistent
    launches in a replayed graph) against the flat form: same outer count,
    k within 1e-9 and node powers within 1e-7."""
    p = generate("core3d_2g", 1).problem
    print(f"persist vs flat solve: core3d_2g {precision}", flush=True)
    res = []
    for env in ({}, {"ADG_PERSIST": 0}):
        e = engine(p, precision, **env)
        e.outer()
        res.append((e.Ke, e.niter, e.powdis()))
        e.close()
    (kp, np_, pp), (kf, nf, pf) = res
    assert abs(np_ - nf) <= 1, (np_, nf)
    assert abs(kp - kf) < 1e-9, (kp, kf)
    m = pf > 0
    assert np.max(np.abs(pp[m] - pf[m]) / pf[m]) < 1e-7


def test_gpu_persistent_timeout_is_a_status():
    """A one-poll wait limit (set from the host) makes the persistent
    kernel's waits give up: the call returns ADG_E_DEVICE_SYNC (or, if every
    block happened to arrive in time, solves), every kernel exits, and a
    context with the default limit then solves normally."""
    p = generate("core2d_small", 1).problem
    print("persist timeout status", flush=True)
    e = engine(p, "fp64", ADG_PSPIN=1)
    assert e.solver_info()["persistent"]
    rc = e.outer(raise_on_error=False)
    assert rc in (ADG_OK, ADG_E_DEVICE_SYNC), rc
    if rc == ADG_E_DEVICE_SYNC:
        assert "persistent" in e.lib.adg_last_error(e.h).decode()
    e.close()
    e2 = Engine(p)
    assert e2.outer(raise_on_error=False) == ADG_OK
    e2.close()
