"""The cross-section update on the device (adg_set_xs_tables /
adg_update_xs: xsec::XS_updt, src/mod_xsec.f90:11-46) through the C-ABI.

Every node cross section is a chain of element-wise passes, so the bar is
bit-exact: against the reference's own output on NEACRP A1 (every feedback
table) at two parameter states, and against the oracle at random node-wise
parameters and rod positions on node boundaries, the core top and fully in.
"""

import numpy as np
import pytest

from conftest import golden
from adpres_amd.engine import ADG_E_XS, ADGError, Engine
from adpres_amd.problem import problem_from_dict, xs_tables_from_dict

pytestmark = pytest.mark.gpu

XS_DEV = ("D", "sigr", "nuf", "sigf", "sigs")


def state(d, i):
    return (float(d[f"s{i}_bcon"]), d[f"s{i}_ftem"], d[f"s{i}_mtem"], d[f"s{i}_cden"], d[f"s{i}_bpos"])


def test_gpu_xs_update_matches_reference():
    d = golden("neacrp_a1_xsupd")
    p = problem_from_dict(d)
    t = xs_tables_from_dict(d)
    e = Engine(p)
    e.set_xs_tables(t)
    for i in (0, 1):
        print(f"neacrp A1 state {i}", flush=True)
        e.update_xs(*state(d, i))
        out = e.get_xs()
        for k in XS_DEV:
            ref = getattr(p, k) if i == 0 else d["s1_" + k]
            assert np.array_equal(out[k], ref), (i, k)
    e.close()


def test_gpu_xs_update_matches_oracle_random_states():
    from oracle.oracle import xs_update
    d = golden("neacrp_a1_xsupd")
    p = problem_from_dict(d)
    t = xs_tables_from_dict(d)
    e = Engine(p)
    e.set_xs_tables(t)
    rng = np.random.default_rng(23)
    coreh = float(np.sum(p.zdel))
    zb = np.cumsum(p.zdel[::-1])
    for trial in range(4):
        N = p.nnod
        ftem = d["s0_ftem"] * rng.uniform(0.7, 1.6, N)
        mtem = d["s0_mtem"] + rng.uniform(-15, 15, N)
        cden = d["s0_cden"] * rng.uniform(0.9, 1.05, N)
        bcon = float(rng.uniform(0, 2500))
        # rod tips inside nodes, on node boundaries, at the top and fully in
        depths = np.concatenate([rng.uniform(0, coreh, 3), zb[rng.integers(0, len(zb), 2)], [0.0, coreh]])
        depths = rng.permutation(depths)[:t["nb"]]
        bpos = (t["coreh"] - t["pos0"] - depths) / t["ssize"]
        print(f"random state {trial}", flush=True)
        rc, ref = xs_update(p, t, bcon, ftem, mtem, cden, bpos)
        assert rc == 0
        e.update_xs(bcon, ftem, mtem, cden, bpos)
        out = e.get_xs()
        for k in XS_DEV:
            assert np.array_equal(out[k], ref[k]), (trial, k)
    e.close()


@pytest.mark.parametrize("lds", ["1", "0"])
@pytest.mark.parametrize("case", [("core3d_8g", 2), ("batch", 7), ("batch", 0)])
def test_gpu_xs_update_synthetic_tables(case, lds, monkeypatch):
    """Synthetic feedback tables (every feedback, rods in an eighth of the
    columns) on the 8-group core, a 3-group and a 4-group core (the kernel's
    run-time and compiled group counts), with the tables staged in LDS and
    read through the caches: bit-exact against the oracle."""
    monkeypatch.setenv("ADG_XS_LDS", lds)
    from adpres_amd.configs import batch_core, feedback_tables, generate
    from oracle.oracle import xs_update
    name, seed = case
    gc = generate(name, seed) if name != "batch" else batch_core(3, seed)
    p = gc.problem
    t, par = feedback_tables(gc)
    print(f"{name} {seed} G={p.ng} N={p.nnod}", flush=True)
    rc, ref = xs_update(p, t, *par)
    assert rc == 0
    e = Engine(p)
    e.set_xs_tables(t)
    e.update_xs(*par)
    out = e.get_xs()
    for k in XS_DEV:
        assert np.array_equal(out[k], ref[k]), k
    assert e.last_timing_ms()[0] > 0
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_gpu_solve_on_device_cross_sections(precision):
    """A solve on cross sections made by adg_update_xs (at 600 ppm, fuel
    temperatures up 30 %, moderator up 3 K, coolant density down 2 %) is the
    solve on the oracle's cross sections for the same parameters uploaded by
    adg_set_xs: identical k, flux and outer count, and k within 1 pcm of the
    oracle's solve."""
    import dataclasses
    from oracle.oracle import Oracle, xs_update
    d = golden("neacrp_a1_xsupd")
    p = problem_from_dict(d)
    t = xs_tables_from_dict(d)
    par = (600.0, d["s0_ftem"] * 1.3, d["s0_mtem"] + 3.0, d["s0_cden"] * 0.98, d["s0_bpos"])
    rc, x = xs_update(p, t, *par)
    assert rc == 0
    q = dataclasses.replace(p, **x)
    res = []
    for dev in (True, False):
        e = Engine(q, precision=precision)
        if dev:
            e.set_xs_tables(t)
            e.update_xs(*par)
        e.outer()
        res.append((e.Ke, e.state()["f0"], e.niter))
        e.close()
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]
    o = Oracle(q)
    assert o.outer() == 0
    assert abs(res[0][0] - o.Ke) < 1e-5


def test_gpu_xs_update_reports_reference_stop():
    """A rod increment driving sigtr below 1e-5 is where the reference
    stops (Dsigr_updt, src/mod_xsec.f90:215): ADG_E_XS."""
    d = golden("neacrp_a1_xsupd")
    p = problem_from_dict(d)
    t = xs_tables_from_dict(d)
    t["rod"] = [a * 0 for a in t["rod"]]
    t["rod"][0] = t["rod"][0] - 10.0
    e = Engine(p)
    e.set_xs_tables(t)
    b = d["s1_bpos"] * 0.0                       # every bank in
    with pytest.raises(ADGError) as ei:
        e.update_xs(float(d["s1_bcon"]), d["s1_ftem"], d["s1_mtem"], d["s1_cden"], b)
    assert ei.value.status == ADG_E_XS
    # the context recovers with the deck's tables
    e.set_xs_tables(xs_tables_from_dict(d))
    e.update_xs(*state(d, 1))
    assert np.array_equal(e.get_xs()["D"], d["s1_D"])
    e.close()


# ---------------------------------------------------------------------------
# branch tables (%XTAB decks): xsec::XStab_updt
# ---------------------------------------------------------------------------

XTAB_DEV = XS_DEV + ("dc",)


def xtab_state(d, i):
    return (float(d[f"s{i}_bcon"]), d[f"s{i}_ftem"], d[f"s{i}_mtem"], d[f"s{i}_cden"], d[f"s{i}_bpos"])


def test_gpu_xtab_update_matches_reference():
    """XStab_updt on the device for the MOX part-3 HELIOS deck (branch
    tables with one to three points per parameter, rodded sets, eight
    banks): bit for bit the reference's node cross sections and
    discontinuity factors at both dumped states."""
    from adpres_amd.problem import branch_tables_from_dict
    d = golden("mox_part3_helios_xtab")
    p = problem_from_dict(d)
    e = Engine(p)
    e.set_xtab_tables(branch_tables_from_dict(d))
    for i in (0, 1):
        print(f"mox part3 helios state {i}", flush=True)
        e.update_xs(*xtab_state(d, i))
        out = e.get_xs()
        for k in XTAB_DEV:
            ref = getattr(p, k) if i == 0 else d["s1_" + k]
            assert np.array_equal(out[k], ref), (i, k)
    e.close()


@pytest.mark.parametrize("lds", ["1", "0"])
def test_gpu_xtab_update_matches_oracle(lds, monkeypatch):
    """Random node-wise parameters inside the branch ranges and in the
    allowed extrapolation, rod tips inside nodes, on node boundaries, at the
    top and fully in; and affine tables with three points on every
    parameter (the moderator-temperature stage the decks do not branch on):
    bit-exact against the oracle, with the tiled kernel (entries through
    LDS) and the one-node-per-thread kernel."""
    monkeypatch.setenv("ADG_XS_LDS", lds)
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle import linear_branch_tables
    from oracle.oracle import xtab_update
    from adpres_amd.problem import branch_tables_from_dict
    d = golden("mox_part3_helios_xtab")
    p = problem_from_dict(d)
    t = branch_tables_from_dict(d)
    e = Engine(p)
    e.set_xtab_tables(t)
    rng = np.random.default_rng(31)
    N = p.nnod
    coreh = float(np.sum(p.zdel))
    zb = np.cumsum(p.zdel[::-1])
    for trial in range(3):
        ftem = rng.uniform(600.0, 1500.0, N)
        mtem = rng.uniform(500.0, 600.0, N)
        cden = rng.uniform(0.60, 0.75, N)
        bcon = [10.0, 1990.0, 2050.0][trial]
        depths = np.concatenate([rng.uniform(0, coreh, 4), zb[rng.integers(0, len(zb), 2)], [0.0, coreh]])
        depths = rng.permutation(depths)[:t["nb"]]
        bpos = (t["coreh"] - t["pos0"] - depths) / t["ssize"]
        print(f"random xtab state {trial}", flush=True)
        rc, ref = xtab_update(p, t, bcon, ftem, mtem, cden, bpos)
        assert rc == 0
        e.update_xs(bcon, ftem, mtem, cden, bpos)
        out = e.get_xs()
        for k in XTAB_DEV:
            assert np.array_equal(out[k], ref[k]), (trial, k)
    # affine tables, every stage of the interpolation
    lt, _, _ = linear_branch_tables(p)
    e.set_xtab_tables(lt)
    par = (1450.0, rng.uniform(450.0, 1400.0, N), rng.uniform(545.0, 600.0, N), rng.uniform(0.55, 0.85, N))
    rc, ref = xtab_update(p, lt, *par)
    assert rc == 0
    e.update_xs(*par)
    out = e.get_xs()
    for k in XTAB_DEV:
        assert np.array_equal(out[k], ref[k]), k
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_gpu_solve_on_device_branch_cross_sections(precision):
    """A solve on branch-table cross sections made by adg_update_xs is the
    solve on the oracle's cross sections and discontinuity factors for the
    same parameters uploaded by adg_set_xs: identical k, flux and outer
    count, and k within 1 pcm of the oracle's solve."""
    import dataclasses
    from oracle.oracle import xtab_update
    from adpres_amd.problem import branch_tables_from_dict
    d = golden("mox_part3_helios_xtab")
    p = problem_from_dict(d)
    t = branch_tables_from_dict(d)
    par = (1500.0,) + xtab_state(d, 1)[1:]      # the perturbed state at 1500 ppm
    rc, x = xtab_update(p, t, *par)
    assert rc == 0
    q = dataclasses.replace(p, **x)
    res = []
    for dev in (True, False):
        e = Engine(q, precision=precision)
        if dev:
            e.set_xtab_tables(t)
            e.update_xs(*par)
        e.outer()
        res.append((e.Ke, e.state()["f0"], e.niter))
        e.close()
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]
    from oracle.oracle import Oracle
    o = Oracle(q)
    assert o.outer() == 0
    assert abs(res[0][0] - o.Ke) < 1e-5


def test_gpu_xtab_update_reports_out_of_range():
    """A boron concentration beyond brInterp's 100 ppm extrapolation is
    where the reference stops: ADG_E_XS, and the context recovers."""
    from adpres_amd.problem import branch_tables_from_dict
    d = golden("mox_part3_helios_xtab")
    p = problem_from_dict(d)
    e = Engine(p)
    e.set_xtab_tables(branch_tables_from_dict(d))
    b, ft, mt, cd, bp = xtab_state(d, 1)
    with pytest.raises(ADGError) as ei:
        e.update_xs(2150.0, ft, mt, cd, bp)
    assert ei.value.status == ADG_E_XS
    e.update_xs(b, ft, mt, cd, bp)
    assert np.array_equal(e.get_xs()["dc"], d["s1_dc"])
    e.close()


@pytest.mark.parametrize("lds", ["1", "0"])
def test_gpu_xtab_update_synthetic_8_groups(lds, monkeypatch):
    """Synthetic branch tables (three points on every parameter, rodded
    sets) on the 8-group core: bit-exact against the oracle."""
    from adpres_amd.configs import branch_tables, generate
    from oracle.oracle import xtab_update
    monkeypatch.setenv("ADG_XS_LDS", lds)
    gc = generate("core3d_8g", 2)
    p = gc.problem
    t, par = branch_tables(gc)
    rc, ref = xtab_update(p, t, *par)
    assert rc == 0
    e = Engine(p)
    e.set_xtab_tables(t)
    e.update_xs(*par)
    out = e.get_xs()
    for k in XTAB_DEV:
        assert np.array_equal(out[k], ref[k]), k
    e.close()



def test_gpu_xs_update_in_a_batch():
    """Per-core cross-section updates in a batch context (adg_set_xs_tables_core /
    adg_update_xs_core): each core of a batch of the 8-group core's cousins
    gets its own tables and parameters; the batch solve then matches the
    single-core solves on the oracle's cross sections of the same states to
    1 pcm."""
    import dataclasses
    from adpres_amd.configs import batch_core, feedback_tables
    from adpres_amd.engine import Batch
    from oracle.oracle import Oracle, xs_update
    cores = [batch_core(3, i) for i in (3, 5, 8)]            # 2-group cores
    probs = [gc.problem for gc in cores]
    b = Batch(probs)
    refs = []
    for j, gc in enumerate(cores):
        t, par = feedback_tables(gc, seed=11 + j)
        b.set_xs_tables(j, t)
        b.update_xs(j, *par)
        rc, x = xs_update(gc.problem, t, *par)
        assert rc == 0
        o = Oracle(dataclasses.replace(gc.problem, **x))
        refs.append((o.outer(), o.Ke))
    st, ni, ke = b.outer()
    for j, (rc, k) in enumerate(refs):
        assert rc == st[j]
        if rc == 0:
            assert abs(ke[j] - k) < 1e-5, j
    assert any(rc == 0 for rc, _ in refs)
    b.close()


def test_gpu_xs_and_maps_reject_bad_input():
    """Host-side checks of the cross-section and output-map entry points:
    an update before any tables, tables for another material count, a bank
    map naming a bank beyond nb, and assembly divisions that do not add up
    to the node counts are refused with ADG_E_NOT_SET / ADG_E_ARG before
    anything is launched; the context stays usable."""
    from adpres_amd.engine import ADG_E_ARG, ADG_E_NOT_SET
    d = golden("neacrp_a1_xsupd")
    p = problem_from_dict(d)
    t = xs_tables_from_dict(d)
    e = Engine(p)
    with pytest.raises(ADGError) as ei:
        e.update_xs(*state(d, 0))
    assert ei.value.status == ADG_E_NOT_SET
    bad = dict(t, nmat=t["nmat"] + 1)
    with pytest.raises(ADGError) as ei:
        e.set_xs_tables(bad)
    assert ei.value.status == ADG_E_ARG
    fb = t["fbmap"].copy()
    fb[fb > 0] = t["nb"] + 1
    with pytest.raises(ADGError) as ei:
        e.set_xs_tables(dict(t, fbmap=fb))
    assert ei.value.status == ADG_E_ARG
    e.set_xs_tables(t)
    e.update_xs(*state(d, 1))
    assert np.array_equal(e.get_xs()["D"], d["s1_D"])
    with pytest.raises(ADGError) as ei:
        e.power_maps(2, 2, 1, [1, 1], [1, 1], [1])
    assert ei.value.status == ADG_E_ARG
    e.close()
