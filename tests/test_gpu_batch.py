"""Batching form: many independent cores in one context (adg_create_batch),
every launch and every graph replay covering all of them.

Bars: a batch of one core is the drop-in path itself (bit-identical); a
batch of uneven cores gives every core what its own single-core solve gives
up to the summation order of the reductions (FP64: k within 1e-9, node
powers within 1e-7 relative; FP32+refinement, whose FP32 iterates carry
that order into the stopping outer: k within 1e-7, powers within 1e-5;
outer count within 2 for both); over the 256-core workload every core the
reference solves has k within 1 pcm of the reference's own FP64 solve of
that core (the sampled cores' node powers within 1e-4 relative), every core
the reference stops as nodally unstable is stopped with that status, in FP64
and in FP32+refinement.
"""

import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from adpres_amd.configs import batch_core
from adpres_amd.engine import ADG_E_NODAL_UNSTABLE, ADG_OK, Batch, Engine

pytestmark = pytest.mark.gpu

SEED = 3


def rel_err(p, ref):
    m = ref > 0
    return float(np.max(np.abs(p[m] - ref[m]) / ref[m]))


def test_gpu_batch_of_one_is_the_drop_in_path():
    p = batch_core(SEED, 0).problem
    e = Engine(p)
    e.outer()
    k1, n1, p1 = e.Ke, e.niter, e.powdis()
    e.close()
    b = Batch([p])
    st, ni, ke = b.outer()
    assert st[0] == ADG_OK and ni[0] == n1 and ke[0] == k1
    assert np.array_equal(b.powdis(0), p1)
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_gpu_batch_matches_single_core_solves(precision):
    cores = [c.problem for c in (batch_core(SEED, i) for i in range(24)) if c.problem.ng == 2][:5]
    assert len(cores) == 5 and len({q.nzz > 2 for q in cores}) == 2   # 2D and 3D, uneven sizes
    single = []
    for q in cores:
        e = Engine(q, precision)
        e.outer()
        single.append((e.Ke, e.niter, e.powdis()))
        e.close()
    ktol, ptol = (1e-9, 1e-7) if precision == "fp64" else (1e-7, 1e-5)
    b = Batch(cores, precision)
    st, ni, ke = b.outer()
    assert np.all(st == ADG_OK)
    for j, (k, n, pw) in enumerate(single):
        assert abs(ke[j] - k) < ktol, j
        assert abs(int(ni[j]) - n) <= 2, j
        assert rel_err(b.powdis(j), pw) < ptol, j
    # a second call resumes every core from its converged state
    st2, ni2, ke2 = b.outer()
    assert np.all(st2 == ADG_OK) and np.all(ni2 < ni // 4 + 2)
    assert np.max(np.abs(ke2 - ke)) < 1e-5
    b.close()


# cores of the batching workload at the edge of the reference's nodal
# instability stop, whose outcome the FP32+refinement form may flip
EDGE_FP32 = {93, 121, 234}


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_gpu_batch256_matches_reference(precision):
    if not os.path.exists(os.path.join(GOLDEN, f"batch_s{SEED}_solve.npz")):
        pytest.skip("reference fixture of the batching workload not generated (tests/golden/make_golden.py batch)")
    d = golden(f"batch_s{SEED}_solve")
    n = int(d["n"])
    cores = [batch_core(SEED, i).problem for i in range(n)]
    keff = np.zeros(n)
    status = np.zeros(n, np.int32)
    powers = {}
    sample = [int(i) for i in d["pow_cores"]]
    peak = {}
    for G in sorted({q.ng for q in cores}):
        idx = [i for i in range(n) if cores[i].ng == G]
        b = Batch([cores[i] for i in idx], precision)
        st, ni, ke = b.outer()
        status[idx] = st
        keff[idx] = ke
        for j, i in enumerate(idx):
            if i in sample:
                powers[i] = b.powdis(j)
            if i in EDGE_FP32:
                peak[i] = float(np.max(b.trace(j, int(ni[j]))["ndmax"]))
        b.close()
    # the reference stops some cores with "two-node nonlinear iteration seems
    # not stable" (no k).  In FP64 the engine stops exactly those, with that
    # status, and solves every other core.  In FP32+refinement the FP32
    # rounding of the iterates can move only the cores of EDGE_FP32, which
    # sit at the edge of that instability, to the other side: every other
    # core has the reference's outcome, and every core both solve meets the
    # k bar.
    ok = np.isfinite(d["Ke"])
    solved = ok & (status == ADG_OK)
    differ = np.nonzero(ok != (status == ADG_OK))[0]
    assert np.all((status == ADG_OK) | (status == ADG_E_NODAL_UNSTABLE)), np.unique(status)
    if precision == "fp64":
        assert len(differ) == 0, (differ, status[differ])
    else:
        assert set(differ.tolist()) <= EDGE_FP32, (differ, status[differ])
        for i in differ:
            # each side of the 1e3 stop threshold on the largest D-hat change
            print(f"core {i}: reference {'solves' if ok[i] else 'stops'} it, FP32+refinement status {status[i]}, "
                  f"peak ndmax {peak[i]:.4g}", flush=True)
            if ok[i]:
                assert status[i] == ADG_E_NODAL_UNSTABLE and peak[i] > 1.e3
            else:
                assert status[i] == ADG_OK and peak[i] <= 1.e3
    err = np.abs(keff[solved] - d["Ke"][solved])
    assert err.max() < 1e-5, (int(np.nonzero(solved)[0][err.argmax()]), float(err.max()))
    for i in sample:
        if solved[i]:
            assert rel_err(powers[i], d[f"pow_{i}"].astype(np.float64)) < 1e-4, i


def test_gpu_batch_of_eight_group_cores_nodal_passes(monkeypatch):
    """Two 8-group cores in one batch (the 68x68x96 core and the same core
    with nu-sigma_f raised by 0.2%), 186 outers (two nodal updates; no
    convergence test).  The cooperative nodal passes and the per-core
    boundary-face lists (their blocks per core, their argmax slots after the
    face blocks') give D-hats, k and node powers bit-identical to the
    per-thread nodal kernels (ADG_NOD_COOP=0) in the same batch; and each
    core's k is within 1e-9 of its own context's.  (The D-hats themselves
    are not compared across the two BiCGSTAB forms: at a few faces of the
    fast group near the outer reflector the interface systems amplify the
    summation-order differences of the flux to 1e-2 relative.)"""
    import dataclasses
    from adpres_amd.configs import generate
    p = generate("core3d_8g", 2).problem
    probs = [p, dataclasses.replace(p, nuf=p.nuf * 1.002, sigf=p.sigf * 1.002)]
    ov = dict(nout=186, serc=0.0, ferc=0.0)
    res = {}
    for coop in ("1", "0"):
        monkeypatch.setenv("ADG_NOD_COOP", coop)
        print(f"batch of two 8-group cores, 186 outers, ADG_NOD_COOP={coop}", flush=True)
        bt = Batch(probs)
        st, ni, ke = bt.outer(**ov)
        assert list(ni) == [186, 186], (st, ni)
        res[coop] = (ke.copy(), [bt.state(b)["dn"] for b in range(2)], [bt.powdis(b) for b in range(2)])
        bt.close()
    (k1, d1, p1), (k0, d0, p0) = res["1"], res["0"]
    assert np.array_equal(k1, k0)
    for b in range(2):
        assert np.array_equal(d1[b], d0[b]), b
        assert np.array_equal(p1[b], p0[b]), b
    monkeypatch.delenv("ADG_NOD_COOP")
    for b, q in enumerate(probs):
        e = Engine(q)
        e.outer(raise_on_error=False, **ov)
        assert e.niter == 186
        assert abs(k1[b] - e.Ke) < 1e-9 * e.Ke, (b, k1[b], e.Ke)
        e.close()


def test_gpu_contexts_on_concurrent_threads_match_sequential():
    """Independent contexts driven from several host threads at once (each on
    its own stream, graph captures serialized inside the library) give the
    same results as the same solves run one after another."""
    import threading
    cores = [batch_core(SEED, i).problem for i in range(6)]
    seq = []
    for q in cores:
        e = Engine(q)
        e.outer()
        seq.append((e.Ke, e.niter, e.powdis()))
        e.close()
    res = [None] * len(cores)

    def run(i):
        e = Engine(cores[i])
        e.outer()
        res[i] = (e.Ke, e.niter, e.powdis())
        e.close()

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(cores))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, (k, n, pw) in enumerate(seq):
        assert res[i] is not None, i
        assert res[i][0] == k and res[i][1] == n and np.array_equal(res[i][2], pw), i
