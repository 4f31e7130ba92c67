"""Gradient of the full nodal model: the transpose of one nodal D-hat update
(adg_nodal_vjp, hand-written kernels) and the full-model adjoint built on it
(adg_backward_nodal), on the 34x34x26-node 2-group core of the
differentiation workload and on the 68x68x96-node 8-group core.

The transpose is pinned by the dot-product identity against central
differences of the forward update itself (adg_nodal_update): for a seeded
cotangent w of the new D-hat and a seeded direction d of one input at a
time -- flux, k, old D-hat, each parameter column --
<w, (N(v + e d) - N(v - e d)) / 2e> = <N^T w, d>.  A dropped term, a wrong
sign or index in any one reverse pass moves one of these sums.  Measured:
1e-10 to 1e-8 relative (a 1e-6 relative step); bar 1e-6.

The full-model gradient is pinned by central differences of the whole
static solve (the outer iteration with nodal updates, both solves of a
difference started from the same converged state and run for the same
1600 outers), on the 64 seeded (node, parameter) pairs of the frozen-model
test, at 1e-4 relative plus the differences' resolution.  Measured: 3e-8 to
5e-6 relative.  The frozen-D-hat gradient misses the same differences by
percents, which the test checks as well, so it cannot pass without the
nodal transpose.
"""

import numpy as np
import pytest
import torch

from adpres_amd.configs import generate
from adpres_amd.diff import CoreSensitivity, keff_and_powers, peak_objective
from adpres_amd.engine import ADGError

pytestmark = pytest.mark.gpu

NPAIRS = 64
C4_PARAMS = 7          # D1, D2, Sa1, Sa2, NuF1, NuF2, S12 (the core has no up-scatter)
K_FLOOR, J_FLOOR = 2e-15, 2e-11
FULL_OUTERS = 1600     # both solves of a difference run the same outers with nodal updates, from the same state
STEP = 2e-3
BETA = 50.0
EPS = 1e-6             # relative step of the transpose's dot-product checks
DOT_BAR = 1e-6


def objective_np(p):
    nf = int((p > 0).sum())
    r = BETA * p * nf
    m = r.max()
    return (m + np.log(np.exp(r - m).sum())) / BETA


def objective_grad_np(p):
    nf = int((p > 0).sum())
    r = BETA * p * nf
    w = np.exp(r - r.max())
    return w / w.sum() * nf


def dot_checks(sens, st, x0, cols, seed):
    """Relative mismatch of the transpose per input (name -> value)."""
    e = sens.engine
    rng = np.random.default_rng(seed)
    G, N = sens.G, sens.N
    w = rng.uniform(-1.0, 1.0, (6, G, N))

    def nod(f0=None, Ke=None, dn=None, x=None):
        e.set_xs(sens.problem_at(x0 if x is None else x))
        e.set_state(st["f0"] if f0 is None else f0, st["fs0"], st["Ke"] if Ke is None else Ke,
                    st["dn"] if dn is None else dn)
        e.nodal_update()
        return e.state()["dn"]

    e.set_xs(sens.problem_at(x0))
    e.set_state(st["f0"], st["fs0"], st["Ke"], st["dn"])
    phib, kb, dnb, gx = e.nodal_vjp(w)
    out = {}

    def rel(fd, ad):
        return abs(fd - ad) / abs(fd)

    d = st["f0"] * rng.uniform(-1.0, 1.0, st["f0"].shape)
    out["flux"] = rel(np.sum(w * (nod(f0=st["f0"] + EPS * d) - nod(f0=st["f0"] - EPS * d))) / (2 * EPS),
                      np.sum(phib * d))
    out["k"] = rel(np.sum(w * (nod(Ke=st["Ke"] * (1 + EPS)) - nod(Ke=st["Ke"] * (1 - EPS)))) / (2 * EPS * st["Ke"]),
                   kb)
    d = rng.uniform(-1.0, 1.0, st["dn"].shape) * (np.abs(st["dn"]).mean() + 1e-3)
    out["dn_old"] = rel(np.sum(w * (nod(dn=st["dn"] + EPS * d) - nod(dn=st["dn"] - EPS * d))) / (2 * EPS),
                        np.sum(dnb * d))
    for col in cols:
        d = np.zeros_like(x0)
        d[:, col] = x0[:, col] * rng.uniform(-1.0, 1.0, N)
        out[f"param{col}"] = rel(np.sum(w * (nod(x=x0 + EPS * d) - nod(x=x0 - EPS * d))) / (2 * EPS),
                                 np.sum(gx.T * d))
    e.set_xs(sens.problem_at(x0))
    e.set_state(st["f0"], st["fs0"], st["Ke"], st["dn"])
    return out


@pytest.fixture(scope="module")
def full():
    sens = CoreSensitivity(generate("core3d_2g", 1).problem, model="full")
    x0 = sens.params()
    k0, p0 = sens.solve(x0)
    st = sens.state()
    gk = sens.gradient(1.0, None)
    steps = sens.engine.nodal_adjoint_steps
    gJ = sens.gradient(0.0, objective_grad_np(p0))
    gk_frozen = np.ascontiguousarray(sens.engine.backward(1.0, None, **sens.adj_iter).T)
    yield dict(sens=sens, x0=x0, st=st, k0=k0, p0=p0, gk=gk, gJ=gJ, gk_frozen=gk_frozen, steps=steps)
    sens.close()


@pytest.mark.parametrize("kern", ["SANM", "PNM"])
def test_gpu_nodal_transpose_two_group(kern):
    p = generate("core3d_2g", 1).problem.with_iter(kern=kern)
    sens = CoreSensitivity(p, model="full")
    x0 = sens.params()
    sens.solve(x0)
    res = dot_checks(sens, sens.state(), x0, range(C4_PARAMS), seed=1)
    sens.close()
    print(kern, {k: f"{v:.2e}" for k, v in res.items()}, flush=True)
    assert max(res.values()) <= DOT_BAR, res


def test_gpu_nodal_transpose_eight_group():
    """The 8-group core with up-scatter (the cooperative forward passes'
    group count): columns D1, D8, Sa4, NuF3, NuF8 and the scatterings
    1->2 (down) and 7->6 (up)."""
    p = generate("core3d_8g").problem
    sens = CoreSensitivity(p, model="full")
    x0 = sens.params()
    e = sens.engine
    e.set_xs(sens.problem_at(x0))
    e.reset()
    e.outer(nout=60, raise_on_error=False)   # a state with developed D-hats
    st = e.state()
    G = sens.G
    sidx = [(g, h) for g in range(G) for h in range(G) if h != g]
    cols = [0, G - 1, G + 3, 2 * G + 2, 3 * G - 1, 3 * G + sidx.index((0, 1)), 3 * G + sidx.index((6, 5))]
    assert np.any(x0[:, cols[-1]] > 0)
    res = dot_checks(sens, st, x0, cols, seed=2)
    sens.close()
    print({k: f"{v:.2e}" for k, v in res.items()}, flush=True)
    assert max(res.values()) <= DOT_BAR, res


def test_gpu_full_model_adjoint_converges(full):
    assert 2 <= full["steps"] < full["sens"].nodal_maxit
    assert full["sens"].engine.nodal_adjoint_resid <= 1e-9


@pytest.mark.parametrize("half", [0, 1])
def test_gpu_full_model_gradients_match_central_differences(full, half):
    sens, x0, st, p0 = full["sens"], full["x0"], full["st"], full["p0"]
    gk, gJ, gf = full["gk"], full["gJ"], full["gk_frozen"]
    rng = np.random.default_rng(4)
    nodes = rng.integers(0, sens.N, NPAIRS)
    cols = rng.integers(0, C4_PARAMS, NPAIRS)
    fuel = p0 > 0
    typical = np.array([np.mean(np.abs(x0[fuel, j])) for j in range(x0.shape[1])])
    sl = slice(half * NPAIRS // 2, (half + 1) * NPAIRS // 2)
    worst_k = worst_J = 0.0
    fds, frz = [], []
    for n, j in zip(nodes[sl], cols[sl]):
        h = STEP * (abs(x0[n, j]) if x0[n, j] != 0 else typical[j])
        xp = x0.copy(); xp[n, j] += h
        xm = x0.copy(); xm[n, j] -= h
        kp, pp = sens.solve_full(xp, st, FULL_OUTERS)
        km, pm = sens.solve_full(xm, st, FULL_OUTERS)
        fdk = (kp - km) / (2 * h)
        fdJ = (objective_np(pp) - objective_np(pm)) / (2 * h)
        ek = abs(gk[n, j] - fdk) / (1e-4 * abs(fdk) + K_FLOOR / h)
        eJ = abs(gJ[n, j] - fdJ) / (1e-4 * abs(fdJ) + J_FLOOR / h)
        worst_k, worst_J = max(worst_k, ek), max(worst_J, eJ)
        fds.append(fdk)
        frz.append(gf[n, j])
        assert ek <= 1.0, (int(n), int(j), gk[n, j], fdk)
        assert eJ <= 1.0, (int(n), int(j), gJ[n, j], fdJ)
    fds, frz = np.array(fds), np.array(frz)
    gap = np.max(np.abs(frz - fds)) / np.max(np.abs(fds))
    print(f"worst k {worst_k:.3f} J {worst_J:.3f} of the bar; frozen-D-hat gradient off by {gap:.3e} "
          f"of the largest difference", flush=True)
    assert gap > 1e-2


def test_gpu_full_model_gradient_pnm_kernel():
    """The same check with the PNM nodal kernel (constant A..H): the
    gradient of k on eight seeded pairs of significant entries in the fuel."""
    sens = CoreSensitivity(generate("core3d_2g", 1).problem.with_iter(kern="PNM"), model="full")
    x0 = sens.params()
    k0, p0 = sens.solve(x0)
    st = sens.state()
    gk = sens.gradient(1.0, None)
    gf = np.ascontiguousarray(sens.engine.backward(1.0, None, **sens.adj_iter).T)
    fuel = np.nonzero(p0 > 0)[0]
    big = np.max(np.abs(gk[fuel, :C4_PARAMS]))
    rng = np.random.default_rng(12)
    rows = []
    while len(rows) < 8:
        n, j = int(rng.choice(fuel)), int(rng.integers(0, C4_PARAMS))
        if abs(gk[n, j]) < 1e-3 * big:
            continue
        h = STEP * abs(x0[n, j])
        xp = x0.copy(); xp[n, j] += h
        xm = x0.copy(); xm[n, j] -= h
        kp, _ = sens.solve_full(xp, st, FULL_OUTERS)
        km, _ = sens.solve_full(xm, st, FULL_OUTERS)
        rows.append((gk[n, j], (kp - km) / (2 * h), gf[n, j], h))
    sens.close()
    r = np.array(rows)
    err = np.abs(r[:, 0] - r[:, 1]) / (1e-4 * np.abs(r[:, 1]) + K_FLOOR / r[:, 3])
    gap = np.max(np.abs(r[:, 2] - r[:, 1])) / np.max(np.abs(r[:, 1]))
    print(f"PNM: worst {err.max():.3f} of the bar; frozen-D-hat gradient off by {gap:.3e}", flush=True)
    assert err.max() <= 1.0, r
    assert gap > 1e-3


def test_gpu_autograd_function_full_model(full):
    sens = full["sens"]
    x = torch.tensor(full["x0"], dtype=torch.float64, requires_grad=True)
    k, p = keff_and_powers(x, sens)
    J = peak_objective(p, BETA) + 0.5 * k
    J.backward()
    ref = sens.gradient(0.5, objective_grad_np(p.detach().numpy()))
    assert np.allclose(x.grad.numpy(), ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    assert abs(float(k.detach()) - full["k0"]) < 1e-12


def test_gpu_full_model_status_paths():
    p = generate("core2d_small", 1).problem
    sens = CoreSensitivity(p, model="full")
    e = sens.engine
    x0 = sens.params()
    sens.solve(x0)
    with pytest.raises(ADGError) as ei:            # no adjoint yet
        e.backward_nodal(1.0, None)
    assert ei.value.status == 7                    # ADG_E_NOT_SET
    e.adjoint(**sens.adj_iter)
    with pytest.raises(ADGError) as ei:            # the FDM kernel has no nodal update
        e.backward_nodal(1.0, None, kern="FDM")
    assert ei.value.status == 5                    # ADG_E_ARG
    with pytest.raises(ADGError) as ei:            # one step: the frozen gradient, not converged
        e.backward_nodal(1.0, None, maxit=1)
    assert ei.value.status == 1                    # ADG_E_MAX_OUTER
    g = sens.gradient(1.0, None)
    assert np.all(np.isfinite(g)) and e.nodal_adjoint_resid <= 1e-9
    sens.close()
    with pytest.raises(ValueError):
        CoreSensitivity(p.with_iter(kern="FDM"), model="full")
