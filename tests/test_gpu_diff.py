"""Differentiation form on the 34x34x26-node 2-group core: the discrete
adjoint, the adjoint-kernel gradients of k-eff and of a smooth peak-power
objective with respect to per-node D1, D2, sigma_a1, sigma_a2, nu-sigma_f1,
nu-sigma_f2 and sigma_s(1->2), the torch.autograd.Function around them, and
the reference's adjoint-flux first-order perturbation estimate (with the
finite-difference kernel, where the reference's adjoint mode converges).

The gradients here are those of the converged CMFD eigenproblem with the
nodal D-hat held at its converged value (model "frozen"; the full nodal
model's gradient has its own tests), so the finite differences solve that
same model (CMFD outers with D-hat frozen, both solves of a difference
warm-started from the same converged state and run for the same 1500
outers, so that their iteration error cancels).  Bar: central differences
with a 0.2% step on 64 seeded (node, parameter) pairs agree within 1e-4
relative, plus the differences' resolution (about 2e-15 in k and 2e-11 in
the objective, divided by the step).
"""

import numpy as np
import pytest
import torch

from conftest import golden
from adpres_amd.configs import generate
from adpres_amd.diff import CoreSensitivity, keff_and_powers, peak_objective

pytestmark = pytest.mark.gpu

NPAIRS = 64
C4_PARAMS = 7          # D1, D2, Sa1, Sa2, NuF1, NuF2, S12 (the core has no up-scatter)
K_FLOOR, J_FLOOR = 2e-15, 2e-11
FD_OUTERS = 1500       # both solves of a difference run the same outers, from the same state
STEP = 2e-3            # relative step: truncation (~1e-5 relative) below the bar, k differences >> 1e-15
BETA = 50.0


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


@pytest.fixture(scope="module")
def core():
    sens = CoreSensitivity(generate("core3d_2g", 1).problem, model="frozen")
    x0 = sens.params()
    k0, p0 = sens.solve(x0)
    yield sens, x0, sens.state(), k0, p0
    sens.close()


def test_gpu_discrete_adjoint_eigenvalue_is_k(core):
    sens, x0, st, k0, p0 = core
    kadj = sens.engine.adjoint(**sens.adj_iter)
    assert abs(kadj - k0) < 1e-11
    psi, pfp = sens.engine.adjoint_flux()
    assert pfp > 0 and np.all(np.isfinite(psi))


def test_gpu_gradients_match_central_differences(core):
    sens, x0, st, k0, p0 = core
    sens.solve(x0)
    p_at = sens.p
    gk = sens.gradient(1.0, None)
    gJ = sens.gradient(0.0, objective_grad_np(p_at))
    rng = np.random.default_rng(4)
    nodes = rng.integers(0, sens.N, NPAIRS)
    cols = rng.integers(0, C4_PARAMS, NPAIRS)
    fuel = p_at > 0
    typical = np.array([np.mean(np.abs(x0[fuel, j])) for j in range(x0.shape[1])])
    worst_k = worst_J = 0.0
    for n, j in zip(nodes, cols):
        h = STEP * (abs(x0[n, j]) if x0[n, j] != 0 else typical[j])
        xp = x0.copy(); xp[n, j] += h
        xm = x0.copy(); xm[n, j] -= h
        kp, pp = sens.solve_frozen(xp, st, outers=FD_OUTERS)
        km, pm = sens.solve_frozen(xm, st, outers=FD_OUTERS)
        fdk = (kp - km) / (2 * h)
        fdJ = (objective_np(pp) - objective_np(pm)) / (2 * h)
        ek = abs(gk[n, j] - fdk) / (1e-4 * abs(fdk) + K_FLOOR / h)
        eJ = abs(gJ[n, j] - fdJ) / (1e-4 * abs(fdJ) + J_FLOOR / h)
        worst_k, worst_J = max(worst_k, ek), max(worst_J, eJ)
        assert ek <= 1.0, (int(n), int(j), gk[n, j], fdk)
        assert eJ <= 1.0, (int(n), int(j), gJ[n, j], fdJ)


def test_gpu_autograd_function(core):
    sens, x0, st, k0, p0 = core
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    k, p = keff_and_powers(x, sens)
    J = peak_objective(p, BETA) + 0.5 * k
    J.backward()
    ref = sens.gradient(0.5, objective_grad_np(p.detach().numpy()))
    assert np.allclose(x.grad.numpy(), ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    assert abs(float(J) - (objective_np(p.detach().numpy()) + 0.5 * float(k))) < 1e-12


def test_gpu_first_order_estimate_with_the_reference_adjoint_flux():
    """ADPRES's own adjoint (outer_ad, %MODE ADJOINT) with the
    finite-difference kernel: D-hat = 0 makes every matrix symmetric, so its
    physical adjoint is the discrete adjoint of the forward operator, and the
    first-order perturbation estimate of dk built on the reference's adjoint
    flux (fixture, reference run to 1e-10) is the exact gradient.  (With
    SANM the reference's adjoint outer on this core stops at outer 109 with
    'two-node nonlinear iteration seems not stable', the same status the
    engine returns in that mode.)"""
    d = golden("core3d_2g_s1_fdm_adjoint")
    sens = CoreSensitivity(generate("core3d_2g", 1).problem.with_iter(kern="FDM", nout=20000, serc=1e-10,
                                                                      ferc=1e-10))
    x0 = sens.params()
    k, _ = sens.solve(x0)
    assert abs(k - float(d["Ke"])) < 1e-9
    g = sens.gradient(1.0, None)
    psi, _ = sens.engine.adjoint_flux()
    ref = d["psi"]
    assert np.max(np.abs(psi / psi.max() - ref / ref.max())) < 1e-6
    sens.engine.set_adjoint_flux(ref)
    fo = np.ascontiguousarray(sens.engine.backward(1.0, None, **sens.adj_iter).T)
    sens.close()
    scale = np.abs(g[:, :C4_PARAMS]).max(axis=0)
    assert np.all(np.abs(fo[:, :C4_PARAMS] - g[:, :C4_PARAMS]) <= 1e-6 * scale)
