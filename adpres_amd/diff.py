"""Differentiation form: k-eff and the PowDis node powers of one core as
differentiable functions of its per-node cross sections.

:class:`CoreSensitivity` holds one core on the GPU (an :class:`Engine`) and
maps a parameter matrix ``x`` of shape (N, P) to the engine's cross
sections.  Columns, for G groups:

    D_1..D_G, sigma_a_1..sigma_a_G, nu-sigma_f_1..nu-sigma_f_G,
    sigma_s(g->h) for g != h in (g, h) row-major order

(G = 2: D1, D2, Sa1, Sa2, NuF1, NuF2, S12, S21).  sigma_f keeps its ratio to
nu-sigma_f (fixed kappa/nu), sigma_r = sigma_a + sum_h sigma_s(g->h) as
``Dsigr_updt`` forms it (src/mod_xsec.f90:199-226), and D enters the
coupling coefficients directly (``coup_coef``, src/mod_cmfd.f90:11-137).

Two models are differentiated:

* ``"full"`` (the default with a nodal kernel): the static solve itself,
  ``cmfd::outer`` with the nodal D-hat updates (src/mod_nodal.f90:18-147),
  converged to a fixed point of both the CMFD eigenproblem and the D-hat
  update.  Its gradient (``adg_backward_nodal``) adds to the CMFD adjoint
  the transpose of the nodal update: the D-hat cotangent w solves
  w = dJ/d(D-hat) + A^T w, A the Jacobian of "D-hat -> CMFD eigenpair ->
  nodal update", by fixed-point iteration, each step one transpose of the
  nodal update (hand-written kernels) and a few outers of the generalized
  adjoint, continued across the steps.
* ``"frozen"``: the converged CMFD eigenproblem with D-hat held at its
  converged value (``adg_backward``).  It differs from the full model's
  gradient by the nodal closure's own sensitivity, up to 6% of the largest
  entry on the 34x34x26 two-group core (4-31% on individual entries); with
  the FDM kernel (no nodal update) the two models are the same.

:class:`KeffPower` is the torch.autograd.Function around either.
"""

from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np
import torch

from .engine import Engine
from .problem import Problem


def param_names(G: int):
    names = [f"D{g + 1}" for g in range(G)] + [f"Sa{g + 1}" for g in range(G)] + [f"NuF{g + 1}" for g in range(G)]
    names += [f"S{g + 1}{h + 1}" for g in range(G) for h in range(G) if h != g]
    return names


class CoreSensitivity:
    """One core, its parameters and the adjoint-based gradients of the
    ``model`` ("full" or "frozen"; default "full" unless the core's kernel
    is FDM)."""

    def __init__(self, p: Problem, precision: str = "fp64", adjoint_iter: Optional[dict] = None,
                 model: Optional[str] = None, nodal_tol: float = 1e-9, nodal_maxit: int = 5000):
        self.base = p
        self.G, self.N = p.ng, p.nnod
        self.model = model or ("frozen" if p.kern == "FDM" else "full")
        if self.model not in ("full", "frozen"):
            raise ValueError(f"model must be 'full' or 'frozen', not {self.model!r}")
        if self.model == "full" and p.kern == "FDM":
            raise ValueError("the full model needs a nodal kernel (PNM or SANM)")
        self.engine = Engine(p, precision)
        self.ratio = np.where(p.nuf > 0, p.sigf / np.where(p.nuf > 0, p.nuf, 1.0), 0.0)
        self.adj_iter = dict(nout=20000, serc=1e-10, ferc=1e-10, nac=1000)
        self.adj_iter.update(adjoint_iter or {})
        # the full model's forward solve: the outer iteration with nodal
        # updates converged tightly, so that the state is a fixed point
        self.fwd_iter = dict(nout=20000, serc=1e-10, ferc=1e-10)
        self.nodal_tol, self.nodal_maxit = nodal_tol, nodal_maxit
        self._adjoint_ok = False
        self.k = None
        self.p = None

    def close(self):
        self.engine.close()

    # -- parameters ----------------------------------------------------------
    def params(self, p: Optional[Problem] = None) -> np.ndarray:
        p = p or self.base
        G = self.G
        cols = [p.D[g] for g in range(G)] + [p.siga[g] for g in range(G)] + [p.nuf[g] for g in range(G)]
        cols += [p.sigs[h, g] for g in range(G) for h in range(G) if h != g]
        return np.ascontiguousarray(np.stack(cols, axis=1))

    def problem_at(self, x: np.ndarray) -> Problem:
        G = self.G
        x = np.asarray(x, np.float64)
        D = np.ascontiguousarray(x[:, :G].T)
        siga = np.ascontiguousarray(x[:, G:2 * G].T)
        nuf = np.ascontiguousarray(x[:, 2 * G:3 * G].T)
        sigs = np.zeros_like(self.base.sigs)
        col = 3 * G
        for g in range(G):
            for h in range(G):
                if h != g:
                    sigs[h, g] = x[:, col]
                    col += 1
        sigr = np.empty_like(siga)
        for g in range(G):
            dum = np.zeros(self.N)
            for h in range(G):
                if h != g:
                    dum = dum + sigs[h, g]
            sigr[g] = siga[g] + dum
        return dataclasses.replace(self.base, sigtr=1.0 / (3.0 * D), siga=siga, nuf=nuf, sigf=nuf * self.ratio,
                                   sigs=sigs, D=D, sigr=sigr)

    # -- forward ---------------------------------------------------------------
    def solve(self, x: np.ndarray, polish: bool = True):
        """Full static solve (nodal D-hat updates included) at x; returns (k,
        powers).  The full model converges it to 1e-10; the frozen model
        (with polish) then iterates the converged state to 1e-11 with D-hat
        fixed, so that the adjoint and the gradients work on an exact
        eigenpair of the model they differentiate."""
        self.engine.set_xs(self.problem_at(x))
        self.engine.reset()
        if self.model == "full":
            self.engine.outer(**self.fwd_iter)
        else:
            self.engine.outer()
            if polish:
                self.solve_frozen(x, self.engine.state())
        self.k, self.p = self.engine.Ke, self.engine.powdis()
        self._adjoint_ok = False
        return self.k, self.p

    def state(self):
        return self.engine.state()

    def solve_frozen(self, x: np.ndarray, state: dict, serc: float = 1e-11, ferc: float = 1e-11,
                     outers: Optional[int] = None):
        """CMFD solve at x with D-hat held at state["dn"] (the model the
        gradients differentiate), warm-started from state; returns (k,
        powers).  With `outers`, exactly that many outer iterations are run
        (two such solves of nearby parameters then carry the same iteration
        error, which cancels in their difference)."""
        self.engine.set_xs(self.problem_at(x))
        self.engine.set_state(state["f0"], state["fs0"], state["Ke"], state["dn"])
        if outers is None:
            self.engine.outer(kern="FDM", serc=serc, ferc=ferc, nout=100000, nac=1000)
        else:
            rc = self.engine.outer(kern="FDM", serc=0.0, ferc=0.0, nout=outers, nac=1000, raise_on_error=False)
            if rc not in (0, 1):   # ADG_OK or ADG_E_MAX_OUTER (the intended stop)
                self.engine._check(rc, "adg_outer")
        self._adjoint_ok = False
        return self.engine.Ke, self.engine.powdis()

    def solve_full(self, x: np.ndarray, state: dict, outers: int):
        """The full model's outer iteration (nodal updates on) at x from
        state for exactly `outers` outers; returns (k, powers).  Two such
        solves of nearby parameters carry the same iteration error, which
        cancels in their difference."""
        self.engine.set_xs(self.problem_at(x))
        self.engine.set_state(state["f0"], state["fs0"], state["Ke"], state["dn"])
        rc = self.engine.outer(serc=0.0, ferc=0.0, nout=outers, raise_on_error=False)
        if rc not in (0, 1):   # ADG_OK or ADG_E_MAX_OUTER (the intended stop)
            self.engine._check(rc, "adg_outer")
        self._adjoint_ok = False
        return self.engine.Ke, self.engine.powdis()

    # -- backward --------------------------------------------------------------
    def gradient(self, dJdk: float, dJdp: Optional[np.ndarray] = None) -> np.ndarray:
        """(N, P) gradient of dJdk * k + <dJdp, powers> at the last solve,
        of the sensitivity's model."""
        if not self._adjoint_ok:
            self.kadj = self.engine.adjoint(**self.adj_iter)
            self._adjoint_ok = True
        if self.model == "full":
            g = self.engine.backward_nodal(dJdk, dJdp, tol=self.nodal_tol, maxit=self.nodal_maxit, **self.adj_iter)
        else:
            g = self.engine.backward(dJdk, dJdp, **self.adj_iter)
        return np.ascontiguousarray(g.T)

    def first_order_dk(self, kern: Optional[str] = None) -> np.ndarray:
        """(N, P) first-order perturbation estimate of dk with the physical
        adjoint of the reference's outer_ad (src/mod_cmfd.f90:602-699: same
        matrices, transposed group coupling, adjoint nodal update), which the
        reference prints for %MODE ADJOINT; exact gradient when D-hat = 0."""
        fwd = self.engine.state()
        self.engine.reset()
        self.engine.outer(adjoint=True, **({"kern": kern} if kern else {}))
        psi = self.engine.state()["f0"]
        self.engine.set_state(fwd["f0"], fwd["fs0"], fwd["Ke"], fwd["dn"])
        self.engine.set_adjoint_flux(psi)
        g = np.ascontiguousarray(self.engine.backward(1.0, None, **self.adj_iter).T)
        self._adjoint_ok = False
        return g


class KeffPower(torch.autograd.Function):
    """(N, P) float64 parameters -> (k-eff, node powers (N,)), with the
    adjoint-kernel backward of the sensitivity's model."""

    @staticmethod
    def forward(ctx, x, sens: CoreSensitivity):
        k, p = sens.solve(x.detach().cpu().numpy())
        ctx.sens = sens
        return torch.tensor(k, dtype=torch.float64), torch.from_numpy(np.array(p))

    @staticmethod
    def backward(ctx, gk, gp):
        dJdk = float(gk) if gk is not None else 0.0
        dJdp = gp.detach().cpu().numpy().astype(np.float64) if gp is not None else None
        if dJdp is not None and not np.any(dJdp):
            dJdp = None
        g = ctx.sens.gradient(dJdk, dJdp)
        return torch.from_numpy(g), None


def keff_and_powers(x: torch.Tensor, sens: CoreSensitivity):
    return KeffPower.apply(x, sens)


def peak_objective(p: torch.Tensor, beta: float = 50.0) -> torch.Tensor:
    """Smooth maximum of the relative node powers: (1/beta) log sum exp(beta p_n / mean),
    the mean taken over nodes with power (a constant of the layout)."""
    nf = int((p > 0).sum())
    return torch.logsumexp(beta * p * nf, dim=0) / beta
