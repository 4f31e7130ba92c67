// Differentiation of the static solution: the discrete adjoint of the
// converged CMFD eigenproblem M phi = (1/k) F phi (D-hat held at its
// converged value), the generalized (fixed-source) adjoint of a response
// of the node powers, and the per-node parameter gradients.
//
// Notation, per core, groups g, h and nodes n:
//   M: block rows A_g (the 7-point matrix of group g, src/mod_cmfd.f90:217-304)
//      minus the in-scatter sigs(h->g); F(g,h) = chi_g nuf_h.
//   psi: M^T psi = (1/k) F^T psi, solved by the outer iteration in adjoint
//        mode with the transposed matrices (k_matrix, transpose != 0).
//   Gamma: (M^T - F^T / k) Gamma = S* with Gamma^T F phi = 0, S* = dJ/dphi
//        of a response J of the node powers (S*^T phi = 0 because the
//        powers are normalised), solved by the same adjoint outer with k
//        fixed, S* as the external source and the psi component projected
//        out after every outer.
//   Lambda = a psi + Gamma with a = (dJ/dk) k^2 / (psi^T F phi).
//   dJ/dp = -Lambda^T (dM/dp - dF/dp / k) phi + (direct dependence of the
//        powers on p).
// Parameters per node, columns of the gradient: D_g (G), sigma_a_g (G),
// nu-sigma_f_g (G; the power cross section sigma_f_g keeps its ratio to
// nu-sigma_f_g), then sigma_s(g->h) for g != h in (g, h) row-major order.
#pragma once

#include "adg_cmfd.h"

namespace adg {

// partial of psi^T F phi = sum_n (sum_g psi_g chi_g) (sum_h nuf_h phi_h)
__global__ void __launch_bounds__(NTF) k_fdot(LV lv, XS xs, const double *phi, const double *psi, double *partout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  const long NC = lv.NC;
  double acc[1] = {0.0};
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double fs = 0.0, pc = 0.0;
    for (int g = 0; g < lv.G; ++g) {
      fs = fs + xs.nuf[(size_t)g * NC + c] * phi[(size_t)g * NC + c];
      pc = pc + psi[(size_t)g * NC + c] * xs.chi[(size_t)g * NC + c];
    }
    acc[0] += pc * fs;
  }
  block_partial<1, 0u>(acc, partout, v.b.slot);
}

// generalized-adjoint source S*_g = dJ/dP * sigf_g V where the node's power
// term phi_g sigf_g V is positive (PowDis clips negative terms to 0)
__global__ void __launch_bounds__(NTF) k_sstar(LV lv, XS xs, const double *phi, const double *dJdP, double *sstar) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  const long NC = lv.NC;
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    for (int g = 0; g < lv.G; ++g) {
      const double w = phi[(size_t)g * NC + c] * xs.sigf[(size_t)g * NC + c] * xs.vdel[c];
      sstar[(size_t)g * NC + c] = w > 0.0 ? dJdP[c] * xs.sigf[(size_t)g * NC + c] * xs.vdel[c] : 0.0;
    }
  }
}

// after a generalized-adjoint outer: partials of |Gamma - Gamma_old|^2,
// |Gamma|^2 and Gamma^T F phi
__global__ void __launch_bounds__(NTF) k_gpt_norm(LV lv, XS xs, const double *gam, const double *gamc,
                                                 const double *phi, double *partout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  if (v.sc->done) return;
  const long NC = lv.NC;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double fs = 0.0;
    for (int g = 0; g < lv.G; ++g) fs = fs + xs.nuf[(size_t)g * NC + c] * phi[(size_t)g * NC + c];
    for (int g = 0; g < lv.G; ++g) {
      const double x = gam[(size_t)g * NC + c], d = x - gamc[(size_t)g * NC + c];
      acc[0] += d * d;
      acc[1] += x * x;
      acc[2] += x * xs.chi[(size_t)g * NC + c] * fs;
    }
  }
  block_partial<3, 0u>(acc, partout, v.b.slot);
}

// one block per core: projection coefficient Gamma^T F phi / psi^T F phi
// (psi^T F phi in aux; the coefficient kept in coef) and the relative change
// ||dGamma|| / ||Gamma||, used as both
// convergence measures of the outer's end test
__global__ void __launch_bounds__(NTF) k_gpt_fin(LV lv, const double *pin) {
  Scal *sc = lv.sc + blockIdx.x;
  if (sc->done) return;
  int first, n;
  core_slots(lv, blockIdx.x, first, n);
  double red[3];
  reduce_partials<3, 0u>(pin, first, n, red);
  if (threadIdx.x == 0) {
    const double err = red[1] > 0.0 ? sqrt(red[0] / red[1]) : 0.0;
    sc->coef = red[2] / sc->aux;
    sc->ser = err;
    sc->fer = err;
    sc->flags = 0;
  }
}

// Gamma -= coef psi, then the adjoint fission source fs0 = sum_g chi_g Gamma_g
__global__ void __launch_bounds__(NTF) k_gpt_proj(LV lv, XS xs, double *gam, const double *psi, double *fs0) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  if (v.sc->done) return;
  const double cf = v.sc->coef;
  const long NC = lv.NC;
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double fs = 0.0;
    for (int g = 0; g < lv.G; ++g) {
      const double x = gam[(size_t)g * NC + c] - cf * psi[(size_t)g * NC + c];
      gam[(size_t)g * NC + c] = x;
      fs = fs + x * xs.chi[(size_t)g * NC + c];
    }
    fs0[c] = fs;
  }
}

// d(df)/dD_n of one face of node n (coup_coef, src/mod_cmfd.f90:11-137):
// interior 2 D Dm / (D hm + Dm h); boundary code 0 2 a D / (2 D + a h) with
// a = 1e30 (written as 2 h / (2 D / a + h)^2), code 1 D / (2 D + h / 2), code 2 0
__device__ __forceinline__ double ddf_dD(bool nb, int bc, double D, double h, double Dm, double hm) {
#pragma clang fp contract(off)
  if (nb) {
    const double den = D * hm + Dm * h;
    return 2.0 * Dm * Dm * h / (den * den);
  }
  if (bc == 0) {
    const double den = 2.0 * D / 1.e30 + h;
    return 2.0 * h / (den * den);
  }
  if (bc == 1) {
    const double den = 2.0 * D + 0.5 * h;
    return 0.5 * h / (den * den);
  }
  return 0.0;
}

// Per-node gradients (layout [param][cell]).  Lambda = apsi * psi + gam
// (gam may be null); dJdP (may be null) carries the direct dependence of the
// response on sigma_f through the node powers.
__global__ void __launch_bounds__(NTF) k_grad(LV lv, XS xs, const double *phi, const double *psi, double apsi,
                                             const double *gam, const double *dJdP, double lambda, double *grad) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  const Geo &q = *v.q;
  const long NC = lv.NC;
  const int G = lv.G;
  const long st[3] = {1, v.b.sy, v.b.sz};
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    const unsigned char fl = q.flag[c];
    if (!(fl & F_ACTIVE)) continue;
    // (i, j, k) of the cell for the node sizes
    const long r0 = c - q.base;
    const int ii = (int)(r0 % q.NXP) - 1, jj = (int)((r0 / q.NXP) % q.NYP) - 1, kk = (int)(r0 / q.sz) - 1;
    const double hn[3] = {q.hx[ii], q.hy[jj], q.hz[kk]};
    double lam[MAXG], lchi = 0.0;
    for (int g = 0; g < G; ++g) {
      lam[g] = apsi * psi[(size_t)g * NC + c] + (gam ? gam[(size_t)g * NC + c] : 0.0);
      lchi = lchi + lam[g] * xs.chi[(size_t)g * NC + c];
    }
    int col = 0;
    // D_g: through the coupling coefficients of the node's six faces
    for (int g = 0; g < G; ++g) {
      const double *ph = phi + (size_t)g * NC, *ps = psi + (size_t)g * NC;
      const double D = xs.D[(size_t)g * NC + c];
      double s = 0.0;
      for (int f = 0; f < 6; ++f) {
        const int u = f >> 1;
        const bool nb = (fl & f_nb(f)) != 0;
        double dA;   // d(Lambda^T A phi) / d(df of this face)
        double Dm = 0.0, hm = 0.0;
        if (nb) {
          const long m = (f & 1) ? c - st[u] : c + st[u];
          const int im = u == 0 ? ii + ((f & 1) ? -1 : 1) : ii, jm = u == 1 ? jj + ((f & 1) ? -1 : 1) : jj,
                    km = u == 2 ? kk + ((f & 1) ? -1 : 1) : kk;
          hm = u == 0 ? q.hx[im] : (u == 1 ? q.hy[jm] : q.hz[km]);
          Dm = xs.D[(size_t)g * NC + m];
          const double lm = apsi * ps[m] + (gam ? gam[(size_t)g * NC + m] : 0.0);
          dA = lam[g] * (ph[c] - ph[m]) / hn[u] + lm * (ph[m] - ph[c]) / hm;
        } else {
          dA = lam[g] * ph[c] / hn[u];
        }
        s = s + dA * ddf_dD(nb, q.bc[f], D, hn[u], Dm, hm);
      }
      grad[(size_t)(col++) * NC + c] = -s;
    }
    // sigma_a_g: dM = e_g e_g^T at the node
    for (int g = 0; g < G; ++g) grad[(size_t)(col++) * NC + c] = -lam[g] * phi[(size_t)g * NC + c];
    // nu-sigma_f_g: dF(h, g) = chi_h; direct term through sigma_f_g = nuf_g * (sigf_g / nuf_g)
    for (int g = 0; g < G; ++g) {
      const double pg = phi[(size_t)g * NC + c];
      double val = lambda * lchi * pg;
      if (dJdP) {
        const double nf = xs.nuf[(size_t)g * NC + c], sf = xs.sigf[(size_t)g * NC + c];
        if (nf != 0.0 && pg * sf * xs.vdel[c] > 0.0) val = val + dJdP[c] * pg * xs.vdel[c] * (sf / nf);
      }
      grad[(size_t)(col++) * NC + c] = val;
    }
    // sigma_s(g -> h): removal from g, in-scatter to h
    for (int g = 0; g < G; ++g)
      for (int h = 0; h < G; ++h) {
        if (h == g) continue;
        grad[(size_t)(col++) * NC + c] = -(lam[g] - lam[h]) * phi[(size_t)g * NC + c];
      }
  }
}

}  // namespace adg
