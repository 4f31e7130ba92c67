// Transpose of one nodal D-hat update (src/mod_nodal.f90:18-147 with
// 282-371, 483-825 and 901-1451): for a cotangent w of the new D-hat, the
// cotangents of what the update reads -- the node fluxes, k, the old D-hat,
// the CMFD coupling coefficients df and the node cross sections -- at the
// state the update starts from (the forward static form, cmode 1).
//
// The forward passes are reversed in the opposite order:
//   pass 2 (per face)        ->  k_vjp_face: the D-hat epilogue, a3, the
//                                 transposed interface (2G x 2G) or boundary
//                                 (G x G) solve, then the cotangents of a2, a4,
//                                 L1, f0, B and A..H of the face's two nodes
//   pass 1 (per node)        ->  k_vjp_node: a4, the transposed G x G solve of
//                                 a2, L1 = h^2/(4D) TLUpd1, Lm2 = h^2/(4D) TLUpd2
//   TLUpd1/2 (node stencil)  ->  k_vjp_tl: the transverse-leakage fits are
//                                 linear in S along the line; each node gathers
//                                 its own and its two neighbours' terms
//   pass 0 (per node)        ->  k_vjp_src: S -> Lxyz -> the currents of the
//                                 node's six faces -> f0, old D-hat, df
//   assembly (flat)          ->  k_vjp_final: the sums over faces and
//                                 directions, df -> D, cross sections -> the
//                                 parameter columns of the gradient
// Every writer of a partial array owns its entries (a node is the "n" node
// of exactly one face and the "p" node of exactly one face per direction,
// and the face kernel writes them into separate slots), so the transpose
// needs no atomics and is deterministic.
//
// Cotangents are per face side, as the D-hat and df arrays are: the forward
// writes one value to both sides of an interior face, so its cotangent is
// the sum of w on the two sides; a partial with respect to one side's entry
// is that entry's own.  The SANM coefficients A..H are differentiated in
// forward mode through their argument alp = sqrt(sigr / D) h / 2.
#pragma once

#include "adg_adjoint.h"
#include "adg_nodal.h"

namespace adg {

// Cross-section cotangents of one node: D (G), sigma_r (G), nu-sigma_f (G),
// sigma_s (G x G, [to][from] as the engine stores it) and k.
constexpr int vjp_px(int G) { return 3 * G + G * G + 1; }

template <int G>
struct XB {
  double D[G], sr[G], nf[G], ss[G * G], k;
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int g = 0; g < G; ++g) { D[g] = 0.0; sr[g] = 0.0; nf[g] = 0.0; }
#pragma unroll
    for (int e = 0; e < G * G; ++e) ss[e] = 0.0;
    k = 0.0;
  }
  __device__ __forceinline__ void store(double *out, long NC, long c) const {
    int col = 0;
    for (int g = 0; g < G; ++g) out[(size_t)(col++) * NC + c] = D[g];
    for (int g = 0; g < G; ++g) out[(size_t)(col++) * NC + c] = sr[g];
    for (int g = 0; g < G; ++g) out[(size_t)(col++) * NC + c] = nf[g];
    for (int e = 0; e < G * G; ++e) out[(size_t)(col++) * NC + c] = ss[e];
    out[(size_t)col * NC + c] = k;
  }
  __device__ __forceinline__ void add(const double *in, long NC, long c) {
    int col = 0;
    for (int g = 0; g < G; ++g) D[g] += in[(size_t)(col++) * NC + c];
    for (int g = 0; g < G; ++g) sr[g] += in[(size_t)(col++) * NC + c];
    for (int g = 0; g < G; ++g) nf[g] += in[(size_t)(col++) * NC + c];
    for (int e = 0; e < G * G; ++e) ss[e] += in[(size_t)(col++) * NC + c];
    k += in[(size_t)col * NC + c];
  }
};

struct VjpArgs {
  NodArgs na;           // the forward state: xs, f0, df, dn (old D-hat), S, a2, a4, L1, kern
  const double *w;      // [6][G][NC] cotangent of the new D-hat
  // face pass: slot 0 = the node below the face ("n"), slot 1 = above ("p")
  double *a2b, *a4b, *l1b, *f0b2;   // [3][2][G][NC]
  double *xb2;                      // [3][2][PX][NC]
  double *dfb2;                     // [6][G][NC]
  // node pass
  double *sdir, *y1, *y2;           // [3][G][NC]: S direct, cotangents of TLUpd1 and TLUpd2
  double *f0b1, *xb1;               // [G][NC], [PX][NC]
  double *sb;                       // [3][G][NC] cotangent of S
  // source pass
  double *dnb, *dfb0, *f0b0;        // [6][G][NC], [6][G][NC], [G][NC]
  // assembly
  double *phib, *grad, *kb;         // [G][NC], [P][NC], [NC]
};

// forward-mode number for the SANM coefficient derivatives
struct Dl {
  double v, d;
};
__device__ __forceinline__ Dl operator+(Dl a, Dl b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dl operator-(Dl a, Dl b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dl operator*(Dl a, Dl b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dl operator/(Dl a, Dl b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Dl operator*(double s, Dl a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ Dl operator/(double s, Dl a) { return {s / a.v, -s * a.d / (a.v * a.v)}; }

// get_ABEFGH (src/mod_nodal.f90:1409-1451) and d(A..H)/d(alp); PNM: constants
__device__ __forceinline__ void nod_coef_d(int kern, double sigr, double D, double h, Coef6 &cv, Coef6 &cd,
                                           double &alpv) {
  if (kern != 2) {
    cv = nod_coef(kern, sigr, D, h);
    cd.A = cd.B = cd.E = cd.F = cd.G = cd.H = 0.0;
    alpv = 0.0;
    return;
  }
  alpv = 0.5 * sqrt(sigr / D) * h;
  const Dl alp{alpv, 1.0};
  const Dl alp2 = alp * alp;
  const Dl sh{sinh(alpv), cosh(alpv)}, ch{cosh(alpv), sinh(alpv)};
  const Dl m0c = sh / alp;
  const Dl m1s = 3.0 * (ch / alp - sh / alp2);
  const Dl m2c = 5.0 * (sh / alp - 3.0 * ch / alp2 + 3.0 * sh / (alp * alp * alp));
  const Dl A = (sh - m1s) / (alp2 * m1s);
  const Dl B = (ch - m0c - m2c) / (alp2 * m2c);
  const Dl E = m0c / m2c - 3.0 / alp2;
  const Dl F = (alp * ch - m1s) / (alp2 * m1s);
  const Dl Gc = (alp * sh - 3.0 * m2c) / (ch - m0c - m2c);
  const Dl H = (alp * ch - m1s) / (sh - m1s);
  cv = nod_coef(kern, sigr, D, h);   // the forward's own values
  cd.A = A.d; cd.B = B.d; cd.E = E.d; cd.F = F.d; cd.G = Gc.d; cd.H = H.d;
}

__device__ __forceinline__ void coef_zero(Coef6 &c) { c.A = c.B = c.E = c.F = c.G = c.H = 0.0; }

// cotangents of A..H of group g -> sigma_r and D through alp
template <int G>
__device__ __forceinline__ void xb_add_cf(XB<G> &x, int g, const Coef6 &bar, const Coef6 &der, double alp,
                                          double sigr, double D) {
  const double ab = bar.A * der.A + bar.B * der.B + bar.E * der.E + bar.F * der.F + bar.G * der.G + bar.H * der.H;
  if (ab == 0.0) return;
  x.sr[g] += ab * alp / (2.0 * sigr);
  x.D[g] += -ab * alp / (2.0 * D);
}

// cotangent of get_B's matrix of node c (B(g,h) = h^2 / (4 D_g) dum(g,h),
// forward form) -> D, sigma_r, sigma_s, nu-sigma_f and k
template <int G>
__device__ __forceinline__ void xb_add_B(XB<G> &x, const double *Bb, const XS &xs, long NC, long c, double h,
                                         double Ke) {
  for (int g = 0; g < G; ++g) {
    const double D = xs.D[(size_t)g * NC + c];
    const double fac = 0.25 * (h * h) / D;
    const double chi = xs.chi[(size_t)g * NC + c];
    for (int hh = 0; hh < G; ++hh) {
      const double bb = Bb[g * G + hh];
      if (bb == 0.0) continue;
      const double nf = xs.nuf[(size_t)hh * NC + c];
      const double dum = (g == hh ? xs.sigr[(size_t)g * NC + c] : -xs.sigs[((size_t)g * G + hh) * NC + c]) -
                         chi * nf / Ke;
      const double db = bb * fac;
      x.D[g] += -bb * fac * dum / D;
      if (g == hh) x.sr[g] += db;
      else x.ss[g * G + hh] += -db;
      x.nf[hh] += -db * chi / Ke;
      x.k += db * chi * nf / (Ke * Ke);
    }
  }
}

// Solve m^T x = b with the in-place Doolittle factors lu_solve left in m
// (L unit lower below the diagonal, U on and above it): U^T z = b, L^T x = z.
template <int M>
__device__ __forceinline__ void lu_solve_T(const double *m, double *b) {
  for (int i = 0; i < M; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s = s - m[k * M + i] * b[k];
    b[i] = s / m[i * M + i];
  }
  for (int i = M - 2; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < M; ++k) s = s - m[k * M + i] * b[k];
    b[i] = s;
  }
}

// TLUpd1 (which = 1) and TLUpd2 (which = 2) of a node as linear forms in
// S(m-1), S(m), S(m+1) along the line: the coefficients of tl1 and tl2
__device__ __forceinline__ void tl_coef(const Line &L, int bclo, int bchi, int which, double &wm, double &w0,
                                        double &wp) {
  wm = w0 = wp = 0.0;
  double tm, tp, p1m, p2m, p1p, p2p, hpp;
  if (L.first) {
    if (bclo == 2) {
      tm = 1.0; tp = L.hp / L.hn;
      p1m = tm + 1.0; p2m = 2.0 * tm + 1.0; p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      wp = which == 1 ? p1m * p2m / hpp : p1m / hpp;
      w0 = -wp;
      return;
    }
    if (which == 1) {
      tp = L.hp / L.hn; p1p = tp + 1.0;
      wp = 1.0 / p1p;
      w0 = -wp;
    }
    return;
  }
  if (L.last) {
    if (bchi == 2) {
      tm = L.hm / L.hn; tp = 1.0;
      p1m = tm + 1.0; p1p = tp + 1.0; p2p = 2.0 * tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      if (which == 1) { w0 = p1p * p2p / hpp; wm = -w0; }
      else { wm = p1p / hpp; w0 = -wm; }
      return;
    }
    if (which == 1) {
      tm = L.hm / L.hn; p1m = tm + 1.0;
      w0 = 1.0 / p1m;
      wm = -w0;
    }
    return;
  }
  tm = L.hm / L.hn; tp = L.hp / L.hn;
  p1m = tm + 1.0; p2m = 2.0 * tm + 1.0; p1p = tp + 1.0; p2p = 2.0 * tp + 1.0;
  hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
  if (which == 1) {
    wp = p1m * p2m / hpp;
    wm = -p1p * p2p / hpp;
    w0 = (p1p * p2p - p1m * p2m) / hpp;
  } else {
    wp = p1m / hpp;
    wm = p1p / hpp;
    w0 = -(p1m + p1p) / hpp;
  }
}

// ---------------------------------------------------------------------------
// reverse of pass 2: one thread per face (the face table, blockIdx.y = u)
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(NT) k_vjp_face(VjpArgs v) {
  const BlkView bv = blk_view(v.na.lv);
  const Geo &q = *bv.q;
  const NodArgs &a = v.na;
  const int u = blockIdx.y;
  const int nu = u == 0 ? q.nxx : (u == 1 ? q.nyy : q.nzz);
  const int e1 = u == 0 ? q.nyy : q.nxx;
  const int e2 = u == 2 ? q.nyy : q.nzz;
  const long nf = (long)(nu + 1) * e1 * e2;
  const long t = bv.b.lo + threadIdx.x;
  if (t >= nf) return;
  const long NC = q.NC;
  const double Ke = bv.sc->Ke;
  const int fi = (int)(t % (nu + 1));
  const long r = t / (nu + 1);
  const int o1 = (int)(r % e1), o2 = (int)(r / e1);
  int i, j, k;
  if (u == 0) { i = fi; j = o1; k = o2; }
  else if (u == 1) { i = o1; j = fi; k = o2; }
  else { i = o1; j = o2; k = fi; }
  const long st = u == 0 ? 1 : (u == 1 ? q.sy : q.sz);
  const long cp = cell_of(q, i, j, k);
  const long cn = cp - st;
  const bool hasp = (q.flag[cp] & F_ACTIVE) != 0;
  const bool hasn = (q.flag[cn] & F_ACTIVE) != 0;
  if (!hasp && !hasn) return;
  constexpr int PX = vjp_px(G);
  const int sf = 2 * u;
  const size_t ub = (size_t)u * G * NC;
  auto slot = [&](int s, int g) { return (((size_t)u * 2 + s) * G + g) * NC; };
  const XS &xs = a.xs;

  if (hasn && hasp) {
    const double hn = hdir(q, u, u == 0 ? i - 1 : i, u == 1 ? j - 1 : j, u == 2 ? k - 1 : k);
    const double hp = hdir(q, u, i, j, k);
    double Bn[G * G], Bp[G * G];
    nod_B<G>(a, NC, cn, hn, Ke, Bn);
    nod_B<G>(a, NC, cp, hp, Ke, Bp);
    Coef6 cfn[G], cfp[G], dn_[G], dp_[G];
    double alpn[G], alpp[G];
    for (int g = 0; g < G; ++g) {
      nod_coef_d(a.kern, xs.sigr[(size_t)g * NC + cn], xs.D[(size_t)g * NC + cn], hn, cfn[g], dn_[g], alpn[g]);
      nod_coef_d(a.kern, xs.sigr[(size_t)g * NC + cp], xs.D[(size_t)g * NC + cp], hp, cfp[g], dp_[g], alpp[g]);
    }
    constexpr int M2 = 2 * G;
    double R[M2 * M2], s[M2];
    for (int g = 0; g < G; ++g) {
      const double Ln1 = a.L1[ub + (size_t)g * NC + cn], Lp1 = a.L1[ub + (size_t)g * NC + cp];
      const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
      const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
      const double Pn = 2.0 * xs.D[(size_t)g * NC + cn] / hn;
      const double Pp = 2.0 * xs.D[(size_t)g * NC + cp] / hp;
      for (int hh = 0; hh < G; ++hh) {
        const double dl = hh == g ? 1.0 : 0.0;
        R[g * M2 + hh] = -Pn * (Bn[g * G + hh] * cfn[g].F + dl);
        R[g * M2 + hh + G] = Pp * (Bp[g * G + hh] * cfp[g].F + dl);
      }
      s[g] = Pn * (3.0 * a2n + cfn[g].G * a4n + cfn[g].F * Ln1) + Pp * (3.0 * a2p + cfp[g].G * a4p - cfp[g].F * Lp1);
    }
    for (int g = 0; g < G; ++g) {
      const double dcn = xs.dc[((size_t)sf * G + g) * NC + cn];
      const double dcp = xs.dc[((size_t)(sf + 1) * G + g) * NC + cp];
      const double Ln1 = a.L1[ub + (size_t)g * NC + cn], Lp1 = a.L1[ub + (size_t)g * NC + cp];
      const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
      const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
      for (int hh = 0; hh < G; ++hh) {
        const double dl = hh == g ? 1.0 : 0.0;
        R[(g + G) * M2 + hh] = dcn * (Bn[g * G + hh] * cfn[g].A + dl);
        R[(g + G) * M2 + hh + G] = dcp * (Bp[g * G + hh] * cfp[g].A + dl);
      }
      s[g + G] = dcp * (a2p + a4p + a.f0[(size_t)g * NC + cp] - cfn[g].A * Ln1) -
                 dcn * (a2n + a4n + a.f0[(size_t)g * NC + cn] + cfp[g].A * Lp1);
    }
    lu_solve<M2>(R, s);   // s: both halves of the solution; R: its factors
    double a1[G], a3[G], Ba1[G];
    for (int g = 0; g < G; ++g) a1[g] = s[g];
    for (int g = 0; g < G; ++g) {
      double Bf = 0.0;
      for (int hh = 0; hh < G; ++hh) Bf = Bf + Bn[g * G + hh] * a1[hh];
      Ba1[g] = Bf + a.L1[ub + (size_t)g * NC + cn];
      a3[g] = cfn[g].A * Ba1[g];
    }
    // reverse
    XB<G> xn, xp;
    xn.zero(); xp.zero();
    Coef6 bn[G], bp[G];
    double Bnb[G * G], Bpb[G * G];
    double a2nb[G], a4nb[G], l1nb[G], f0nb[G], a2pb[G], a4pb[G], l1pb[G], f0pb[G], a1b[G], a3b[G];
    for (int g = 0; g < G; ++g) {
      coef_zero(bn[g]); coef_zero(bp[g]);
      a2nb[g] = a4nb[g] = l1nb[g] = f0nb[g] = a2pb[g] = a4pb[g] = l1pb[g] = f0pb[g] = a1b[g] = a3b[g] = 0.0;
    }
    for (int e = 0; e < G * G; ++e) Bnb[e] = Bpb[e] = 0.0;
    for (int g = 0; g < G; ++g) {
      const size_t on = ((size_t)sf * G + g) * NC + cn, op = ((size_t)(sf + 1) * G + g) * NC + cp;
      const double wv = v.w[on] + v.w[op];
      const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
      const double Dn = xs.D[(size_t)g * NC + cn];
      const double Q = a1[g] + 3.0 * a2n + cfn[g].H * a3[g] + cfn[g].G * a4n;
      const double jp = -2.0 * Dn / hn * Q;
      const double fn = a.f0[(size_t)g * NC + cn], fp = a.f0[(size_t)g * NC + cp];
      const double dfv = a.df[on];
      const double den = fn + fp;
      const double val = (dfv * (fn - fp) - jp) / den;
      const double numb = wv / den, denb = -wv * val / den;
      v.dfb2[on] = numb * (fn - fp);
      f0nb[g] += numb * dfv + denb;
      f0pb[g] += -numb * dfv + denb;
      const double jpb = -numb;
      xn.D[g] += jpb * Q * (-2.0 / hn);
      const double Qb = jpb * (-2.0 * Dn / hn);
      a1b[g] += Qb;
      a2nb[g] += 3.0 * Qb;
      bn[g].H += Qb * a3[g];
      a3b[g] += Qb * cfn[g].H;
      bn[g].G += Qb * a4n;
      a4nb[g] += Qb * cfn[g].G;
    }
    for (int g = 0; g < G; ++g) {
      bn[g].A += a3b[g] * Ba1[g];
      const double tb = a3b[g] * cfn[g].A;
      for (int hh = 0; hh < G; ++hh) {
        Bnb[g * G + hh] += tb * a1[hh];
        a1b[hh] += tb * Bn[g * G + hh];
      }
      l1nb[g] += tb;
    }
    double sb[M2];
    for (int e = 0; e < M2; ++e) sb[e] = e < G ? a1b[e] : 0.0;
    lu_solve_T<M2>(R, sb);
    for (int g = 0; g < G; ++g) {
      const double Dn = xs.D[(size_t)g * NC + cn], Dp = xs.D[(size_t)g * NC + cp];
      const double Pn = 2.0 * Dn / hn, Pp = 2.0 * Dp / hp;
      const double dcn = xs.dc[((size_t)sf * G + g) * NC + cn];
      const double dcp = xs.dc[((size_t)(sf + 1) * G + g) * NC + cp];
      const double Ln1 = a.L1[ub + (size_t)g * NC + cn], Lp1 = a.L1[ub + (size_t)g * NC + cp];
      const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
      const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
      double Pnb = 0.0, Ppb = 0.0;
      for (int hh = 0; hh < G; ++hh) {
        const double dl = hh == g ? 1.0 : 0.0;
        const double r1 = -sb[g] * s[hh], r2 = -sb[g] * s[G + hh];
        const double r3 = -sb[G + g] * s[hh], r4 = -sb[G + g] * s[G + hh];
        Pnb += r1 * -(Bn[g * G + hh] * cfn[g].F + dl);
        Bnb[g * G + hh] += r1 * (-Pn * cfn[g].F) + r3 * dcn * cfn[g].A;
        bn[g].F += r1 * (-Pn * Bn[g * G + hh]);
        Ppb += r2 * (Bp[g * G + hh] * cfp[g].F + dl);
        Bpb[g * G + hh] += r2 * Pp * cfp[g].F + r4 * dcp * cfp[g].A;
        bp[g].F += r2 * Pp * Bp[g * G + hh];
        bn[g].A += r3 * dcn * Bn[g * G + hh];
        bp[g].A += r4 * dcp * Bp[g * G + hh];
      }
      const double sg = sb[g];
      Pnb += sg * (3.0 * a2n + cfn[g].G * a4n + cfn[g].F * Ln1);
      a2nb[g] += 3.0 * sg * Pn;
      bn[g].G += sg * Pn * a4n;
      a4nb[g] += sg * Pn * cfn[g].G;
      bn[g].F += sg * Pn * Ln1;
      l1nb[g] += sg * Pn * cfn[g].F;
      Ppb += sg * (3.0 * a2p + cfp[g].G * a4p - cfp[g].F * Lp1);
      a2pb[g] += 3.0 * sg * Pp;
      bp[g].G += sg * Pp * a4p;
      a4pb[g] += sg * Pp * cfp[g].G;
      bp[g].F += -sg * Pp * Lp1;
      l1pb[g] += -sg * Pp * cfp[g].F;
      const double s2 = sb[G + g];
      a2pb[g] += s2 * dcp;
      a4pb[g] += s2 * dcp;
      f0pb[g] += s2 * dcp;
      bn[g].A += -s2 * dcp * Ln1;
      l1nb[g] += -s2 * dcp * cfn[g].A;
      a2nb[g] += -s2 * dcn;
      a4nb[g] += -s2 * dcn;
      f0nb[g] += -s2 * dcn;
      bp[g].A += -s2 * dcn * Lp1;
      l1pb[g] += -s2 * dcn * cfp[g].A;
      xn.D[g] += Pnb * 2.0 / hn;
      xp.D[g] += Ppb * 2.0 / hp;
    }
    xb_add_B<G>(xn, Bnb, xs, NC, cn, hn, Ke);
    xb_add_B<G>(xp, Bpb, xs, NC, cp, hp, Ke);
    for (int g = 0; g < G; ++g) {
      xb_add_cf<G>(xn, g, bn[g], dn_[g], alpn[g], xs.sigr[(size_t)g * NC + cn], xs.D[(size_t)g * NC + cn]);
      xb_add_cf<G>(xp, g, bp[g], dp_[g], alpp[g], xs.sigr[(size_t)g * NC + cp], xs.D[(size_t)g * NC + cp]);
    }
    for (int g = 0; g < G; ++g) {
      v.a2b[slot(0, g) + cn] = a2nb[g]; v.a4b[slot(0, g) + cn] = a4nb[g];
      v.l1b[slot(0, g) + cn] = l1nb[g]; v.f0b2[slot(0, g) + cn] = f0nb[g];
      v.a2b[slot(1, g) + cp] = a2pb[g]; v.a4b[slot(1, g) + cp] = a4pb[g];
      v.l1b[slot(1, g) + cp] = l1pb[g]; v.f0b2[slot(1, g) + cp] = f0pb[g];
    }
    xn.store(v.xb2 + ((size_t)u * 2 + 0) * PX * NC, NC, cn);
    xp.store(v.xb2 + ((size_t)u * 2 + 1) * PX * NC, NC, cp);
    return;
  }

  // boundary face: the first (hasp) or last (hasn) node of a line
  const bool first = hasp;
  const long c = first ? cp : cn;
  const int ii = first ? i : (u == 0 ? i - 1 : i), jj = first ? j : (u == 1 ? j - 1 : j),
            kk = first ? k : (u == 2 ? k - 1 : k);
  const double h = hdir(q, u, ii, jj, kk);
  const int bc = first ? q.bc[2 * u + 1] : q.bc[2 * u];
  const int side = first ? sf + 1 : sf;   // the node's own side of this face
  double Bm[G * G];
  nod_B<G>(a, NC, c, h, Ke, Bm);
  Coef6 cf[G], dcf[G];
  double alp[G];
  for (int g = 0; g < G; ++g)
    nod_coef_d(a.kern, xs.sigr[(size_t)g * NC + c], xs.D[(size_t)g * NC + c], h, cf[g], dcf[g], alp[g]);
  double Mx[G * G], b[G];
  for (int g = 0; g < G; ++g) {
    const double L1v = a.L1[ub + (size_t)g * NC + c];
    const double a2v = a.a2[ub + (size_t)g * NC + c], a4v = a.a4[ub + (size_t)g * NC + c];
    const double P = 2.0 * xs.D[(size_t)g * NC + c] / h;
    const double dc = xs.dc[((size_t)side * G + g) * NC + c];
    const double A = cf[g].A, F = cf[g].F, Gg = cf[g].G, H = cf[g].H;
    const double f0v = a.f0[(size_t)g * NC + c];
    for (int hh = 0; hh < G; ++hh) {
      const double Bv = Bm[g * G + hh];
      const double dl = hh == g ? 1.0 : 0.0;
      if (first) {
        if (bc == 2) Mx[g * G + hh] = P * (Bv * F + dl);
        else if (bc == 1) Mx[g * G + hh] = -dc * (dl + A * Bv) - 2.0 * P * (A * Bv * H + dl);
        else Mx[g * G + hh] = dc * (dl + A * Bv);
      } else {
        if (bc == 2) Mx[g * G + hh] = -P * (Bv * F + dl);
        else if (bc == 1) Mx[g * G + hh] = dc * (dl + A * Bv) + 2.0 * P * (A * Bv * H + dl);
        else Mx[g * G + hh] = dc * (dl + A * Bv);
      }
    }
    if (first) {
      if (bc == 2) b[g] = P * (3.0 * a2v + Gg * a4v - F * L1v);
      else if (bc == 1) b[g] = 2.0 * P * (A * H * L1v - 3.0 * a2v - Gg * a4v) - dc * (a2v + a4v + f0v - A * L1v);
      else b[g] = dc * (a2v + a4v + f0v - A * L1v);
    } else {
      if (bc == 2) b[g] = P * (3.0 * a2v + Gg * a4v + F * L1v);
      else if (bc == 1) b[g] = -2.0 * P * (A * H * L1v + 3.0 * a2v + Gg * a4v) - dc * (a2v + a4v + f0v + A * L1v);
      else b[g] = -dc * (a2v + a4v + f0v + A * L1v);
    }
  }
  lu_solve<G>(Mx, b);
  double a1[G], a3[G], Ba1[G];
  for (int g = 0; g < G; ++g) a1[g] = b[g];
  for (int g = 0; g < G; ++g) {
    double Bf = 0.0;
    for (int hh = 0; hh < G; ++hh) Bf = Bf + Bm[g * G + hh] * a1[hh];
    Ba1[g] = Bf + a.L1[ub + (size_t)g * NC + c];
    a3[g] = cf[g].A * Ba1[g];
  }
  XB<G> x;
  x.zero();
  Coef6 cb[G];
  double Bb[G * G], a2b[G], a4b[G], l1b[G], f0b[G], a1b[G], a3b[G];
  for (int g = 0; g < G; ++g) {
    coef_zero(cb[g]);
    a2b[g] = a4b[g] = l1b[g] = f0b[g] = a1b[g] = a3b[g] = 0.0;
  }
  for (int e = 0; e < G * G; ++e) Bb[e] = 0.0;
  const double sgn = first ? -1.0 : 1.0;   // jp's a2 and a4 terms: - for the first node, + for the last
  for (int g = 0; g < G; ++g) {
    const size_t o = ((size_t)side * G + g) * NC + c;
    const double wv = v.w[o];
    const double a2v = a.a2[ub + (size_t)g * NC + c], a4v = a.a4[ub + (size_t)g * NC + c];
    const double Dv = xs.D[(size_t)g * NC + c];
    const double Q = a1[g] + sgn * 3.0 * a2v + cf[g].H * a3[g] + sgn * cf[g].G * a4v;
    const double jp = -2.0 * Dv / h * Q;
    const double f0v = a.f0[(size_t)g * NC + c];
    // first: v = -(jp / f0 + df); last: v = -(jp / f0 - df)
    const double jpb = -wv / f0v;
    f0b[g] += wv * jp / (f0v * f0v);
    v.dfb2[o] = first ? -wv : wv;
    x.D[g] += jpb * Q * (-2.0 / h);
    const double Qb = jpb * (-2.0 * Dv / h);
    a1b[g] += Qb;
    a2b[g] += sgn * 3.0 * Qb;
    cb[g].H += Qb * a3[g];
    a3b[g] += Qb * cf[g].H;
    cb[g].G += sgn * Qb * a4v;
    a4b[g] += sgn * Qb * cf[g].G;
  }
  for (int g = 0; g < G; ++g) {
    cb[g].A += a3b[g] * Ba1[g];
    const double tb = a3b[g] * cf[g].A;
    for (int hh = 0; hh < G; ++hh) {
      Bb[g * G + hh] += tb * a1[hh];
      a1b[hh] += tb * Bm[g * G + hh];
    }
    l1b[g] += tb;
  }
  double bb[G];
  for (int g = 0; g < G; ++g) bb[g] = a1b[g];
  lu_solve_T<G>(Mx, bb);
  for (int g = 0; g < G; ++g) {
    const double L1v = a.L1[ub + (size_t)g * NC + c];
    const double a2v = a.a2[ub + (size_t)g * NC + c], a4v = a.a4[ub + (size_t)g * NC + c];
    const double P = 2.0 * xs.D[(size_t)g * NC + c] / h;
    const double dc = xs.dc[((size_t)side * G + g) * NC + c];
    const double A = cf[g].A, F = cf[g].F, Gg = cf[g].G, H = cf[g].H;
    const double f0v = a.f0[(size_t)g * NC + c];
    double Pb = 0.0;
    for (int hh = 0; hh < G; ++hh) {
      const double Bv = Bm[g * G + hh];
      const double dl = hh == g ? 1.0 : 0.0;
      const double mb = -bb[g] * a1[hh];
      if (bc == 2) {
        const double sp = first ? 1.0 : -1.0;   // Mx = +-P (B F + delta)
        Pb += mb * sp * (Bv * F + dl);
        Bb[g * G + hh] += mb * sp * P * F;
        cb[g].F += mb * sp * P * Bv;
      } else if (bc == 1) {
        const double sp = first ? -1.0 : 1.0;   // Mx = +-(dc (delta + A B) + 2 P (A B H + delta))
        Bb[g * G + hh] += mb * sp * (dc * A + 2.0 * P * A * H);
        cb[g].A += mb * sp * (dc * Bv + 2.0 * P * Bv * H);
        cb[g].H += mb * sp * 2.0 * P * A * Bv;
        Pb += mb * sp * 2.0 * (A * Bv * H + dl);
      } else {
        Bb[g * G + hh] += mb * dc * A;
        cb[g].A += mb * dc * Bv;
      }
    }
    const double r = bb[g];
    if (bc == 2) {
      // first: P (3 a2 + G a4 - F L1); last: P (3 a2 + G a4 + F L1)
      const double sl = first ? -1.0 : 1.0;
      Pb += r * (3.0 * a2v + Gg * a4v + sl * F * L1v);
      a2b[g] += 3.0 * r * P;
      cb[g].G += r * P * a4v;
      a4b[g] += r * P * Gg;
      cb[g].F += sl * r * P * L1v;
      l1b[g] += sl * r * P * F;
    } else if (bc == 1) {
      if (first) {
        // 2 P (A H L1 - 3 a2 - G a4) - dc (a2 + a4 + f0 - A L1)
        Pb += 2.0 * r * (A * H * L1v - 3.0 * a2v - Gg * a4v);
        cb[g].A += 2.0 * r * P * H * L1v + r * dc * L1v;
        cb[g].H += 2.0 * r * P * A * L1v;
        l1b[g] += 2.0 * r * P * A * H + r * dc * A;
        a2b[g] += -6.0 * r * P - r * dc;
        cb[g].G += -2.0 * r * P * a4v;
        a4b[g] += -2.0 * r * P * Gg - r * dc;
        f0b[g] += -r * dc;
      } else {
        // -2 P (A H L1 + 3 a2 + G a4) - dc (a2 + a4 + f0 + A L1)
        Pb += -2.0 * r * (A * H * L1v + 3.0 * a2v + Gg * a4v);
        cb[g].A += -2.0 * r * P * H * L1v - r * dc * L1v;
        cb[g].H += -2.0 * r * P * A * L1v;
        l1b[g] += -2.0 * r * P * A * H - r * dc * A;
        a2b[g] += -6.0 * r * P - r * dc;
        cb[g].G += -2.0 * r * P * a4v;
        a4b[g] += -2.0 * r * P * Gg - r * dc;
        f0b[g] += -r * dc;
      }
    } else {
      // first: dc (a2 + a4 + f0 - A L1); last: -dc (a2 + a4 + f0 + A L1)
      const double sd = first ? 1.0 : -1.0;
      a2b[g] += sd * r * dc;
      a4b[g] += sd * r * dc;
      f0b[g] += sd * r * dc;
      cb[g].A += -r * dc * L1v;
      l1b[g] += -r * dc * A;
    }
    (void)f0v;
    x.D[g] += Pb * 2.0 / h;
  }
  xb_add_B<G>(x, Bb, xs, NC, c, h, Ke);
  for (int g = 0; g < G; ++g)
    xb_add_cf<G>(x, g, cb[g], dcf[g], alp[g], xs.sigr[(size_t)g * NC + c], xs.D[(size_t)g * NC + c]);
  const int sl = first ? 1 : 0;
  for (int g = 0; g < G; ++g) {
    v.a2b[slot(sl, g) + c] = a2b[g];
    v.a4b[slot(sl, g) + c] = a4b[g];
    v.l1b[slot(sl, g) + c] = l1b[g];
    v.f0b2[slot(sl, g) + c] = f0b[g];
  }
  x.store(v.xb2 + ((size_t)u * 2 + sl) * PX * NC, NC, c);
}

// ---------------------------------------------------------------------------
// reverse of pass 1: one thread per node, the three directions in turn
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(NT) k_vjp_node(VjpArgs v) {
  const BlkView bv = blk_view(v.na.lv);
  const Geo &q = *bv.q;
  const NodArgs &a = v.na;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = bv.b.lo + threadIdx.x;
  if (t >= bv.b.hi || t >= nxy * q.nzz) return;
  int i, j, k;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  const long c = cell_of(q, i, j, k);
  if (!(q.flag[c] & F_ACTIVE)) return;
  const long NC = q.NC;
  const double Ke = bv.sc->Ke;
  const XS &xs = a.xs;
  XB<G> x;
  x.zero();
  double f0b[G];
  for (int g = 0; g < G; ++g) f0b[g] = 0.0;
  for (int u = 0; u < 3; ++u) {
    const Line L = line_at(q, u, i, j, k, c);
    if (L.first && L.last) continue;
    const int bclo = q.bc[2 * u + 1], bchi = q.bc[2 * u];
    const double h = L.hn;
    const double *Su = a.S + (size_t)u * G * NC;
    double B[G * G], M[G * G], rhs[G], fac[G], t1[G], t2[G], Lm2[G];
    Coef6 cf[G], dcf[G];
    double alp[G];
    nod_B<G>(a, NC, c, h, Ke, B);
    for (int g = 0; g < G; ++g) {
      nod_coef_d(a.kern, xs.sigr[(size_t)g * NC + c], xs.D[(size_t)g * NC + c], h, cf[g], dcf[g], alp[g]);
      const double D = xs.D[(size_t)g * NC + c];
      fac[g] = 0.25 * (h * h) / D;
      t1[g] = tl1(L, Su + (size_t)g * NC, c, bclo, bchi);
      t2[g] = tl2(L, Su + (size_t)g * NC, c, bclo, bchi);
      Lm2[g] = fac[g] * t2[g];
      const double S = fac[g] * (Su[(size_t)g * NC + c] - xs.exsrc[(size_t)g * NC + c]);
      double Bf = 0.0;
      for (int hh = 0; hh < G; ++hh) {
        M[g * G + hh] = (hh == g) ? B[g * G + hh] * cf[g].E + 3.0 : B[g * G + hh] * cf[g].E;
        Bf = Bf + B[g * G + hh] * a.f0[(size_t)hh * NC + c];
      }
      rhs[g] = Bf - cf[g].E * Lm2[g] + S;
    }
    lu_solve<G>(M, rhs);   // rhs = a2; M = its factors
    double a2bt[G], a4bt[G], l1bt[G];
    for (int g = 0; g < G; ++g) {
      const size_t s0 = (((size_t)u * 2 + 0) * G + g) * NC + c, s1 = (((size_t)u * 2 + 1) * G + g) * NC + c;
      a2bt[g] = v.a2b[s0] + v.a2b[s1];
      a4bt[g] = v.a4b[s0] + v.a4b[s1];
      l1bt[g] = v.l1b[s0] + v.l1b[s1];
    }
    Coef6 cb[G];
    double Bb[G * G], facb[G], Lm2b[G], y1[G];
    for (int g = 0; g < G; ++g) {
      coef_zero(cb[g]);
      facb[g] = l1bt[g] * t1[g];   // L1 = fac tl1
      y1[g] = l1bt[g] * fac[g];
      Lm2b[g] = 0.0;
    }
    for (int e = 0; e < G * G; ++e) Bb[e] = 0.0;
    // a4 = B_coef (B a2 + Lm2)
    for (int g = 0; g < G; ++g) {
      double Ba2 = 0.0;
      for (int hh = 0; hh < G; ++hh) Ba2 = Ba2 + B[g * G + hh] * rhs[hh];
      cb[g].B += a4bt[g] * (Ba2 + Lm2[g]);
      const double tb = a4bt[g] * cf[g].B;
      for (int hh = 0; hh < G; ++hh) {
        Bb[g * G + hh] += tb * rhs[hh];
        a2bt[hh] += tb * B[g * G + hh];
      }
      Lm2b[g] += tb;
    }
    // a2 = M^-1 rhs
    double rb[G];
    for (int g = 0; g < G; ++g) rb[g] = a2bt[g];
    lu_solve_T<G>(M, rb);
    for (int g = 0; g < G; ++g) {
      const double Sraw = Su[(size_t)g * NC + c] - xs.exsrc[(size_t)g * NC + c];
      for (int hh = 0; hh < G; ++hh) {
        const double mb = -rb[g] * rhs[hh];
        Bb[g * G + hh] += mb * cf[g].E + rb[g] * a.f0[(size_t)hh * NC + c];
        cb[g].E += mb * B[g * G + hh];
        f0b[hh] += rb[g] * B[g * G + hh];
      }
      cb[g].E += -rb[g] * Lm2[g];
      Lm2b[g] += -rb[g] * cf[g].E;
      facb[g] += rb[g] * Sraw;
      const double sd = rb[g] * fac[g];
      facb[g] += Lm2b[g] * t2[g];
      const double y2 = Lm2b[g] * fac[g];
      const double D = xs.D[(size_t)g * NC + c];
      x.D[g] += facb[g] * (-fac[g] / D);
      const size_t o = ((size_t)u * G + g) * NC + c;
      v.sdir[o] = sd;
      v.y1[o] = y1[g];
      v.y2[o] = y2;
    }
    xb_add_B<G>(x, Bb, xs, NC, c, h, Ke);
    for (int g = 0; g < G; ++g)
      xb_add_cf<G>(x, g, cb[g], dcf[g], alp[g], xs.sigr[(size_t)g * NC + c], xs.D[(size_t)g * NC + c]);
  }
  for (int g = 0; g < G; ++g) v.f0b1[(size_t)g * NC + c] = f0b[g];
  x.store(v.xb1, NC, c);
}

// ---------------------------------------------------------------------------
// reverse of TLUpd1/2: S cotangent of node c along each direction gathers
// its own and its two line neighbours' leakage-fit terms; blockIdx.y = group
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(NT) k_vjp_tl(VjpArgs v) {
  const BlkView bv = blk_view(v.na.lv);
  const Geo &q = *bv.q;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = bv.b.lo + threadIdx.x;
  if (t >= bv.b.hi || t >= nxy * q.nzz) return;
  int i, j, k;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  const long c = cell_of(q, i, j, k);
  if (!(q.flag[c] & F_ACTIVE)) return;
  const long NC = q.NC;
  const int G = q.G, g = blockIdx.y;
  for (int u = 0; u < 3; ++u) {
    const int bclo = q.bc[2 * u + 1], bchi = q.bc[2 * u];
    const long st = u == 0 ? 1 : (u == 1 ? q.sy : q.sz);
    const size_t ob = ((size_t)u * G + g) * NC;
    const Line L = line_at(q, u, i, j, k, c);
    double acc = v.sdir[ob + c];
    double wm, w0, wp;
    if (!(L.first && L.last)) {
      tl_coef(L, bclo, bchi, 1, wm, w0, wp);
      acc += w0 * v.y1[ob + c];
      tl_coef(L, bclo, bchi, 2, wm, w0, wp);
      acc += w0 * v.y2[ob + c];
    }
    if (!L.first) {   // the lower neighbour's fits read S(c) as their S(m+1)
      const long m = c - st;
      const Line Lm = line_at(q, u, u == 0 ? i - 1 : i, u == 1 ? j - 1 : j, u == 2 ? k - 1 : k, m);
      if (!(Lm.first && Lm.last)) {
        tl_coef(Lm, bclo, bchi, 1, wm, w0, wp);
        acc += wp * v.y1[ob + m];
        tl_coef(Lm, bclo, bchi, 2, wm, w0, wp);
        acc += wp * v.y2[ob + m];
      }
    }
    if (!L.last) {    // the upper neighbour's fits read S(c) as their S(m-1)
      const long m = c + st;
      const Line Lp = line_at(q, u, u == 0 ? i + 1 : i, u == 1 ? j + 1 : j, u == 2 ? k + 1 : k, m);
      if (!(Lp.first && Lp.last)) {
        tl_coef(Lp, bclo, bchi, 1, wm, w0, wp);
        acc += wm * v.y1[ob + m];
        tl_coef(Lp, bclo, bchi, 2, wm, w0, wp);
        acc += wm * v.y2[ob + m];
      }
    }
    v.sb[ob + c] = acc;
  }
}

// ---------------------------------------------------------------------------
// reverse of pass 0 (get_source + Lxyz, forward form): one thread per
// (node, group); blockIdx.y = group
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lbar(const double *sb, int G, long NC, int g, long c, int v) {
  double s = 0.0;
  for (int u = 0; u < 3; ++u)
    if (u != v) s += sb[((size_t)u * G + g) * NC + c];
  return s;
}

__global__ void __launch_bounds__(NT) k_vjp_src(VjpArgs v) {
  const BlkView bv = blk_view(v.na.lv);
  const Geo &q = *bv.q;
  const NodArgs &a = v.na;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = bv.b.lo + threadIdx.x;
  if (t >= bv.b.hi || t >= nxy * q.nzz) return;
  int i, j, k;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  const long c = cell_of(q, i, j, k);
  const unsigned char fl = q.flag[c];
  if (!(fl & F_ACTIVE)) return;
  const long NC = q.NC;
  const int G = q.G, g = blockIdx.y;
  const double *f = a.f0 + (size_t)g * NC;
  const double fn = f[c];
  const long st[3] = {1, q.sy, q.sz};
  const double hc[3] = {q.hx[i], q.hy[j], q.hz[k]};
  double f0b = 0.0;
  for (int u = 0; u < 3; ++u) {
    const int sp = 2 * u, sm = 2 * u + 1;
    const double Lb = lbar(v.sb, G, NC, g, c, u);
    const double jpb = Lb / hc[u], jmb = -Lb / hc[u];
    const size_t op = ((size_t)sp * G + g) * NC + c, om = ((size_t)sm * G + g) * NC + c;
    const double dfp = a.df[op], dnp = a.dn[op], dfm = a.df[om], dnm = a.dn[om];
    double dfbp = 0.0, dnbp = 0.0, dfbm = 0.0, dnbm = 0.0;
    if (fl & f_nb(sp)) {
      const long m = c + st[u];
      const double fp = f[m];
      dfbp = jpb * -(fp - fn);
      dnbp = jpb * -(fp + fn);
      f0b += jpb * (dfp - dnp);
      // the upper neighbour's lower-face current reads f(c) as its fm
      const double hm = u == 0 ? q.hx[i + 1] : (u == 1 ? q.hy[j + 1] : q.hz[k + 1]);
      const size_t mo = ((size_t)sm * G + g) * NC + m;
      f0b += lbar(v.sb, G, NC, g, m, u) * -(a.df[mo] - a.dn[mo]) / hm;
    } else if (q.bc[sp] != 2) {
      dfbp = jpb * fn;
      dnbp = -jpb * fn;
      f0b += jpb * (dfp - dnp);
    }
    if (fl & f_nb(sm)) {
      const long m = c - st[u];
      const double fm = f[m];
      dfbm = jmb * -(fn - fm);
      dnbm = jmb * -(fn + fm);
      f0b += jmb * (-dfm - dnm);
      // the lower neighbour's upper-face current reads f(c) as its fp
      const double hm = u == 0 ? q.hx[i - 1] : (u == 1 ? q.hy[j - 1] : q.hz[k - 1]);
      const size_t mo = ((size_t)sp * G + g) * NC + m;
      f0b += lbar(v.sb, G, NC, g, m, u) * (-a.df[mo] - a.dn[mo]) / hm;
    } else if (q.bc[sm] != 2) {
      dfbm = -jmb * fn;
      dnbm = -jmb * fn;
      f0b += jmb * (-dfm - dnm);
    }
    v.dfb0[op] = dfbp; v.dnb[op] = dnbp;
    v.dfb0[om] = dfbm; v.dnb[om] = dnbm;
  }
  v.f0b0[(size_t)g * NC + c] = f0b;
}

// ---------------------------------------------------------------------------
// assembly (flat launch over the core's cells): flux cotangent, parameter
// columns (D_g, sigma_a_g, nu-sigma_f_g, sigma_s(g->h) for h != g, as the
// gradient of the CMFD adjoint) and the per-cell k share
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(NTF) k_vjp_final(VjpArgs v, LV lv) {
  const BlkView bv = blk_view(lv);
  const Geo &q = *bv.q;
  const XS &xs = v.na.xs;
  const long NC = lv.NC;
  constexpr int PX = vjp_px(G);
  const long st[3] = {1, bv.b.sy, bv.b.sz};
  for (long c = bv.b.lo + threadIdx.x; c < bv.b.hi; c += NTF) {
    const unsigned char fl = q.flag[c];
    if (!(fl & F_ACTIVE)) continue;
    const long r0 = c - q.base;
    const int ii = (int)(r0 % q.NXP) - 1, jj = (int)((r0 / q.NXP) % q.NYP) - 1, kk = (int)(r0 / q.sz) - 1;
    const double hn[3] = {q.hx[ii], q.hy[jj], q.hz[kk]};
    XB<G> x;
    x.zero();
    x.add(v.xb1, NC, c);
    for (int u = 0; u < 3; ++u)
      for (int s = 0; s < 2; ++s) x.add(v.xb2 + ((size_t)u * 2 + s) * PX * NC, NC, c);
    for (int g = 0; g < G; ++g) {
      double fb = v.f0b0[(size_t)g * NC + c] + v.f0b1[(size_t)g * NC + c];
      for (int u = 0; u < 3; ++u)
        for (int s = 0; s < 2; ++s) fb += v.f0b2[(((size_t)u * 2 + s) * G + g) * NC + c];
      v.phib[(size_t)g * NC + c] = fb;
      // df of the node's six faces -> D (coup_coef; the two sides of an
      // interior face hold one value, so both sides' cotangents count)
      const double D = xs.D[(size_t)g * NC + c];
      for (int f = 0; f < 6; ++f) {
        const int u = f >> 1;
        const bool nb = (fl & f_nb(f)) != 0;
        const size_t o = ((size_t)f * G + g) * NC + c;
        double dft = v.dfb0[o] + v.dfb2[o];
        double Dm = 0.0, hm = 0.0;
        if (nb) {
          const long m = (f & 1) ? c - st[u] : c + st[u];
          const size_t om = ((size_t)(f ^ 1) * G + g) * NC + m;
          dft += v.dfb0[om] + v.dfb2[om];
          const int im = u == 0 ? ii + ((f & 1) ? -1 : 1) : ii, jm = u == 1 ? jj + ((f & 1) ? -1 : 1) : jj,
                    km = u == 2 ? kk + ((f & 1) ? -1 : 1) : kk;
          hm = u == 0 ? q.hx[im] : (u == 1 ? q.hy[jm] : q.hz[km]);
          Dm = xs.D[(size_t)g * NC + m];
        }
        x.D[g] += dft * ddf_dD(nb, q.bc[f], D, hn[u], Dm, hm);
      }
    }
    int col = 0;
    for (int g = 0; g < G; ++g) v.grad[(size_t)(col++) * NC + c] = x.D[g];
    for (int g = 0; g < G; ++g) v.grad[(size_t)(col++) * NC + c] = x.sr[g];
    for (int g = 0; g < G; ++g) v.grad[(size_t)(col++) * NC + c] = x.nf[g];
    for (int g = 0; g < G; ++g)
      for (int h = 0; h < G; ++h) {
        if (h == g) continue;
        v.grad[(size_t)(col++) * NC + c] = x.ss[h * G + g] + x.sr[g];
      }
    v.kb[c] = x.k;
  }
}

// sum of one value per cell (flat launch) into the block partials
__global__ void __launch_bounds__(NTF) k_cell_sum(LV lv, const double *val, double *partout) {
  const BlkView bv = blk_view(lv);
  double acc[1] = {0.0};
  for (long c = bv.b.lo + threadIdx.x; c < bv.b.hi; c += NTF) acc[0] += val[c];
  block_partial<1, 0u>(acc, partout, bv.b.slot);
}

// out = a + b over K arrays of NC (flat launch; a or b may be null, out may be a)
__global__ void __launch_bounds__(NTF) k_add2(LV lv, const double *a, const double *b, int K, double *out) {
  const BlkView bv = blk_view(lv);
  const long NC = lv.NC;
  for (long c = bv.b.lo + threadIdx.x; c < bv.b.hi; c += NTF)
    for (int kk = 0; kk < K; ++kk) {
      const size_t o = (size_t)kk * NC + c;
      out[o] = (a ? a[o] : 0.0) + (b ? b[o] : 0.0);
    }
}

// |a - b|^2 and |a|^2 over K arrays of NC (flat launch)
__global__ void __launch_bounds__(NTF) k_diff_norm(LV lv, const double *a, const double *b, int K,
                                                  double *partout) {
  const BlkView bv = blk_view(lv);
  const long NC = lv.NC;
  double acc[2] = {0.0, 0.0};
  for (long c = bv.b.lo + threadIdx.x; c < bv.b.hi; c += NTF)
    for (int kk = 0; kk < K; ++kk) {
      const double x = a[(size_t)kk * NC + c], d = x - b[(size_t)kk * NC + c];
      acc[0] += d * d;
      acc[1] += x * x;
    }
  block_partial<2, 0u>(acc, partout, bv.b.slot);
}

// D-hat part of the CMFD adjoint's gradient: d/d dn(s, g, c) of
// -Lambda^T M phi, Lambda = apsi psi + gam, for the node's six face sides
// (the matrix of k_matrix: lower side +dn (phi_c + phi_m) / h, upper side
// -dn (phi_c + phi_m) / h, phi_m = 0 across a boundary)
__global__ void __launch_bounds__(NTF) k_grad_dn(LV lv, const double *phi, const double *psi, double apsi,
                                                const double *gam, double *gdn) {
  const BlkView bv = blk_view(lv);
  const Geo &q = *bv.q;
  const long NC = lv.NC;
  const int G = lv.G;
  const long st[3] = {1, bv.b.sy, bv.b.sz};
  for (long c = bv.b.lo + threadIdx.x; c < bv.b.hi; c += NTF) {
    const unsigned char fl = q.flag[c];
    if (!(fl & F_ACTIVE)) continue;
    const long r0 = c - q.base;
    const int ii = (int)(r0 % q.NXP) - 1, jj = (int)((r0 / q.NXP) % q.NYP) - 1, kk = (int)(r0 / q.sz) - 1;
    const double hn[3] = {q.hx[ii], q.hy[jj], q.hz[kk]};
    for (int g = 0; g < G; ++g) {
      const double *ph = phi + (size_t)g * NC;
      const double lam = apsi * psi[(size_t)g * NC + c] + (gam ? gam[(size_t)g * NC + c] : 0.0);
      for (int f = 0; f < 6; ++f) {
        const int u = f >> 1;
        double sum = ph[c];
        if (fl & f_nb(f)) sum += ph[(f & 1) ? c - st[u] : c + st[u]];
        const double val = lam * sum / hn[u];
        gdn[((size_t)f * G + g) * NC + c] = (f & 1) ? -val : val;
      }
    }
  }
}

}  // namespace adg
