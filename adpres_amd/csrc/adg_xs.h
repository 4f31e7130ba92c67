// Node cross sections from the feedback tables, on the device:
// xsec::XS_updt (src/mod_xsec.f90:11-46) with its passes base_updt (:172),
// bcon_updt / ftem_updt / mtem_updt / cden_updt (:336-436), crod_updt
// (:230-287), Dsigr_updt (:199-226) and check_xs (:86-150).
//
// Every output element of the reference is a chain of independent passes
// over the same element (base value, then one "+ table * (parameter -
// reference value)" per feedback, then the rod increment, then the clamp of
// rodded columns), so one thread per node runs the chain for all of its
// entries in the same order and rounds exactly as the reference does:
// no contraction, IEEE division and square root.  The kernel reads per node
// its material, the node-wise parameters and its rod column's bank and
// level, and writes D, sigr, nuf, sigf and sigs in the device layout
// (group-major over the cells of the padded boxes).  It is bound by those
// writes: (4G + G^2) doubles per node against 4 reads.
#pragma once
#include "adg_common.h"

namespace adg {

// kinds of feedback table, in the order XS_updt applies them
enum { XK_BASE = 0, XK_BORON = 1, XK_FTEM = 2, XK_MTEM = 3, XK_CDEN = 4, XK_ROD = 5, XK_N = 6 };

struct XsArgs {
  const double *tab;        // [XK_N][4 G M + G G M]: sigtr, siga, nuf, sigf (m + M g), sigs (m + M (g + G h))
  const double *zdum;       // (nzz) depth from the core top of each level's top, summed top-down as crod_updt does
  const double *zdel;       // (nzz) level heights
  const int *mat0;          // (N) material - 1
  const int *rodk;          // (N) bank | (level << 16), 0 outside rodded columns
  const double *ftem, *mtem, *cden, *bpos;   // (N), (N), (N), (nb): NULL where off
  const long *cell;         // (N) node -> cell
  double *D, *sigr, *nuf, *sigf, *sigs;      // device arrays, group stride NC
  int *err;                 // set to 1 where the reference would stop
  long NC;
  int N, G, M, nzz, on;     // on: bit k for table kind k present
  double bcon, rbcon, rftem, rmtem, rcden, rod0, ssize;   // rod0 = coreh - pos0
};

template <int GG>   // group count known at compile time, 0: run time
__global__ void __launch_bounds__(256) k_xs_update(XsArgs a) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const int G = GG > 0 ? GG : a.G, M = a.M;
  const long K = (long)(4 * G + G * G) * M;
  const int m = a.mat0[n];
  const long cell = a.cell[n];
  // the feedback multipliers (parameter - reference value), in order
  double dv[4] = {0.0, 0.0, 0.0, 0.0};
  if (a.on & (1 << XK_BORON)) dv[0] = a.bcon - a.rbcon;
  if (a.on & (1 << XK_FTEM)) dv[1] = sqrt(a.ftem[n]) - sqrt(a.rftem);
  if (a.on & (1 << XK_MTEM)) dv[2] = a.mtem[n] - a.rmtem;
  if (a.on & (1 << XK_CDEN)) dv[3] = a.cden[n] - a.rcden;
  // rod: crod_updt walks the column from the top and stops after the first
  // level whose [dum, dum + zdel] holds the tip; levels above it are fully
  // rodded (the whole increment), that level gets it times the rodded
  // fraction, levels below none.  The whole column is clamped.
  int rmode = 0;           // 0 none, 1 whole, 2 partial
  double vf = 0.0;
  bool incol = false;
  if ((a.on & (1 << XK_ROD)) && a.rodk[n]) {
    incol = true;
    const int b = a.rodk[n] & 0xffff, k = a.rodk[n] >> 16;
    const double rodh = a.rod0 - a.bpos[b - 1] * a.ssize;
    bool above = false;
    for (int kk = a.nzz - 1; kk > k; --kk)
      if (rodh >= a.zdum[kk] && rodh <= a.zdum[kk] + a.zdel[kk]) {
        above = true;
        break;
      }
    if (!above) {
      if (rodh >= a.zdum[k] && rodh <= a.zdum[k] + a.zdel[k]) {
        rmode = 2;
        vf = (rodh - a.zdum[k]) / a.zdel[k];
      } else {
        rmode = 1;
      }
    }
  }
  bool bad = false;
  for (int g = 0; g < G; ++g) {
    double x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long o = (long)q * G * M + m + (long)M * g;
      double v = a.tab[o];
#pragma unroll
      for (int kd = 1; kd <= 4; ++kd)
        if (a.on & (1 << kd)) v = v + a.tab[kd * K + o] * dv[kd - 1];
      if (rmode == 1) v = v + a.tab[XK_ROD * K + o];
      else if (rmode == 2) v = v + vf * a.tab[XK_ROD * K + o];
      if (incol && q >= 1 && v < 0.0) v = 0.0;
      x[q] = v;
    }
    double dum = 0.0;
    for (int h = 0; h < G; ++h) {
      const long o = 4L * G * M + m + (long)M * (g + G * h);
      double v = a.tab[o];
#pragma unroll
      for (int kd = 1; kd <= 4; ++kd)
        if (a.on & (1 << kd)) v = v + a.tab[kd * K + o] * dv[kd - 1];
      if (rmode == 1) v = v + a.tab[XK_ROD * K + o];
      else if (rmode == 2) v = v + vf * a.tab[XK_ROD * K + o];
      if (incol && v < 0.0) v = 0.0;
      bad |= v < 0.0;
      a.sigs[(long)(h * G + g) * a.NC + cell] = v;   // device [to][from][cell]
      if (h != g) dum = dum + v;
    }
    // the thresholds are single-precision literals in the reference
    bad |= x[0] < (double)1.e-5f;
    const double D = 1.0 / (3.0 * x[0]);
    const double sr = x[1] + dum;
    bad |= D < (double)1.e-20f || sr < 0.0 || x[2] < 0.0;
    a.D[g * a.NC + cell] = D;
    a.sigr[g * a.NC + cell] = sr;
    a.nuf[g * a.NC + cell] = x[2];
    a.sigf[g * a.NC + cell] = x[3];
  }
  if (bad) *a.err = 1;
}

}  // namespace adg
