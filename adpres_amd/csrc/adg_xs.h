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
// no contraction, IEEE division and square root.  The tables (a few tens
// of KB) are staged in LDS by each block of a grid of a few blocks per CU
// when they fit, and read through the caches otherwise.  The kernel reads per node
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

// One node: the element chains of XS_updt for all of its entries.  T holds
// the tables at offsets ks[kind] (ks < 0: kind absent or folded into the
// base slot); dv the per-node feedback multipliers.
template <int GG>
__device__ __forceinline__ void xs_node(const XsArgs &a, int n, const double *T, const long *ks) {
#pragma clang fp contract(off)
  const int G = GG > 0 ? GG : a.G, M = a.M;
  const int m = a.mat0[n];
  const long cell = a.cell[n];
  // the feedback multipliers (parameter - reference value), in order
  double dv[4] = {0.0, 0.0, 0.0, 0.0};
  if (ks[XK_BORON] >= 0) dv[0] = a.bcon - a.rbcon;
  if (ks[XK_FTEM] >= 0) dv[1] = sqrt(a.ftem[n]) - sqrt(a.rftem);
  if (ks[XK_MTEM] >= 0) dv[2] = a.mtem[n] - a.rmtem;
  if (ks[XK_CDEN] >= 0) dv[3] = a.cden[n] - a.rcden;
  // rod: crod_updt walks the column from the top and stops after the first
  // level whose [dum, dum + zdel] holds the tip; levels above it are fully
  // rodded (the whole increment), that level gets it times the rodded
  // fraction, levels below none.  The whole column is clamped.  The walk
  // stopped above level k exactly when 0 <= rodh <= zdum[k]: the tested
  // intervals of the levels above are [zdum[kk], zdum[kk] + zdel[kk]] with
  // zdum[kk] + zdel[kk] rounded to the same double as zdum[kk - 1] (the host
  // summed zdum with that very addition, from zdum[nzz - 1] = 0), so they
  // tile [0, zdum[k]] without gaps.
  int rmode = 0;           // 0 none, 1 whole, 2 partial
  double vf = 0.0;
  bool incol = false;
  if (ks[XK_ROD] >= 0 && a.rodk[n]) {
    incol = true;
    const int b = a.rodk[n] & 0xffff, k = a.rodk[n] >> 16;
    const double rodh = a.rod0 - a.bpos[b - 1] * a.ssize;
    const bool above = k < a.nzz - 1 && rodh >= 0.0 && rodh <= a.zdum[k];
    if (!above) {
      if (rodh >= a.zdum[k] && rodh <= a.zdum[k] + a.zdel[k]) {
        rmode = 2;
        vf = (rodh - a.zdum[k]) / a.zdel[k];
      } else {
        rmode = 1;
      }
    }
  }
  const double *TR = T + (ks[XK_ROD] >= 0 ? ks[XK_ROD] : 0);
  bool bad = false;
  for (int g = 0; g < G; ++g) {
    double x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long o = (long)q * G * M + m + (long)M * g;
      double v = T[o];
#pragma unroll
      for (int kd = 1; kd <= 4; ++kd)
        if (ks[kd] >= 0) v = v + T[ks[kd] + o] * dv[kd - 1];
      if (rmode == 1) v = v + TR[o];
      else if (rmode == 2) v = v + vf * TR[o];
      if (incol && q >= 1 && v < 0.0) v = 0.0;
      x[q] = v;
    }
    double dum = 0.0;
    // unrolled for a compiled group count, so that the table reads of a
    // row's entries are in flight together
#pragma unroll
    for (int h = 0; h < G; ++h) {
      const long o = 4L * G * M + m + (long)M * (g + G * h);
      double v = T[o];
#pragma unroll
      for (int kd = 1; kd <= 4; ++kd)
        if (ks[kd] >= 0) v = v + T[ks[kd] + o] * dv[kd - 1];
      if (rmode == 1) v = v + TR[o];
      else if (rmode == 2) v = v + vf * TR[o];
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

// Tables read from global memory (through the caches): one node per thread.
template <int GG>   // group count known at compile time, 0: run time
__global__ void __launch_bounds__(256) k_xs_update(XsArgs a) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const int G = GG > 0 ? GG : a.G;
  const long K = (long)(4 * G + G * G) * a.M;
  long ks[XK_N];
  for (int kd = 0; kd < XK_N; ++kd) ks[kd] = (a.on & (1 << kd)) || kd == 0 ? kd * K : -1;
  xs_node<GG>(a, n, a.tab, ks);
}

// Tables staged in LDS once per block (a grid of a few blocks per CU walks
// the nodes; blocks of 512 threads, so that the LDS a table takes still
// leaves many waves per CU), with the boron pass folded into the base slot: base + c * dv
// is the same operation on the same operands for every node of a material,
// so it is done once per block.  LDS slots: base(+boron), then the fuel
// temperature, moderator temperature, coolant density and rod tables that
// are present.
template <int GG>
__global__ void __launch_bounds__(512) k_xs_update_lds(XsArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double sh[];
  const int G = GG > 0 ? GG : a.G;
  const long K = (long)(4 * G + G * G) * a.M;
  long ks[XK_N];
  ks[XK_BASE] = 0;
  ks[XK_BORON] = -1;
  long next = K;
  for (int kd = XK_FTEM; kd < XK_N; ++kd) {
    ks[kd] = (a.on & (1 << kd)) ? next : -1;
    if (ks[kd] >= 0) next += K;
  }
  const bool boron = a.on & (1 << XK_BORON);
  const double dvb = a.bcon - a.rbcon;
  for (long i = threadIdx.x; i < K; i += blockDim.x) {
    double v = a.tab[i];
    if (boron) v = v + a.tab[K + i] * dvb;
    sh[i] = v;
    for (int kd = XK_FTEM; kd < XK_N; ++kd)
      if (ks[kd] >= 0) sh[ks[kd] + i] = a.tab[kd * K + i];
  }
  __syncthreads();
  // the per-node chain sees boron through the base slot (ks[XK_BORON] < 0)
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += gridDim.x * blockDim.x)
    xs_node<GG>(a, n, sh, ks);
}

}  // namespace adg
