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

// The rod state of a node of a rodded column (bank | level << 16 in rodk):
// 0 unrodded, 1 whole increment, 2 the tip node with its rodded fraction vf
__device__ __forceinline__ void rod_state(int rodk, const double *bpos, double rod0, double ssize,
                                          const double *zdum, const double *zdel, int nzz, int &rmode, double &vf) {
#pragma clang fp contract(off)
  const int b = rodk & 0xffff, k = rodk >> 16;
  const double rodh = rod0 - bpos[b - 1] * ssize;
  const bool above = k < nzz - 1 && rodh >= 0.0 && rodh <= zdum[k];
  rmode = 0;
  vf = 0.0;
  if (!above) {
    if (rodh >= zdum[k] && rodh <= zdum[k] + zdel[k]) {
      rmode = 2;
      vf = (rodh - zdum[k]) / zdel[k];
    } else {
      rmode = 1;
    }
  }
}

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
  const bool incol = ks[XK_ROD] >= 0 && a.rodk[n];
  if (incol) rod_state(a.rodk[n], a.bpos, a.rod0, a.ssize, a.zdum, a.zdel, a.nzz, rmode, vf);
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

// ---------------------------------------------------------------------------
// Branch tables (%XTAB decks): xsec::XStab_updt (src/mod_xsec.f90:48-84),
// brInterp (:520-788) per node and crod_tab_updt (:291-376).  Each material
// holds nd x nb x nf x nm branch records (coolant density, boron, fuel and
// moderator temperature points) of E = 4G + G^2 + 6G doubles (sigtr, siga,
// nuf, sigf, sigs (from, to), dc (g, face)).  A node brackets its four
// parameters once, then every entry runs brInterp's chain on the eight
// corner records: moderator stage on all eight, fuel on four pairs, boron
// on two, density on one, in the reference's order; rodded nodes blend the
// rodded set's result in as crod_tab_updt does.
// ---------------------------------------------------------------------------

struct BrMat {
  int nd, nb, nf, nm, trod, pad;
  long poff, xoff, roff;   // offsets of its points, records and rodded records
};

struct XtabArgs {
  const BrMat *mats;
  const double *points, *xsec, *rxsec;
  const double *zdum, *zdel;
  const int *mat0, *rodk;
  const double *ftem, *mtem, *cden, *bpos;
  const long *cell;
  double *D, *sigr, *nuf, *sigf, *sigs, *dc;
  int *err;
  long NC;
  int N, G, nzz, rods;
  double bcon, rod0, ssize;
};

// brInterp's bracket search (:553-660): inside the range the first interval
// holding x; outside, the end interval when x is within 20 % (rel) or
// 100 ppm of the end point (default-real literals there); else false
__device__ __forceinline__ bool br_bracket(const double *p, int n, double x, bool rel, int &i1, int &i2) {
#pragma clang fp contract(off)
  i1 = 0;
  i2 = 0;
  if (n <= 1) return true;
  if (x >= p[0] && x <= p[n - 1]) {
    for (int s = 1; s < n; ++s)
      if (x >= p[s - 1] && x <= p[s]) {
        i1 = s - 1;
        i2 = s;
        break;
      }
    return true;
  }
  if (x < p[0] && (rel ? (p[0] - x) / p[0] < (double)0.2f : (p[0] - x) < (double)100.f)) {
    i1 = 0;
    i2 = 1;
    return true;
  }
  if (x > p[n - 1] && (rel ? (x - p[n - 1]) / p[n - 1] < (double)0.2f : (x - p[n - 1]) < (double)100.f)) {
    i1 = n - 2;
    i2 = n - 1;
    return true;
  }
  return false;
}

struct BrPos {
  long o[8];      // corner records at v1 (doubles from the material's first record)
  long dv;        // v2 record - v1 record
  double rm, rf, rb, rd;
  int nd, nb, nf, nm;
};

__device__ __forceinline__ bool br_setup(const XtabArgs &a, const BrMat &mm, int E, double cd, double bc, double ft,
                                         double mt, BrPos &P) {
#pragma clang fp contract(off)
  const double *pd = a.points + mm.poff, *pb = pd + mm.nd, *pf = pb + mm.nb, *pm = pf + mm.nf;
  int s1, s2, t1, t2, u1, u2, v1, v2;
  bool ok = br_bracket(pd, mm.nd, cd, true, s1, s2);
  ok &= br_bracket(pb, mm.nb, bc, false, t1, t2);
  ok &= br_bracket(pf, mm.nf, ft, true, u1, u2);
  ok &= br_bracket(pm, mm.nm, mt, true, v1, v2);
  if (!ok) return false;
  P.nd = mm.nd; P.nb = mm.nb; P.nf = mm.nf; P.nm = mm.nm;
  P.rm = mm.nm > 1 ? (mt - pm[v1]) / (pm[v2] - pm[v1]) : 0.0;
  P.rf = mm.nf > 1 ? (ft - pf[u1]) / (pf[u2] - pf[u1]) : 0.0;
  P.rb = mm.nb > 1 ? (bc - pb[t1]) / (pb[t2] - pb[t1]) : 0.0;
  P.rd = mm.nd > 1 ? (cd - pd[s1]) / (pd[s2] - pd[s1]) : 0.0;
  const long snb = (long)mm.nd * mm.nb, snf = snb * mm.nf;
  const int cs[8] = {s1, s1, s1, s1, s2, s2, s2, s2}, ct[8] = {t1, t1, t2, t2, t1, t1, t2, t2};
  const int cu[8] = {u1, u2, u1, u2, u1, u2, u1, u2};
#pragma unroll
  for (int c = 0; c < 8; ++c) P.o[c] = (long)E * (cs[c] + (long)mm.nd * ct[c] + snb * cu[c] + snf * v1);
  P.dv = (long)E * snf * (v2 - v1);
  return true;
}

// one entry e of brInterp's chain on table block T
__device__ __forceinline__ double br_elem(const double *T, const BrPos &P, int e) {
#pragma clang fp contract(off)
  double c[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double A = T[P.o[k] + e];
    if (P.nm > 1) {
      const double B = T[P.o[k] + P.dv + e];
      c[k] = A + P.rm * (B - A);
    } else {
      c[k] = A;
    }
  }
  if (P.nf > 1) {
#pragma unroll
    for (int k = 0; k < 8; k += 2) c[k] = c[k] + P.rf * (c[k + 1] - c[k]);
  }
  if (P.nb > 1) {
    c[0] = c[0] + P.rb * (c[2] - c[0]);
    c[4] = c[4] + P.rb * (c[6] - c[4]);
  }
  if (P.nd > 1) c[0] = c[0] + P.rd * (c[4] - c[0]);
  return c[0];
}

template <int GG>
__global__ void __launch_bounds__(256) k_xtab_update(XtabArgs a) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= a.N) return;
  const int G = GG > 0 ? GG : a.G, E = 4 * G + G * G + 6 * G;
  const BrMat mm = a.mats[a.mat0[n]];
  const long cell = a.cell[n];
  BrPos P;
  if (!br_setup(a, mm, E, a.cden[n], a.bcon, a.ftem[n], a.mtem[n], P)) {
    *a.err = 1;
    return;
  }
  // rod state as in k_xs_update (crod_tab_updt walks the column the same way)
  int rmode = 0;
  double vf = 0.0;
  const bool incol = a.rods && a.rodk[n];
  if (incol) rod_state(a.rodk[n], a.bpos, a.rod0, a.ssize, a.zdum, a.zdel, a.nzz, rmode, vf);
  if (rmode && mm.trod != 1) {     // a rodded node of a material without rodded data
    *a.err = 1;
    return;
  }
  const double *TX = a.xsec + mm.xoff, *TR = a.rxsec + (rmode ? mm.roff : 0);
  auto val = [&](int e) {
    double x = br_elem(TX, P, e);
    if (rmode) {
      const double r = br_elem(TR, P, e);
      x = rmode == 2 ? (1.0 - vf) * x + vf * r : r;
    }
    return x;
  };
  bool bad = false;
  for (int g = 0; g < G; ++g) {
    const double st = val(g);
    double sa = val(G + g), nf = val(2 * G + g), sf = val(3 * G + g);
    if (incol) {
      if (sa < 0.0) sa = 0.0;
      if (nf < 0.0) nf = 0.0;
      if (sf < 0.0) sf = 0.0;
    }
    double dum = 0.0;
    for (int h = 0; h < G; ++h) {
      double v = val(4 * G + g + G * h);
      if (incol && v < 0.0) v = 0.0;
      bad |= v < 0.0;
      a.sigs[(long)(h * G + g) * a.NC + cell] = v;
      if (h != g) dum = dum + v;
    }
    bad |= st < (double)1.e-5f;
    const double D = 1.0 / (3.0 * st);
    const double sr = sa + dum;
    bad |= D < (double)1.e-20f || sr < 0.0 || nf < 0.0;
    a.D[g * a.NC + cell] = D;
    a.sigr[g * a.NC + cell] = sr;
    a.nuf[g * a.NC + cell] = nf;
    a.sigf[g * a.NC + cell] = sf;
  }
  for (int f = 0; f < 6; ++f)
    for (int g = 0; g < G; ++g) {
      double v = val(4 * G + G * G + g + G * f);
      if (incol && v < 0.0) v = 0.0;
      a.dc[(long)(f * G + g) * a.NC + cell] = v;
    }
  if (bad) *a.err = 1;
}

// The same update in tiles of XT_NODES nodes per block, so that both the
// table reads and the output writes are coalesced: a node's entries are
// computed by consecutive lanes (the eight corner records of one node are
// read as contiguous runs), kept in LDS, and written out node-fastest per
// output plane.  Phase 0 brackets each node and finds its rod state, phase
// 1 runs the per-entry chains, phase 2 forms D and sigr from the tile (the
// reference's h order) and writes every plane.
constexpr int XT_NODES = 32;

template <int GG>
__global__ void __launch_bounds__(512) k_xtab_tile(XtabArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double tile[];              // [XT_NODES][E]
  __shared__ BrPos pos[XT_NODES];
  __shared__ double vfr[XT_NODES];
  __shared__ long cl[XT_NODES], tx[XT_NODES], tr[XT_NODES];
  __shared__ int st[XT_NODES];                  // bit0 valid, bit1 in rodded column, bits 2-3 rod mode
  const int G = GG > 0 ? GG : a.G, E = 4 * G + G * G + 6 * G;
  const int n0 = blockIdx.x * XT_NODES, nn = min(XT_NODES, a.N - n0);
  const int tid = threadIdx.x;
  if (tid < nn) {
    const int n = n0 + tid;
    const BrMat mm = a.mats[a.mat0[n]];
    int flags = 0;
    BrPos P;
    if (br_setup(a, mm, E, a.cden[n], a.bcon, a.ftem[n], a.mtem[n], P)) {
      int rmode = 0;
      double vf = 0.0;
      const bool incol = a.rods && a.rodk[n];
      if (incol) rod_state(a.rodk[n], a.bpos, a.rod0, a.ssize, a.zdum, a.zdel, a.nzz, rmode, vf);
      if (rmode && mm.trod != 1) {
        *a.err = 1;
      } else {
        flags = 1 | (incol ? 2 : 0) | (rmode << 2);
        vfr[tid] = vf;
        tx[tid] = mm.xoff;
        tr[tid] = rmode ? mm.roff : 0;
      }
    } else {
      *a.err = 1;
    }
    pos[tid] = P;
    st[tid] = flags;
    cl[tid] = a.cell[n];
  }
  __syncthreads();
  for (int i = tid; i < nn * E; i += blockDim.x) {
    const int j = i / E, e = i - j * E;
    const int f = st[j];
    if (!(f & 1)) continue;
    const int rmode = f >> 2;
    double x = br_elem(a.xsec + tx[j], pos[j], e);
    if (rmode) {
      const double r = br_elem(a.rxsec + tr[j], pos[j], e);
      x = rmode == 2 ? (1.0 - vfr[j]) * x + vfr[j] * r : r;
    }
    if ((f & 2) && e >= G && x < 0.0) x = 0.0;    // the rodded column's clamp (not sigtr)
    tile[i] = x;
  }
  __syncthreads();
  const int NP = 4 * G + G * G + 6 * G;          // output planes: D, sigr, nuf, sigf, sigs, dc
  bool bad = false;
  for (int i = tid; i < nn * NP; i += blockDim.x) {
    const int p = i / nn, j = i - p * nn;
    if (!(st[j] & 1)) continue;
    const double *t = tile + (long)j * E;
    const long cell = cl[j];
    if (p < G) {
      const double s = t[p];
      bad |= s < (double)1.e-5f;
      const double D = 1.0 / (3.0 * s);
      bad |= D < (double)1.e-20f;
      a.D[p * a.NC + cell] = D;
    } else if (p < 2 * G) {
      const int g = p - G;
      double dum = 0.0;
      for (int h = 0; h < G; ++h)
        if (h != g) dum = dum + t[4 * G + g + G * h];
      const double sr = t[G + g] + dum;
      bad |= sr < 0.0;
      a.sigr[g * a.NC + cell] = sr;
    } else if (p < 3 * G) {
      const double v = t[p];
      bad |= v < 0.0;
      a.nuf[(p - 2 * G) * a.NC + cell] = v;
    } else if (p < 4 * G) {
      a.sigf[(p - 3 * G) * a.NC + cell] = t[p];
    } else if (p < 4 * G + G * G) {
      const int q = p - 4 * G, g = q % G, h = q / G;   // record index g + G h: from g, to h
      const double v = t[p];
      bad |= v < 0.0;
      a.sigs[(long)(h * G + g) * a.NC + cell] = v;
    } else {
      a.dc[(long)(p - 4 * G - G * G) * a.NC + cell] = t[p];
    }
  }
  if (bad) *a.err = 1;
}

}  // namespace adg
