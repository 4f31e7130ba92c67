// CMFD kernels: coupling coefficients, 7-point matrix, sources, fused
// BiCGSTAB and the outer-iteration reductions (src/mod_cmfd.f90).
//
// Every kernel runs over a batch of cores (one core is a batch of one): the
// launch table gives each block its core and its share of that core's
// cells, and all scalars (k, BiCGSTAB coefficients, norms, the schedule)
// are per core.  Reductions are segmented by core: a block's partial goes to
// its slot, and a consumer block reduces only its own core's slots.
#pragma once

#include "adg_common.h"

namespace adg {

// the core's first block writes its per-core scalars
__device__ __forceinline__ bool core_leader(const BlkView &v) {
  return v.b.slot == v.b.pfirst && threadIdx.x == 0;
}

// ---------------------------------------------------------------------------
// coup_coef (src/mod_cmfd.f90:11-137) and matrix_setup (:217-304)
// ---------------------------------------------------------------------------

// local node index t of core q (k slowest, then j, then i) -> (i, j, k) and cell
__device__ __forceinline__ bool decode_cell(long t, const Geo &q, int &i, int &j, int &k, long &c) {
  const long nxy = (long)q.nxx * q.nyy;
  if (t >= nxy * q.nzz) return false;
  k = (int)(t / nxy);
  long r = t - (long)k * nxy;
  j = (int)(r / q.nxx);
  i = (int)(r - (long)j * q.nxx);
  c = q.base + ((long)(k + 1) * q.NYP + (j + 1)) * q.NXP + (i + 1);
  return true;
}

__device__ __forceinline__ double df_bnd(int bc, double D, double h) {
#pragma clang fp contract(off)
  const double alb = 1.e30;
  if (bc == 0) return 2.0 * alb * D / (2.0 * D + alb * h);
  if (bc == 1) return D / (2.0 * D + 0.5 * h);
  return 0.0;
}

__device__ __forceinline__ double df_int(double D, double Dnb, double hnb, double h) {
#pragma clang fp contract(off)
  return 2.0 * D * Dnb / (D * hnb + Dnb * h);
}

// one thread per (node, group) of the per-node table; blockIdx.y = g
__global__ void __launch_bounds__(NT) k_coupcoef(LV lv, XS xs, double *df) {
  const BlkView v = blk_view(lv);
  const Geo &q = *v.q;
  int i, j, k;
  long c;
  const long t = v.b.lo + threadIdx.x;
  if (t >= v.b.hi || !decode_cell(t, q, i, j, k, c)) return;
  const int g = blockIdx.y;
  const long NC = q.NC;
  const unsigned char fl = q.flag[c];
  if (!(fl & F_ACTIVE)) return;
  const double *Dg = xs.D + (size_t)g * NC;
  const double D = Dg[c];
  const double h[3] = {q.hx[i], q.hy[j], q.hz[k]};
  const long st[3] = {1, q.sy, q.sz};
  for (int s = 0; s < 6; ++s) {
    const int u = s >> 1;
    const bool up = (s & 1) == 0;
    double val;
    if (fl & f_nb(s)) {
      const long cn = up ? c + st[u] : c - st[u];
      const double hnb = (u == 0) ? q.hx[up ? i + 1 : i - 1] : (u == 1) ? q.hy[up ? j + 1 : j - 1]
                                                                         : q.hz[up ? k + 1 : k - 1];
      val = df_int(D, Dg[cn], hnb, h[u]);
    } else {
      val = df_bnd(q.bc[s], D, h[u]);
    }
    df[((size_t)s * q.G + g) * NC + c] = val;
  }
}

// Matrix slots per row: 0 z-, 1 y-, 2 x-, 3 diag, 4 x+, 5 y+, 6 z+.  Absent
// neighbours (and every slot of an inactive cell) hold 0.
// sched != 0: rebuild only in a nodal-update outer of the core (device-side schedule).
// transpose != 0: the transposed matrix, row n slot s = the forward matrix's
// row m (the neighbour across face s) slot opposite(s), evaluated from m's
// own face coefficients so that it is the exact transpose.
__device__ __forceinline__ double offdiag(const double *df, const double *dn, int G, long NC, int g, long c, int s,
                                          double h) {
#pragma clang fp contract(off)
  const double f = df[((size_t)s * G + g) * NC + c], d = dn[((size_t)s * G + g) * NC + c];
  return (s & 1) ? -(f - d) / h : -(f + d) / h;   // lower faces 1/3/5: -(df - dn)/h; upper: -(df + dn)/h
}

template <typename T>
__global__ void __launch_bounds__(NT) k_matrix(LV lv, XS xs, const double *df, const double *dn, double *A64,
                                               T *AT, int sched, int transpose) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  const Geo &q = *v.q;
  if (sched && !nodal_outer(v.sc)) return;
  int i, j, k;
  long c;
  const long t = v.b.lo + threadIdx.x;
  if (t >= v.b.hi || !decode_cell(t, q, i, j, k, c)) return;
  const int g = blockIdx.y;
  const long NC = q.NC;
  const int G = q.G;
  const unsigned char fl = q.flag[c];
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  if (fl & F_ACTIVE) {
    double f[6], d[6];
    for (int s = 0; s < 6; ++s) {
      f[s] = df[((size_t)s * G + g) * NC + c];
      d[s] = dn[((size_t)s * G + g) * NC + c];
    }
    const double hx = q.hx[i], hy = q.hy[j], hz = q.hz[k];
    a[3] = (f[0] + f[1] - d[0] + d[1]) / hx + (f[2] + f[3] - d[2] + d[3]) / hy +
           (f[4] + f[5] - d[4] + d[5]) / hz + xs.sigr[(size_t)g * NC + c];
    if (!transpose) {
      if (fl & f_nb(5)) a[0] = -(f[5] - d[5]) / hz;
      if (fl & f_nb(3)) a[1] = -(f[3] - d[3]) / hy;
      if (fl & f_nb(1)) a[2] = -(f[1] - d[1]) / hx;
      if (fl & f_nb(0)) a[4] = -(f[0] + d[0]) / hx;
      if (fl & f_nb(2)) a[5] = -(f[2] + d[2]) / hy;
      if (fl & f_nb(4)) a[6] = -(f[4] + d[4]) / hz;
    } else {
      // slot -> (face of n, neighbour offset, opposite face, neighbour size)
      if (fl & f_nb(5)) a[0] = offdiag(df, dn, G, NC, g, c - q.sz, 4, q.hz[k - 1]);
      if (fl & f_nb(3)) a[1] = offdiag(df, dn, G, NC, g, c - q.sy, 2, q.hy[j - 1]);
      if (fl & f_nb(1)) a[2] = offdiag(df, dn, G, NC, g, c - 1, 0, q.hx[i - 1]);
      if (fl & f_nb(0)) a[4] = offdiag(df, dn, G, NC, g, c + 1, 1, q.hx[i + 1]);
      if (fl & f_nb(2)) a[5] = offdiag(df, dn, G, NC, g, c + q.sy, 3, q.hy[j + 1]);
      if (fl & f_nb(4)) a[6] = offdiag(df, dn, G, NC, g, c + q.sz, 5, q.hz[k + 1]);
    }
  }
  for (int s = 0; s < 7; ++s) {
    if (A64) A64[((size_t)g * 7 + s) * NC + c] = a[s];
    if (AT) AT[((size_t)g * 7 + s) * NC + c] = (T)a[s];
  }
}

// y(c) = A_g x in the reference's slot order (sp_matvec, src/mod_cmfd.f90:1260-1265);
// absent slots contribute an exact +0.
template <typename TA, typename TX>
__device__ __forceinline__ double stencil(const TA *__restrict__ A, long NC, const TX *__restrict__ x,
                                          long c, long sy, long sz) {
#pragma clang fp contract(off)
  double s = 0.0;
  s = s + (double)A[0 * NC + c] * (double)x[c - sz];
  s = s + (double)A[1 * NC + c] * (double)x[c - sy];
  s = s + (double)A[2 * NC + c] * (double)x[c - 1];
  s = s + (double)A[3 * NC + c] * (double)x[c];
  s = s + (double)A[4 * NC + c] * (double)x[c + 1];
  s = s + (double)A[5 * NC + c] * (double)x[c + sy];
  s = s + (double)A[6 * NC + c] * (double)x[c + sz];
  return s;
}

__global__ void __launch_bounds__(NTF) k_matvec(LV lv, const double *A64, int g, const double *x, double *y) {
  const BlkView v = blk_view(lv);
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF)
    y[c] = stencil(A64 + (size_t)g * 7 * lv.NC, lv.NC, x, c, v.b.sy, v.b.sz);
}

// ---------------------------------------------------------------------------
// sources: TSrc / TSrcAd (src/mod_cmfd.f90:1006-1064)
// ---------------------------------------------------------------------------

// smode 0: TSrc (forward, :1006-1033), 1: TSrcAd (:1037-1064), 2: TSrcTr
// (transient, :1068-1096: fission source weighted by 1 - beta + dfis, no 1/k)
// (compiled per group count GG, so that the in-scatter loads of a cell are
// issued together).  smask[(core * 2 + adjoint) * G + g]: the in-scatter
// source groups h whose sigma_s term is nonzero somewhere in the core (bit
// h) and whether the external source is (SMASK_EX); the terms left out are
// exactly +0 (sigma_s = 0 times a finite flux, or a zero source), so the
// sums are the same as with every term.
constexpr unsigned SMASK_EX = 1u << 30;

template <int TAB, int GG>
__global__ void __launch_bounds__(NTF) k_src(LV lv, XS xs, int g, int smode, int first_group, const double *f0,
                                            double *f0c, const double *fs0, double *fs0c, double *b,
                                            const unsigned *smask) {
#pragma clang fp contract(off)
  const BlkView v = blk_view_t<TAB>(lv);
  if (v.sc->done) return;
  const double Ke = v.sc->Ke;
  const long NC = lv.NC;
  constexpr int G = GG;
  const int adjoint = smode == 1;
  const unsigned msk = smask[((size_t)v.b.core * 2 + adjoint) * G + g];
  const bool ex = (msk & SMASK_EX) != 0;
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    if (first_group) fs0c[c] = fs0[c];
    f0c[(size_t)g * NC + c] = f0[(size_t)g * NC + c];
    // every in-scatter term's loads issued first, then summed in order
    double sgv[G], fv[G];
#pragma unroll
    for (int h = 0; h < G; ++h) {
      const bool on = h != g && ((msk >> h) & 1u);
      sgv[h] = on ? (adjoint ? xs.sigs[((size_t)h * G + g) * NC + c] : xs.sigs[((size_t)g * G + h) * NC + c]) : 0.0;
      fv[h] = on ? f0[(size_t)h * NC + c] : 0.0;
    }
    double s0 = 0.0;
#pragma unroll
    for (int h = 0; h < G; ++h) {
      if (h == g || !((msk >> h) & 1u)) continue;
      s0 = s0 + sgv[h] * fv[h];
    }
    const double e = ex ? xs.exsrc[(size_t)g * NC + c] : 0.0;
    if (smode == 2) {
      b[c] = xs.wtr[c] * xs.chi[(size_t)g * NC + c] * fs0[c] + s0 + e;
      continue;
    }
    const double w = adjoint ? xs.nuf[(size_t)g * NC + c] : xs.chi[(size_t)g * NC + c];
    b[c] = w * fs0[c] / Ke + s0 + e;
  }
}

// ---------------------------------------------------------------------------
// BiCGSTAB (src/mod_cmfd.f90:1203-1243): a setup kernel, two fused stencil
// kernels per iteration and a final update.  TV is the vector/matrix type
// of the iteration (double, or float for the FP32 form); the residual of
// the setup is always formed in FP64 from the FP64 matrix and FP64 flux.
//   setup: r = b - A x (FP64), rs = r, d = 0            dot (rs, r)
// For TV = double, d is the flux itself; for float, d is the correction that
// the final kernel adds to the FP64 flux.
// ---------------------------------------------------------------------------

template <typename TV, int TAB>
__global__ void __launch_bounds__(NTF) k_bicg_setup(LV lv, const double *A64g, const double *b, const double *x,
                                                   TV *r, TV *rs, TV *d, double *pout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view_t<TAB>(lv);
  if (v.sc->done) return;
  double acc[1] = {0.0};
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    const double rr = b[c] - stencil(A64g, lv.NC, x, c, v.b.sy, v.b.sz);
    const TV rv = (TV)rr;
    r[c] = rv;
    rs[c] = rv;
    if (d) d[c] = (TV)0;
    acc[0] += (double)rv * (double)rv;
  }
  block_partial<1, 0u, TAB == 0>(acc, pout, v.b.slot);
  if (core_leader(v)) {   // rho = alpha = omega = 1 before iteration 1
    v.sc->rho_s[0] = 1.0;
    v.sc->alpha_s[0] = 1.0;
    v.sc->omega_s[0] = 1.0;
  }
}

// BiCGSTAB in two stencil kernels per iteration.  With the reference's
// recurrences (src/mod_cmfd.f90:1213-1241) for iteration i,
//   KB(i): s = r - alpha_i v at c and its six neighbours, t = A s;
//          dots (t,t), (t,s), (rs,t), (rs,s)
//   KA(i+1): omega_i = (t,s)/(t,t),  rho_{i+1} = (rs,s) - omega_i (rs,t)
//          [= (rs, r_i) with r_i = s - omega_i t],
//          beta = (rho_{i+1}/rho_i)(alpha_i/omega_i); at c and its neighbours
//          r_i = s - omega_i t and p_{i+1} = r_i + beta (p_i - omega_i v_i);
//          x += alpha_i p_i + omega_i s at c; v_{i+1} = A p_{i+1}; dot (rs, v)
//   KB(i+1): alpha_{i+1} = rho_{i+1} / (rs, v_{i+1}), ...
//   KT:    the last iteration's x update.
// In exact arithmetic these are the reference's iterates; the x/r update of
// the three-kernel form is recomputed where the next stencil reads it.  r, p
// and v ping-pong between two buffers because neighbours read old values.
//
// Each cell's loads (4 vectors at 7 points and 7 coefficients in KA, 2
// vectors at 7 points and 7 coefficients in KB) are all issued before any
// arithmetic, so a wavefront waits for one memory round trip per cell
// instead of one per stencil point; the first iteration (p = r) is a
// separate instantiation so that no branch splits the load batch.  The FP64
// kernels are allowed 128 VGPRs (at most 4 waves per SIMD), which the batch
// of 35 double loads needs.

// occupancy bound that lets one cell's loads be in flight together
template <typename TV>
struct StencilWaves {
  static constexpr int value = sizeof(TV) == 8 ? 4 : 8;
};

// one cell's operands of KA / KB
template <typename TV, bool FIRST>
struct KAIn {
  TV R[7], V[7], T[7], P[7], Ac[7], rsc, xc;
};
template <typename TV, bool FIRST>
__device__ __forceinline__ void ka_load(KAIn<TV, FIRST> &L, const TV *__restrict__ AT, const TV *__restrict__ r,
                                        const TV *__restrict__ p, const TV *__restrict__ v,
                                        const TV *__restrict__ t, const TV *__restrict__ rs,
                                        const TV *__restrict__ x, long NC, const long (&off)[7], long c) {
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    L.R[k] = r[c + off[k]];
    if (!FIRST) {
      L.V[k] = v[c + off[k]];
      L.T[k] = t[c + off[k]];
      L.P[k] = p[c + off[k]];
    }
    L.Ac[k] = AT[k * NC + c];
  }
  L.rsc = rs[c];
  L.xc = FIRST ? (TV)0 : x[c];
}

template <typename TV>
struct KBIn {
  TV R[7], V[7], Ac[7], rsc;
};
template <typename TV>
__device__ __forceinline__ void kb_load(KBIn<TV> &L, const TV *__restrict__ AT, const TV *__restrict__ r,
                                        const TV *__restrict__ v, const TV *__restrict__ rs, long NC,
                                        const long (&off)[7], long c) {
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    L.R[k] = r[c + off[k]];
    L.V[k] = v[c + off[k]];
    L.Ac[k] = AT[k * NC + c];
  }
  L.rsc = rs[c];
}

// stores of the stencil kernels' vector outputs; with ADG_NT_STORE they are
// streaming (non-temporal) stores
template <typename TV>
__device__ __forceinline__ void vstore(TV *p, TV v) {
#if ADG_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// one cell of KA from its loaded operands
template <typename TV, bool FIRST>
__device__ __forceinline__ void ka_cell(const KAIn<TV, FIRST> &L, long c, TV ta, TV tw, TV tb, TV *__restrict__ x,
                                        TV *__restrict__ rout, TV *__restrict__ pout, TV *__restrict__ vout,
                                        double &acc) {
#pragma clang fp contract(off)
  TV pn[7], rn;
  if (FIRST) {
#pragma unroll
    for (int k = 0; k < 7; ++k) pn[k] = L.R[k];
    rn = L.R[3];
  } else {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const TV s_ = (TV)(L.R[k] - ta * L.V[k]);
      const TV r_ = (TV)(s_ - tw * L.T[k]);
      if (k == 3) rn = r_;
      pn[k] = (TV)(r_ + tb * (L.P[k] - tw * L.V[k]));
    }
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 7; ++k) s = s + (double)(L.Ac[k] * pn[k]);
  const TV vc = (TV)s;
  if (!FIRST) {
    const TV sv = (TV)(L.R[3] - ta * L.V[3]);
    vstore<TV>(x + c, (TV)(L.xc + ta * L.P[3] + tw * sv));
  }
  vstore<TV>(rout + c, rn);
  vstore<TV>(pout + c, pn[3]);
  vstore<TV>(vout + c, vc);
  acc += (double)L.rsc * (double)vc;
}

// one cell of KB from its loaded operands
template <typename TV>
__device__ __forceinline__ void kb_cell(const KBIn<TV> &L, long c, TV alpha, TV *__restrict__ t, double (&acc)[4]) {
#pragma clang fp contract(off)
  const double rsc = (double)L.rsc;
  TV sv[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) sv[k] = (TV)(L.R[k] - alpha * L.V[k]);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < 7; ++k) s = s + (double)(L.Ac[k] * sv[k]);
  const TV tc = (TV)s;
  vstore<TV>(t + c, tc);
  acc[0] += (double)tc * (double)tc;
  acc[1] += (double)tc * (double)sv[3];
  acc[2] += rsc * (double)tc;
  acc[3] += rsc * (double)sv[3];
}

// With a single core (TAB == 0) the first cell's operands are requested
// before the scalar prologue (the done flag, the partial sums of the
// previous kernel and their reduction), which they do not depend on, so the
// two memory round trips overlap.  After convergence this costs the first
// cell's loads of each remaining launch of the replayed chunk; with a batch
// table the loads wait for the done flag, so converged cores read nothing.
template <typename TV, bool FIRST, int TAB>
__global__ void __launch_bounds__(NTF) __attribute__((amdgpu_waves_per_eu(1, StencilWaves<TV>::value)))
k_bicgA(LV lv, const TV *__restrict__ AT, int it, const TV *__restrict__ r, const TV *__restrict__ p,
        const TV *__restrict__ v, const TV *__restrict__ t, const TV *__restrict__ rs, TV *__restrict__ x,
        TV *__restrict__ rout, TV *__restrict__ pout, TV *__restrict__ vout, const double *__restrict__ pin,
        double *__restrict__ partout) {
#pragma clang fp contract(off)
  const BlkView bv = blk_view_t<TAB>(lv);
  Scal *__restrict__ sc = bv.sc;
  const long NC = lv.NC, sy = bv.b.sy, sz = bv.b.sz;
  const long off[7] = {-sz, -sy, -1, 0, 1, sy, sz};   // slot order z-, y-, x-, c, x+, y+, z+
  const long c0 = bv.b.lo + threadIdx.x;
  KAIn<TV, FIRST> L;
  if (TAB == 0 && c0 < bv.b.hi) ka_load<TV, FIRST>(L, AT, r, p, v, t, rs, x, NC, off, c0);
  if (sc->done) return;
  const int prv = (it - 1) & 1, cur = it & 1;
  double rho, alpha_d = 0.0, omega_d = 0.0, beta = 0.0;
  if (FIRST) {
    double red[1];
    reduce_partials<1, 0u, TAB == 0>(pin, bv.b.pfirst, bv.b.pn, red);
    rho = red[0];
  } else {
    double red[4];   // (t,t), (t,s), (rs,t), (rs,s) of iteration it-1
    reduce_partials<4, 0u, TAB == 0>(pin, bv.b.pfirst, bv.b.pn, red);
    omega_d = red[1] / red[0];
    rho = red[3] - omega_d * red[2];
    alpha_d = sc->alpha_s[prv];
    beta = (rho / sc->rho_s[prv]) * (alpha_d / omega_d);
  }
  const TV ta = (TV)alpha_d, tw = (TV)omega_d, tb = (TV)beta;
  double acc[1] = {0.0};
  // ADG_KUNROLL cells per thread in flight: their loads are all issued
  // before the first is used; the cells are still summed in order
  for (long c = c0; c < bv.b.hi; c += ADG_KUNROLL * NTF) {
    if (TAB != 0 || c != c0) ka_load<TV, FIRST>(L, AT, r, p, v, t, rs, x, NC, off, c);
#if ADG_KUNROLL == 2
    KAIn<TV, FIRST> L2;
    const long c2 = c + NTF;
    const bool h2 = c2 < bv.b.hi;
    if (h2) ka_load<TV, FIRST>(L2, AT, r, p, v, t, rs, x, NC, off, c2);
#endif
    ka_cell<TV, FIRST>(L, c, ta, tw, tb, x, rout, pout, vout, acc[0]);
#if ADG_KUNROLL == 2
    if (h2) ka_cell<TV, FIRST>(L2, c2, ta, tw, tb, x, rout, pout, vout, acc[0]);
#endif
  }
  block_partial<1, 0u, TAB == 0>(acc, partout, bv.b.slot);
  if (core_leader(bv)) {
    sc->rho_s[cur] = rho;
    if (!FIRST) sc->omega_s[prv] = omega_d;
  }
}

template <typename TV, int TAB>
__global__ void __launch_bounds__(NTF) __attribute__((amdgpu_waves_per_eu(1, StencilWaves<TV>::value)))
k_bicgB(LV lv, const TV *__restrict__ AT, int it, const TV *__restrict__ r, const TV *__restrict__ v,
        TV *__restrict__ t, const TV *__restrict__ rs, const double *__restrict__ pin, double *__restrict__ partout) {
#pragma clang fp contract(off)
  const BlkView bv = blk_view_t<TAB>(lv);
  Scal *__restrict__ sc = bv.sc;
  const long NC = lv.NC, sy = bv.b.sy, sz = bv.b.sz;
  const long off[7] = {-sz, -sy, -1, 0, 1, sy, sz};
  const long c0 = bv.b.lo + threadIdx.x;
  KBIn<TV> L;
  if (TAB == 0 && c0 < bv.b.hi) kb_load<TV>(L, AT, r, v, rs, NC, off, c0);
  if (sc->done) return;
  double red[1];
  reduce_partials<1, 0u, TAB == 0>(pin, bv.b.pfirst, bv.b.pn, red);
  const int cur = it & 1;
  const double alpha_d = sc->rho_s[cur] / red[0];
  const TV alpha = (TV)alpha_d;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long c = c0; c < bv.b.hi; c += ADG_KUNROLL * NTF) {
    if (TAB != 0 || c != c0) kb_load<TV>(L, AT, r, v, rs, NC, off, c);
#if ADG_KUNROLL == 2
    KBIn<TV> L2;
    const long c2 = c + NTF;
    const bool h2 = c2 < bv.b.hi;
    if (h2) kb_load<TV>(L2, AT, r, v, rs, NC, off, c2);
#endif
    kb_cell<TV>(L, c, alpha, t, acc);
#if ADG_KUNROLL == 2
    if (h2) kb_cell<TV>(L2, c2, alpha, t, acc);
#endif
  }
  block_partial<4, 0u, TAB == 0>(acc, partout, bv.b.slot);
  if (core_leader(bv)) sc->alpha_s[cur] = alpha_d;
}

// last iteration's x update: x += alpha p + omega s; for the FP32 form the
// accumulated correction is then added to the FP64 flux
template <typename TV, int TAB>
__global__ void __launch_bounds__(NTF) k_bicgT(LV lv, int it, const TV *r, const TV *p, const TV *v, TV *x,
                                              double *x64, const double *pin) {
#pragma clang fp contract(off)
  const BlkView bv = blk_view_t<TAB>(lv);
  if (bv.sc->done) return;
  double red[4];
  reduce_partials<4, 0u, TAB == 0>(pin, bv.b.pfirst, bv.b.pn, red);
  const double omega_d = red[1] / red[0];
  const TV ta = (TV)bv.sc->alpha_s[it & 1], tw = (TV)omega_d;
  for (long c = bv.b.lo + threadIdx.x; c < bv.b.hi; c += NTF) {
    const TV sv = (TV)(r[c] - ta * v[c]);
    const TV xn = (TV)(x[c] + ta * p[c] + tw * sv);
    if (x64) x64[c] = x64[c] + (double)xn;
    else x[c] = xn;
  }
}

// ---------------------------------------------------------------------------
// FSrc/FSrcAd (src/mod_cmfd.f90:956-1002), the error vector and its l2 norm
// (:480-481, l2norm :1100), RelEg (:1172) -- then extrapolation
// (fiss_extrp :308-335), Integrate (:1120), the k update (:485) and RelE
// (:1143).
// ---------------------------------------------------------------------------

// fs0 = FSrc(f0) for the cores whose outer call starts from the first-call
// state (every core with force != 0)
__global__ void __launch_bounds__(NTF) k_fsrc_init(LV lv, XS xs, int adjoint, const double *f0, double *fs0,
                                                  int force) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  if (!force && !v.sc->init) return;
  const long NC = lv.NC;
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double fs = 0.0;
    for (int g = 0; g < lv.G; ++g) {
      const double w = adjoint ? xs.chi[(size_t)g * NC + c] : xs.nuf[(size_t)g * NC + c];
      fs = fs + f0[(size_t)g * NC + c] * w;
    }
    fs0[c] = fs;
  }
}

// f = Integrate(fs0), e1 = Integrate(1) at the start of outer (:456-458):
// partials here, finished by k_fin_integrals
__global__ void __launch_bounds__(NTF) k_integrals(LV lv, XS xs, const double *fs0, double *partout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  double acc[2] = {0.0, 0.0};
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    acc[0] += xs.vdel[c] * fs0[c];
    acc[1] += xs.vdel[c];
  }
  block_partial<2, 0u>(acc, partout, v.b.slot);
}

// one block per core (blockIdx.x = core)
__global__ void __launch_bounds__(NTF) k_fin_integrals(LV lv, const double *pin) {
  Scal *sc = lv.sc + blockIdx.x;
  int first, n;
  core_slots(lv, blockIdx.x, first, n);
  double red[2];
  reduce_partials<2, 0u>(pin, first, n, red);
  if (threadIdx.x == 0) {
    sc->f = red[0];
    sc->e1 = red[1];
  }
}

// FSrc, errn = fs0 - fs0c, partials of |errn|^2 and of RelEg
__global__ void __launch_bounds__(NTF) k_fsrc(LV lv, XS xs, int adjoint, const double *f0, const double *f0c,
                                             double *fs0, const double *fs0c, double *errn, double *partout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  if (v.sc->done) return;
  const long NC = lv.NC;
  double acc[2] = {0.0, 0.0};   // sum errn^2, max flux rel. error
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double fs = 0.0;
    for (int g = 0; g < lv.G; ++g) {
      const double w = adjoint ? xs.chi[(size_t)g * NC + c] : xs.nuf[(size_t)g * NC + c];
      const double fg = f0[(size_t)g * NC + c];
      fs = fs + fg * w;
      if (fabs(fg) > 1.e-10) acc[1] = fmax(acc[1], fabs(fg - f0c[(size_t)g * NC + c]) / fabs(fg));
    }
    fs0[c] = fs;
    const double e = fs - fs0c[c];
    errn[c] = e;
    acc[0] += e * e;
  }
  block_partial<2, 2u>(acc, partout, v.b.slot);
}

// e2 = l2norm(errn) and the extrapolation (fiss_extrp, :308-335) when
// MOD(p, nac) == 0; partials of Integrate(fs0) and of RelE
__global__ void __launch_bounds__(NTF) k_post(LV lv, XS xs, double *fs0, const double *fs0c, const double *errn,
                                             const double *pin, double *partout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  Scal *sc = v.sc;
  if (sc->done) return;
  double red[2];
  reduce_partials<2, 2u>(pin, v.b.pfirst, v.b.pn, red);
  const double e2 = sqrt(red[0]);
  const int extrap = ((sc->p + 1) % sc->nac) == 0;
  double coef = 0.0;
  if (extrap) {
    const double domiR = e2 / sc->e1;
    coef = domiR / (1.0 - domiR);
  }
  double acc[2] = {0.0, 0.0};   // integral, max fission-source rel. error
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double fs = fs0[c];
    if (extrap) {
      fs = fs + coef * errn[c];
      fs0[c] = fs;
    }
    acc[0] += xs.vdel[c] * fs;
    if (fabs(fs) > 1.e-10) acc[1] = fmax(acc[1], fabs(fs - fs0c[c]) / fabs(fs));
  }
  block_partial<2, 2u>(acc, partout, v.b.slot);
  if (core_leader(v)) {
    sc->e2 = e2;
    sc->fer = red[1];
    sc->flags = extrap;
  }
}

// one block per core: k = k * f / fc (:484-485; not in the fixed-source
// outer, outer_fs :513-598, where k keeps its value), ser, e1 = e2
__global__ void __launch_bounds__(NTF) k_fin_outer(LV lv, const double *pin) {
  Scal *sc = lv.sc + blockIdx.x;
  if (sc->done) return;
  int first, n;
  core_slots(lv, blockIdx.x, first, n);
  double red[2];
  reduce_partials<2, 2u>(pin, first, n, red);
  if (threadIdx.x == 0) {
    const double fc = sc->f;
    sc->f = red[0];
    if (!sc->fixed_k) sc->Ke = sc->Ke * red[0] / fc;
    sc->ser = red[1];
    sc->e1 = sc->e2;
  }
}

// end of one outer, one thread per core: trace, error checks, convergence
// test (:491-496).  The trace arrays hold `cap` outers per core.
__global__ void k_end(Scal *scs, int ncore, int cap, double *tk, double *tser, double *tfer, double *tnd,
                      int *tflags, int *tloc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= ncore) return;
  Scal *sc = scs + b;
  if (sc->done) return;
  const int nodal = nodal_outer(sc) ? 1 : 0;
  const int p = sc->p + 1;
  const int flags = (sc->flags & 1) | (nodal ? 2 : 0);
  if (p <= cap) {
    const size_t t = (size_t)b * cap + (p - 1);
    tk[t] = sc->Ke;
    tser[t] = sc->ser;
    tfer[t] = sc->fer;
    tnd[t] = sc->ndmax;
    tflags[t] = flags;
    tloc[3 * t] = sc->im;
    tloc[3 * t + 1] = sc->jm;
    tloc[3 * t + 2] = sc->km;
  }
  sc->p = p;
  if (nodal && sc->lu_key != ~0ull) {
    sc->status = 3;
    sc->done = 1;
    return;
  }
  if (nodal && sc->ndmax > 1.e3) {
    sc->status = 2;
    sc->done = 1;
    return;
  }
  if (sc->ser < sc->serc && sc->fer < sc->ferc && sc->ndmax < 1.e-2) sc->done = 1;
}

// PowDis (src/mod_cmfd.f90:1290-1333): node powers, then normalisation
__global__ void __launch_bounds__(NTF) k_powdis(LV lv, XS xs, const double *f0, double *pw, double *partout) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(lv);
  const long NC = lv.NC;
  double acc[1] = {0.0};
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF) {
    double pc = 0.0;
    for (int g = 0; g < lv.G; ++g) {
      double w = f0[(size_t)g * NC + c] * xs.sigf[(size_t)g * NC + c] * xs.vdel[c];
      if (w < 0.0) w = 0.0;
      pc = pc + w;
    }
    pw[c] = pc;
    acc[0] += pc;
  }
  block_partial<1, 0u>(acc, partout, v.b.slot);
}

__global__ void __launch_bounds__(NTF) k_scale(LV lv, double *pw, const double *pin) {
  const BlkView v = blk_view(lv);
  double red[1];
  reduce_partials<1, 0u>(pin, v.b.pfirst, v.b.pn, red);
  const double tp = red[0];
  if (core_leader(v)) v.sc->tpow = tp;
  for (long c = v.b.lo + threadIdx.x; c < v.b.hi; c += NTF)
    pw[c] = pw[c] / tp;
}

}  // namespace adg
