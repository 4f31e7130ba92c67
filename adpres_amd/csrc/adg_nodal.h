// Nodal (SANM / PNM) two-node update of the D-hat coupling coefficients
// (src/mod_nodal.f90), reorganised for the GPU.
//
// In the reference every quantity a line sweep needs is a function of the
// state at the start of the update (flux, k, old D-hat, cross sections):
// the sweep only carries a2/a4/L1 of the previous node forward and writes
// each face's D-hat exactly once.  The update therefore splits into three
// data-parallel passes with identical arithmetic:
//   pass 0  per (node, group):      zeroth transverse leakages S1..S3 (get_source)
//   pass 1  per (node, direction):  B, A..H, Lm2 -> a2 (G x G LU), a4, L1 (TLUpd1)
//   pass 2  per (face, direction):  a1 from the 2G x 2G interface system (or the
//                                   G x G boundary system), a3, surface current,
//                                   new D-hat, max |change| with its location
// The argmax is reduced with ties broken by the reference's sweep order
// (x lines, then y, then z), so ndmax and its (i,j,k) match exactly.
#pragma once

#include "adg_common.h"

namespace adg {

// lv.blk is the per-node table for pass 0/1 and the per-face table for pass 2
struct NodArgs {
  LV lv;
  XS xs;
  const double *f0, *df;
  double *dn;
  double *S;            // [3][G][NC]
  double *a2, *a4, *L1; // [3][G][NC]
  int kern;             // 1 PNM, 2 SANM
  int cmode;            // 1 forward, 0 adjoint, 2 transient
  int sched;            // 1: run only in a nodal-update outer (device-side schedule)
  int coop;             // pass 2 leaves the interior faces to k_nod_pass2c
  double *part;
  // cooperative passes: the node terms of get_B's rows (S1 of the reference's
  // dum, [cell][g][h], made once per update by k_nod_bd) and each node's A,
  // F, G, H per direction ([u][cell][g][4], made by k_nod_pass1c), read by the
  // interface pass instead of being made again for every face; node-major,
  // so that a lane's row or coefficients are one contiguous read
  double *Bd, *cf4;
};

__device__ __forceinline__ double nod_dum(const NodArgs &a, int G, long NC, long c, double Ke, int g, int hh);

__device__ __forceinline__ long cell_of(const Geo &q, int i, int j, int k) {
  return q.base + ((long)(k + 1) * q.NYP + (j + 1)) * q.NXP + (i + 1);
}

__device__ __forceinline__ double hdir(const Geo &q, int u, int i, int j, int k) {
  return u == 0 ? q.hx[i] : (u == 1 ? q.hy[j] : q.hz[k]);
}

// ---------------------------------------------------------------------------
// pass 0: Lxyz + get_source (src/mod_nodal.f90:901-1043)
// ---------------------------------------------------------------------------

// Lxyz (src/mod_nodal.f90:901-1005): zeroth transverse leakages of cell c,
// group g, from the CMFD currents of its six faces
__device__ __forceinline__ void nod_lxyz(const Geo &q, const double *f, const double *df, const double *dn,
                                         unsigned char fl, long c, int i, int j, int k, int g, double (&L)[3]) {
#pragma clang fp contract(off)
  const long NC = q.NC;
  const int G = q.G;
  const double fn = f[c];
  const long st[3] = {1, q.sy, q.sz};
  for (int u = 0; u < 3; ++u) {
    const int sp = 2 * u, sm = 2 * u + 1;
    const double dfp = df[((size_t)sp * G + g) * NC + c], dnp = dn[((size_t)sp * G + g) * NC + c];
    const double dfm = df[((size_t)sm * G + g) * NC + c], dnm = dn[((size_t)sm * G + g) * NC + c];
    double jp, jm;
    if (!(fl & f_nb(sp))) {
      jp = (q.bc[sp] == 2) ? 0.0 : dfp * fn - dnp * fn;
    } else {
      const double fp = f[c + st[u]];
      jp = -dfp * (fp - fn) - dnp * (fp + fn);
    }
    if (!(fl & f_nb(sm))) {
      jm = (q.bc[sm] == 2) ? 0.0 : -dfm * fn - dnm * fn;
    } else {
      const double fm = f[c - st[u]];
      jm = -dfm * (fn - fm) - dnm * (fn + fm);
    }
    L[u] = (jp - jm) / hdir(q, u, i, j, k);
  }
}

// get_source: S_u = sum of the other two leakages; the transient form
// (cmode 2, :1031-1035) subtracts the external source here instead of in
// get_a2matvec
__global__ void __launch_bounds__(NT) k_nod_src(NodArgs a) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(a.lv);
  const Geo &q = *v.q;
  int i, j, k;
  long c;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = v.b.lo + threadIdx.x;
  if (t >= v.b.hi || t >= nxy * q.nzz) return;
  if (v.sc->done || (a.sched && !nodal_outer(v.sc))) return;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  c = cell_of(q, i, j, k);
  const unsigned char fl = q.flag[c];
  if (!(fl & F_ACTIVE)) return;
  const int g = blockIdx.y;
  const long NC = q.NC;
  const int G = q.G;
  double L[3];
  nod_lxyz(q, a.f0 + (size_t)g * NC, a.df, a.dn, fl, c, i, j, k, g, L);
  if (a.cmode == 2) {
    const double ex = a.xs.exsrc[(size_t)g * NC + c];
    a.S[((size_t)0 * G + g) * NC + c] = L[1] + L[2] - ex;
    a.S[((size_t)1 * G + g) * NC + c] = L[0] + L[2] - ex;
    a.S[((size_t)2 * G + g) * NC + c] = L[0] + L[1] - ex;
  } else {
    a.S[((size_t)0 * G + g) * NC + c] = L[1] + L[2];
    a.S[((size_t)1 * G + g) * NC + c] = L[0] + L[2];
    a.S[((size_t)2 * G + g) * NC + c] = L[0] + L[1];
  }
}

// get_B's node terms of every row of a node (S1 of the reference's dum,
// src/mod_nodal.f90:1345-1405), once per update for the cooperative passes,
// node-major: one thread per (node, row g), the node fastest, so that each
// cross-section load of a wavefront reads whole lines of consecutive cells
// (for G = 8 a wavefront holds rows 2w and 2w + 1 of 32 nodes, and a node's
// two rows fill one stored line).  The node-table entry of NT nodes is
// covered by blockIdx.y = 0 .. ceil(NT / (NT / G)) - 1 blocks of NT / G nodes.
template <int G>
__global__ void __launch_bounds__(NT) k_nod_bd(NodArgs a) {
#pragma clang fp contract(off)
  constexpr int NPB = NT / G;
  const BlkView v = blk_view(a.lv);
  const Geo &q = *v.q;
  if ((int)threadIdx.x >= NPB * G) return;
  const int local = (int)blockIdx.y * NPB + (int)threadIdx.x % NPB, g = (int)threadIdx.x / NPB;
  if (local >= NT) return;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = v.b.lo + local;
  if (t >= v.b.hi || t >= nxy * q.nzz) return;
  if (v.sc->done || (a.sched && !nodal_outer(v.sc))) return;
  int i, j, k;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  const long c = cell_of(q, i, j, k);
  if (!(q.flag[c] & F_ACTIVE)) return;
  const double Ke = v.sc->Ke;
  // the whole row first (every load before any store: the output may not
  // alias the inputs as far as the compiler knows), then stored at once
  double row[G];
#pragma unroll
  for (int hh = 0; hh < G; ++hh) row[hh] = nod_dum(a, G, q.NC, c, Ke, g, hh);
  double *out = a.Bd + ((size_t)c * G + g) * G;
#pragma unroll
  for (int hh = 0; hh < G; ++hh) out[hh] = row[hh];
}

// Lxyz of every node and group into out[u][g][cell] (the leakages the
// host's transient reactivity reads, src/mod_trans.f90:675-677)
__global__ void __launch_bounds__(NT) k_leak(NodArgs a, double *out) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(a.lv);
  const Geo &q = *v.q;
  int i, j, k;
  long c;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = v.b.lo + threadIdx.x;
  if (t >= v.b.hi || t >= nxy * q.nzz) return;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  c = cell_of(q, i, j, k);
  const unsigned char fl = q.flag[c];
  if (!(fl & F_ACTIVE)) return;
  const int g = blockIdx.y;
  const long NC = q.NC;
  const int G = q.G;
  double L[3];
  nod_lxyz(q, a.f0 + (size_t)g * NC, a.df, a.dn, fl, c, i, j, k, g, L);
  for (int u = 0; u < 3; ++u) out[((size_t)u * G + g) * NC + c] = L[u];
}

// ---------------------------------------------------------------------------
// per-node helpers
// ---------------------------------------------------------------------------

// get_B (src/mod_nodal.f90:1345-1405): the node term dum of entry (g, hh)
__device__ __forceinline__ double nod_dum(const NodArgs &a, int G, long NC, long c, double Ke, int g, int hh) {
#pragma clang fp contract(off)
  const XS &x = a.xs;
  double dum;
  if (a.cmode == 1) {
    if (g == hh) dum = x.sigr[(size_t)g * NC + c] - x.chi[(size_t)g * NC + c] * x.nuf[(size_t)hh * NC + c] / Ke;
    else dum = -x.sigs[((size_t)g * G + hh) * NC + c] - x.chi[(size_t)g * NC + c] * x.nuf[(size_t)hh * NC + c] / Ke;
  } else if (a.cmode == 2) {
    // transient (:1379-1391): fission weighted by 1 - beta + dfis, no 1/k;
    // the diagonal term carries chi twice, as the reference has it
    const double wc = x.wtr[c] * x.chi[(size_t)g * NC + c];
    if (g == hh) dum = x.sigr[(size_t)g * NC + c] - wc * x.chi[(size_t)g * NC + c] * x.nuf[(size_t)hh * NC + c];
    else dum = -x.sigs[((size_t)g * G + hh) * NC + c] - wc * x.nuf[(size_t)hh * NC + c];
  } else {
    if (g == hh) dum = x.sigr[(size_t)g * NC + c] - x.chi[(size_t)g * NC + c] * x.nuf[(size_t)hh * NC + c] / Ke;
    else dum = -x.sigs[((size_t)hh * G + g) * NC + c] - x.chi[(size_t)hh * NC + c] * x.nuf[(size_t)g * NC + c] / Ke;
  }
  return dum;
}

// get_B, row g: B(g, 0..G-1) = h^2 / (4 D_g) dum(g, .)
template <int G>
__device__ __forceinline__ void nod_B_row(const NodArgs &a, long NC, long c, double h, double Ke, int g, double *B) {
#pragma clang fp contract(off)
  const double D = a.xs.D[(size_t)g * NC + c];
  for (int hh = 0; hh < G; ++hh) {
    const double dum = nod_dum(a, G, NC, c, Ke, g, hh);
    B[hh] = 0.25 * (h * h) / D * dum;
  }
}

// the same row from the dum terms k_nod_bd stored (a.Bd)
template <int G>
__device__ __forceinline__ void nod_B_row_pre(const NodArgs &a, long NC, long c, double h, int g, double *B) {
#pragma clang fp contract(off)
  const double D = a.xs.D[(size_t)g * NC + c];
  const double fac = 0.25 * (h * h) / D;
#pragma unroll
  for (int hh = 0; hh < G; ++hh) B[hh] = fac * a.Bd[((size_t)c * G + g) * G + hh];
}

// get_B, all rows; B row-major [g*G+h]
template <int G>
__device__ __forceinline__ void nod_B(const NodArgs &a, long NC, long c, double h, double Ke, double *B) {
  for (int g = 0; g < G; ++g) nod_B_row<G>(a, NC, c, h, Ke, g, B + g * G);
}

// get_ABEFGH (src/mod_nodal.f90:1409-1451) or the PNM constants (:180-182)
struct Coef6 {
  double A, B, E, F, G, H;
};

__device__ __forceinline__ Coef6 nod_coef(int kern, double sigr, double D, double h) {
#pragma clang fp contract(off)
  Coef6 r;
  if (kern != 2) {
    r.A = 1.0 / 15.0; r.B = 1.0 / 35.0; r.E = 2.0 / 7.0; r.F = 2.0 / 5.0; r.G = 10.0; r.H = 6.0;
    return r;
  }
  const double alp = 0.5 * sqrt(sigr / D) * h;
  const double alp2 = alp * alp;
  const double sh = sinh(alp), ch = cosh(alp);
  const double m0c = sh / alp;
  const double m1s = 3.0 * (ch / alp - sh / alp2);
  const double m2c = 5.0 * (sh / alp - 3.0 * ch / alp2 + 3.0 * sh / (alp * alp * alp));
  r.A = (sh - m1s) / (alp2 * m1s);
  r.B = (ch - m0c - m2c) / (alp2 * m2c);
  r.E = (m0c / m2c - 3.0 / alp2);
  r.F = (alp * ch - m1s) / (alp2 * m1s);
  r.G = (alp * sh - 3.0 * m2c) / (ch - m0c - m2c);
  r.H = (alp * ch - m1s) / (sh - m1s);
  return r;
}

// LU_solve (src/mod_nodal.f90:829-897): Doolittle without pivoting, done in
// place (L below the diagonal) with the reference's operation order.
// Returns false if an input diagonal is below the reference's threshold.
template <int M>
__device__ __forceinline__ bool lu_solve(double *m, double *b) {
#pragma clang fp contract(off)
  bool ok = true;
  for (int i = 0; i < M; ++i)
    if (fabs(m[i * M + i]) < (double)10e-5f) ok = false;
  for (int i = 0; i < M; ++i) {
    for (int j = i + 1; j < M; ++j) {
      const double piv = m[j * M + i] / m[i * M + i];
      for (int k = i + 1; k < M; ++k) m[j * M + k] = m[j * M + k] - piv * m[i * M + k];
      m[j * M + i] = piv;
    }
  }
  // forward: y(i) = b(i) - sum_k<i L(i,k) y(k)
  for (int i = 1; i < M; ++i) {
    double isum = 0.0;
    for (int k = 0; k < i; ++k) isum = isum + m[i * M + k] * b[k];
    b[i] = b[i] - isum;
  }
  b[M - 1] = b[M - 1] / m[(M - 1) * M + (M - 1)];
  for (int i = M - 2; i >= 0; --i) {
    double isum = 0.0;
    for (int k = i + 1; k < M; ++k) isum = isum + m[i * M + k] * b[k];
    b[i] = (b[i] - isum) / m[i * M + i];
  }
  return ok;
}

// Neighbourhood of a node along direction u.
struct Line {
  bool first, last;    // no lower / no upper neighbour
  long cm, cp;         // neighbour cells (valid when present)
  double hm, hn, hp;
};

__device__ __forceinline__ Line line_at(const Geo &q, int u, int i, int j, int k, long c) {
  Line L;
  const unsigned char fl = q.flag[c];
  const long st = u == 0 ? 1 : (u == 1 ? q.sy : q.sz);
  L.first = !(fl & f_nb(2 * u + 1));
  L.last = !(fl & f_nb(2 * u));
  L.cm = c - st;
  L.cp = c + st;
  const double *h = u == 0 ? q.hx : (u == 1 ? q.hy : q.hz);
  const int ii = u == 0 ? i : (u == 1 ? j : k);
  L.hn = h[ii];
  L.hm = L.first ? 0.0 : h[ii - 1];
  L.hp = L.last ? 0.0 : h[ii + 1];
  return L;
}

// TLUpd1 (src/mod_nodal.f90:1047-1199) without the h^2/(4D) factor
__device__ __forceinline__ double tl1(const Line &L, const double *S, long c, int bclo, int bchi) {
#pragma clang fp contract(off)
  const double Sn = S[c];
  double tm, tp, p1m, p2m, p1p, p2p, hpp;
  if (L.first) {
    if (bclo == 2) {
      tm = 1.0; tp = L.hp / L.hn;
      p1m = tm + 1.0; p2m = 2.0 * tm + 1.0; p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      return (p1m * p2m * (S[L.cp] - Sn)) / hpp;
    }
    tp = L.hp / L.hn; p1p = tp + 1.0;
    return (S[L.cp] - Sn) / p1p;
  }
  if (L.last) {
    if (bchi == 2) {
      tm = L.hm / L.hn; tp = 1.0;
      p1m = tm + 1.0; p1p = tp + 1.0; p2p = 2.0 * tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      return (p1p * p2p * (Sn - S[L.cm])) / hpp;
    }
    tm = L.hm / L.hn; p1m = tm + 1.0;
    return (Sn - S[L.cm]) / p1m;
  }
  tm = L.hm / L.hn; tp = L.hp / L.hn;
  p1m = tm + 1.0; p2m = 2.0 * tm + 1.0; p1p = tp + 1.0; p2p = 2.0 * tp + 1.0;
  hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
  return (p1m * p2m * (S[L.cp] - Sn) + p1p * p2p * (Sn - S[L.cm])) / hpp;
}

// TLUpd2 (src/mod_nodal.f90:1203-1341) without the h^2/(4D) factor
__device__ __forceinline__ double tl2(const Line &L, const double *S, long c, int bclo, int bchi) {
#pragma clang fp contract(off)
  const double Sn = S[c];
  double tm, tp, p1m, p1p, hpp;
  if (L.first) {
    if (bclo == 2) {
      tm = 1.0; tp = L.hp / L.hn;
      p1m = tm + 1.0; p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      return (p1m * (S[L.cp] - Sn)) / hpp;
    }
    return 0.0;
  }
  if (L.last) {
    if (bchi == 2) {
      tm = L.hm / L.hn; tp = 1.0;
      p1m = tm + 1.0; p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      return (p1p * (S[L.cm] - Sn)) / hpp;
    }
    return 0.0;
  }
  tm = L.hm / L.hn; tp = L.hp / L.hn;
  p1m = tm + 1.0; p1p = tp + 1.0;
  hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
  return (p1m * (S[L.cp] - Sn) + p1p * (S[L.cm] - Sn)) / hpp;
}

__device__ __forceinline__ unsigned long long lu_key(unsigned long long order, int i, int j, int k) {
  return (order << 30) | ((unsigned long long)(i + 1) << 20) | ((unsigned long long)(j + 1) << 10) |
         (unsigned long long)(k + 1);
}

// Sweep-order index of the face below cell (i,j,k) along u, i.e. the face
// at which the reference handles this node as "p"; faces of one direction
// are numbered in the reference's line order, directions x < y < z.
__device__ __forceinline__ unsigned long long face_order(const Geo &q, int u, int fi, int j_or_i, int k_or_j) {
  const unsigned long long nx = q.nxx, ny = q.nyy, nz = q.nzz;
  const unsigned long long X = (nx + 1) * ny * nz, Y = (ny + 1) * nx * nz;
  if (u == 0) return ((unsigned long long)k_or_j * ny + j_or_i) * (nx + 1) + fi;          // (k, j, i)
  if (u == 1) return X + ((unsigned long long)k_or_j * nx + j_or_i) * (ny + 1) + fi;      // (k, i, j)
  return X + Y + ((unsigned long long)k_or_j * nx + j_or_i) * (nz + 1) + fi;              // (j, i, k)
}

// ---------------------------------------------------------------------------
// pass 1: per (node, direction) -- get_a2matvec (:769-825), LU_solve, get_a4
// (:740-765) and TLUpd1 -- blockIdx.y = direction
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(NT) k_nod_pass1(NodArgs a) {
#pragma clang fp contract(off)
  const BlkView v = blk_view(a.lv);
  const Geo &q = *v.q;
  Scal *sc = v.sc;
  const long nxy = (long)q.nxx * q.nyy;
  const long t = v.b.lo + threadIdx.x;
  if (t >= v.b.hi || t >= nxy * q.nzz) return;
  if (sc->done || (a.sched && !nodal_outer(sc))) return;
  int i, j, k;
  k = (int)(t / nxy);
  { long r = t - (long)k * nxy; j = (int)(r / q.nxx); i = (int)(r - (long)j * q.nxx); }
  const long c = cell_of(q, i, j, k);
  if (!(q.flag[c] & F_ACTIVE)) return;
  const int u = blockIdx.y;
  const long NC = q.NC;
  const Line L = line_at(q, u, i, j, k, c);
  const int bclo = q.bc[2 * u + 1], bchi = q.bc[2 * u];
  if (L.first && L.last) {   // single-node line: the reference reads an undefined neighbour
    atomicMin(&sc->lu_key, lu_key(0, i, j, k));
    return;
  }
  const double h = L.hn, Ke = sc->Ke;
  double B[G * G], M[G * G], rhs[G], Lm2[G];
  Coef6 cf[G];
  nod_B<G>(a, NC, c, h, Ke, B);
  for (int g = 0; g < G; ++g) cf[g] = nod_coef(a.kern, a.xs.sigr[(size_t)g * NC + c], a.xs.D[(size_t)g * NC + c], h);
  const double *Su = a.S + (size_t)u * G * NC;
  for (int g = 0; g < G; ++g) {
    const double D = a.xs.D[(size_t)g * NC + c];
    // get_a2matvec (:788-805): the external source enters here except in the
    // transient form, where get_source has taken it out of S already
    const double S = a.cmode == 2 ? 0.25 * (h * h) / D * Su[(size_t)g * NC + c]
                                  : 0.25 * (h * h) / D * (Su[(size_t)g * NC + c] - a.xs.exsrc[(size_t)g * NC + c]);
    double Bf = 0.0;
    for (int hh = 0; hh < G; ++hh) {
      M[g * G + hh] = (hh == g) ? B[g * G + hh] * cf[g].E + 3.0 : B[g * G + hh] * cf[g].E;
      Bf = Bf + B[g * G + hh] * a.f0[(size_t)hh * NC + c];
    }
    Lm2[g] = 0.25 * (h * h) / D * tl2(L, Su + (size_t)g * NC, c, bclo, bchi);
    rhs[g] = Bf - cf[g].E * Lm2[g] + S;
  }
  if (!lu_solve<G>(M, rhs)) {
    const int fi = u == 0 ? i : (u == 1 ? j : k);
    const int o1 = u == 0 ? j : i, o2 = u == 2 ? j : k;
    atomicMin(&sc->lu_key, lu_key(face_order(q, u, fi, o1, o2), i, j, k));
  }
  for (int g = 0; g < G; ++g) {
    double Bf = 0.0;
    for (int hh = 0; hh < G; ++hh) Bf = Bf + B[g * G + hh] * rhs[hh];
    const size_t o = ((size_t)u * G + g) * NC + c;
    a.a2[o] = rhs[g];
    a.a4[o] = cf[g].B * (Bf + Lm2[g]);
    const double D = a.xs.D[(size_t)g * NC + c];
    a.L1[o] = 0.25 * (h * h) / D * tl1(L, Su + (size_t)g * NC, c, bclo, bchi);
  }
}

// ---------------------------------------------------------------------------
// pass 2: per face -- get_a1matvec(_first/_last) (:483-698), LU_solve,
// get_a3 (:702-736) and nodal_coup_upd (:282-371).  blockIdx.y = direction.
// ---------------------------------------------------------------------------
// One face of direction u, face index t of that direction (faces of
// direction u: (n_u + 1) x (other two extents), the face index fastest):
// the new D-hat of the face and the running argmax (best, bkey, bloc).
template <int G>
__device__ __forceinline__ void face_update(const NodArgs &a, const Geo &q, Scal *sc, int u, long t, double &best,
                                            double &bkey, double &bloc) {
#pragma clang fp contract(off)
  const long NC = q.NC;
  const int nu = u == 0 ? q.nxx : (u == 1 ? q.nyy : q.nzz);
  const int e1 = u == 0 ? q.nyy : q.nxx;          // second index extent (j for x; i for y, z)
  {
    const int fi = (int)(t % (nu + 1));
    const long r = t / (nu + 1);
    const int o1 = (int)(r % e1), o2 = (int)(r / e1);
    int i, j, k;
    if (u == 0) { i = fi; j = o1; k = o2; }
    else if (u == 1) { i = o1; j = fi; k = o2; }
    else { i = o1; j = o2; k = fi; }
    const long st = u == 0 ? 1 : (u == 1 ? q.sy : q.sz);
    const long cp = cell_of(q, i, j, k);    // node above the face ("p")
    const long cn = cp - st;                // node below the face ("n")
    const bool hasp = (q.flag[cp] & F_ACTIVE) != 0;
    const bool hasn = (q.flag[cn] & F_ACTIVE) != 0;
    if ((hasp || hasn) && !(a.coop && hasp && hasn)) {
      const double Ke = sc->Ke;
      const int sf = 2 * u;                 // upper face of n (x+, y+, z+), sf+1 lower face of p
      const unsigned long long ord = face_order(q, u, fi, o1, o2);
      const size_t ub = (size_t)u * G * NC;
      int li = i, lj = j, lk = k;           // location reported: n, or p for a first node
      if (hasn) { if (u == 0) li = i - 1; else if (u == 1) lj = j - 1; else lk = k - 1; }
      double a1[G], a3[G];
      Coef6 cfn[G], cfp[G];
      double Bn[G * G], Bp[G * G];
      if (hasn && hasp) {
        const double hn = hdir(q, u, u == 0 ? i - 1 : i, u == 1 ? j - 1 : j, u == 2 ? k - 1 : k);
        const double hp = hdir(q, u, i, j, k);
        nod_B<G>(a, NC, cn, hn, Ke, Bn);
        nod_B<G>(a, NC, cp, hp, Ke, Bp);
        for (int g = 0; g < G; ++g) {
          cfn[g] = nod_coef(a.kern, a.xs.sigr[(size_t)g * NC + cn], a.xs.D[(size_t)g * NC + cn], hn);
          cfp[g] = nod_coef(a.kern, a.xs.sigr[(size_t)g * NC + cp], a.xs.D[(size_t)g * NC + cp], hp);
        }
        constexpr int M2 = 2 * G;
        double R[M2 * M2], s[M2];
        for (int g = 0; g < G; ++g) {
          const double Ln1 = a.L1[ub + (size_t)g * NC + cn], Lp1 = a.L1[ub + (size_t)g * NC + cp];
          const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
          const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
          const double Pn = 2.0 * a.xs.D[(size_t)g * NC + cn] / hn;
          const double Pp = 2.0 * a.xs.D[(size_t)g * NC + cp] / hp;
          for (int hh = 0; hh < G; ++hh) {
            if (hh == g) {
              R[g * M2 + g] = -Pn * (Bn[g * G + hh] * cfn[g].F + 1.0);
              R[g * M2 + g + G] = Pp * (Bp[g * G + hh] * cfp[g].F + 1.0);
            } else {
              R[g * M2 + hh] = -Pn * Bn[g * G + hh] * cfn[g].F;
              R[g * M2 + hh + G] = Pp * Bp[g * G + hh] * cfp[g].F;
            }
          }
          s[g] = Pn * (3.0 * a2n + cfn[g].G * a4n + cfn[g].F * Ln1) + Pp * (3.0 * a2p + cfp[g].G * a4p - cfp[g].F * Lp1);
        }
        for (int g = 0; g < G; ++g) {
          const double dcn = a.xs.dc[((size_t)sf * G + g) * NC + cn];
          const double dcp = a.xs.dc[((size_t)(sf + 1) * G + g) * NC + cp];
          const double Ln1 = a.L1[ub + (size_t)g * NC + cn], Lp1 = a.L1[ub + (size_t)g * NC + cp];
          const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
          const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
          for (int hh = 0; hh < G; ++hh) {
            if (hh == g) {
              R[(g + G) * M2 + g] = dcn * (Bn[g * G + hh] * cfn[g].A + 1.0);
              R[(g + G) * M2 + g + G] = dcp * (Bp[g * G + hh] * cfp[g].A + 1.0);
            } else {
              R[(g + G) * M2 + hh] = dcn * Bn[g * G + hh] * cfn[g].A;
              R[(g + G) * M2 + hh + G] = dcp * Bp[g * G + hh] * cfp[g].A;
            }
          }
          // the reference pairs An with Ln1 and Ap with Lp1 here (:693-694)
          s[g + G] = dcp * (a2p + a4p + a.f0[(size_t)g * NC + cp] - cfn[g].A * Ln1) -
                     dcn * (a2n + a4n + a.f0[(size_t)g * NC + cn] + cfp[g].A * Lp1);
        }
        if (!lu_solve<M2>(R, s)) atomicMin(&sc->lu_key, lu_key(ord, li, lj, lk));
        for (int g = 0; g < G; ++g) a1[g] = s[g];
        for (int g = 0; g < G; ++g) {
          double Bf = 0.0;
          for (int hh = 0; hh < G; ++hh) Bf = Bf + Bn[g * G + hh] * a1[hh];
          a3[g] = cfn[g].A * (Bf + a.L1[ub + (size_t)g * NC + cn]);
        }
        for (int g = 0; g < G; ++g) {
          const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
          const double jp = -2.0 * a.xs.D[(size_t)g * NC + cn] / hn *
                            (a1[g] + 3.0 * a2n + cfn[g].H * a3[g] + cfn[g].G * a4n);
          const size_t on = ((size_t)sf * G + g) * NC + cn, op = ((size_t)(sf + 1) * G + g) * NC + cp;
          const double ndpr = a.dn[on];
          const double fn = a.f0[(size_t)g * NC + cn], fp = a.f0[(size_t)g * NC + cp];
          const double v = (a.df[on] * (fn - fp) - jp) / (fn + fp);
          a.dn[on] = v;
          a.dn[op] = v;
          const double nder = fabs(v - ndpr);
          if (nder > best) { best = nder; bkey = (double)(ord * G + g); }
        }
      } else if (hasp) {
        // first node of a line (get_coefs_first, :375-407)
        const double hp = hdir(q, u, i, j, k);
        nod_B<G>(a, NC, cp, hp, Ke, Bp);
        for (int g = 0; g < G; ++g)
          cfp[g] = nod_coef(a.kern, a.xs.sigr[(size_t)g * NC + cp], a.xs.D[(size_t)g * NC + cp], hp);
        const int bc = q.bc[2 * u + 1];
        double Mx[G * G], b[G];
        for (int g = 0; g < G; ++g) {
          const double Lp1 = a.L1[ub + (size_t)g * NC + cp];
          const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
          const double Pp = 2.0 * a.xs.D[(size_t)g * NC + cp] / hp;
          const double dc = a.xs.dc[((size_t)(sf + 1) * G + g) * NC + cp];
          const double Ap = cfp[g].A, Fp = cfp[g].F, Gp = cfp[g].G, Hp = cfp[g].H;
          const double f0p = a.f0[(size_t)g * NC + cp];
          for (int hh = 0; hh < G; ++hh) {
            const double Bv = Bp[g * G + hh];
            if (bc == 2) Mx[g * G + hh] = (hh == g) ? Pp * (Bv * Fp + 1.0) : Pp * Bv * Fp;
            else if (bc == 1)
              Mx[g * G + hh] = (hh == g) ? -dc * (1.0 + Ap * Bv) - 2.0 * Pp * (Ap * Bv * Hp + 1.0)
                                         : -dc * Ap * Bv - 2.0 * Pp * Ap * Bv * Hp;
            else Mx[g * G + hh] = (hh == g) ? dc * (1.0 + Ap * Bv) : dc * Ap * Bv;
          }
          if (bc == 2) b[g] = Pp * (3.0 * a2p + Gp * a4p - Fp * Lp1);
          else if (bc == 1) b[g] = 2.0 * Pp * (Ap * Hp * Lp1 - 3.0 * a2p - Gp * a4p) - dc * (a2p + a4p + f0p - Ap * Lp1);
          else b[g] = dc * (a2p + a4p + f0p - Ap * Lp1);
        }
        if (!lu_solve<G>(Mx, b)) atomicMin(&sc->lu_key, lu_key(ord, li, lj, lk));
        for (int g = 0; g < G; ++g) a1[g] = b[g];
        for (int g = 0; g < G; ++g) {
          double Bf = 0.0;
          for (int hh = 0; hh < G; ++hh) Bf = Bf + Bp[g * G + hh] * a1[hh];
          a3[g] = cfp[g].A * (Bf + a.L1[ub + (size_t)g * NC + cp]);
        }
        for (int g = 0; g < G; ++g) {
          const double a2p = a.a2[ub + (size_t)g * NC + cp], a4p = a.a4[ub + (size_t)g * NC + cp];
          const double jp = -2.0 * a.xs.D[(size_t)g * NC + cp] / hp *
                            (a1[g] - 3.0 * a2p + cfp[g].H * a3[g] - cfp[g].G * a4p);
          const size_t op = ((size_t)(sf + 1) * G + g) * NC + cp;
          const double ndpr = a.dn[op];
          const double v = -(jp / a.f0[(size_t)g * NC + cp] + a.df[op]);
          a.dn[op] = v;
          const double nder = fabs(v - ndpr);
          if (nder > best) { best = nder; bkey = (double)(ord * G + g); }
        }
      } else {
        // last node of a line (get_coefs_last, :411-441)
        const int ni = u == 0 ? i - 1 : i, nj = u == 1 ? j - 1 : j, nk = u == 2 ? k - 1 : k;
        const double hn = hdir(q, u, ni, nj, nk);
        nod_B<G>(a, NC, cn, hn, Ke, Bn);
        for (int g = 0; g < G; ++g)
          cfn[g] = nod_coef(a.kern, a.xs.sigr[(size_t)g * NC + cn], a.xs.D[(size_t)g * NC + cn], hn);
        const int bc = q.bc[2 * u];
        double Mx[G * G], b[G];
        for (int g = 0; g < G; ++g) {
          const double Ln1 = a.L1[ub + (size_t)g * NC + cn];
          const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
          const double Pn = 2.0 * a.xs.D[(size_t)g * NC + cn] / hn;
          const double dc = a.xs.dc[((size_t)sf * G + g) * NC + cn];
          const double An = cfn[g].A, Fn = cfn[g].F, Gn = cfn[g].G, Hn = cfn[g].H;
          const double f0n = a.f0[(size_t)g * NC + cn];
          for (int hh = 0; hh < G; ++hh) {
            const double Bv = Bn[g * G + hh];
            if (bc == 2) Mx[g * G + hh] = (hh == g) ? -Pn * (Bv * Fn + 1.0) : -Pn * Bv * Fn;
            else if (bc == 1)
              Mx[g * G + hh] = (hh == g) ? dc * (1.0 + An * Bv) + 2.0 * Pn * (An * Bv * Hn + 1.0)
                                         : dc * An * Bv + 2.0 * Pn * An * Bv * Hn;
            else Mx[g * G + hh] = (hh == g) ? dc * (1.0 + An * Bv) : dc * An * Bv;
          }
          if (bc == 2) b[g] = Pn * (3.0 * a2n + Gn * a4n + Fn * Ln1);
          else if (bc == 1) b[g] = -2.0 * Pn * (An * Hn * Ln1 + 3.0 * a2n + Gn * a4n) - dc * (a2n + a4n + f0n + An * Ln1);
          else b[g] = -dc * (a2n + a4n + f0n + An * Ln1);
        }
        if (!lu_solve<G>(Mx, b)) atomicMin(&sc->lu_key, lu_key(ord, li, lj, lk));
        for (int g = 0; g < G; ++g) a1[g] = b[g];
        for (int g = 0; g < G; ++g) {
          double Bf = 0.0;
          for (int hh = 0; hh < G; ++hh) Bf = Bf + Bn[g * G + hh] * a1[hh];
          a3[g] = cfn[g].A * (Bf + a.L1[ub + (size_t)g * NC + cn]);
        }
        for (int g = 0; g < G; ++g) {
          const double a2n = a.a2[ub + (size_t)g * NC + cn], a4n = a.a4[ub + (size_t)g * NC + cn];
          const double jp = -2.0 * a.xs.D[(size_t)g * NC + cn] / hn *
                            (a1[g] + 3.0 * a2n + cfn[g].H * a3[g] + cfn[g].G * a4n);
          const size_t on = ((size_t)sf * G + g) * NC + cn;
          const double ndpr = a.dn[on];
          const double v = -(jp / a.f0[(size_t)g * NC + cn] - a.df[on]);
          a.dn[on] = v;
          const double nder = fabs(v - ndpr);
          if (nder > best) { best = nder; bkey = (double)(ord * G + g); }
        }
      }
      bloc = (double)(((li + 1) << 20) | ((lj + 1) << 10) | (lk + 1));
    }
  }
}

// block argmax (largest value, then smallest sweep key) into the block's
// partial slot
__device__ __forceinline__ void face_argmax(double best, double bkey, double bloc, double *slot3) {
  __shared__ double sv[NT], sk[NT], sl[NT];
  sv[threadIdx.x] = best;
  sk[threadIdx.x] = bkey;
  sl[threadIdx.x] = bloc;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int o = threadIdx.x + w;
      if (sv[o] > sv[threadIdx.x] || (sv[o] == sv[threadIdx.x] && sk[o] < sk[threadIdx.x])) {
        sv[threadIdx.x] = sv[o]; sk[threadIdx.x] = sk[o]; sl[threadIdx.x] = sl[o];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    slot3[0] = sv[0];
    slot3[1] = sk[0];
    slot3[2] = sl[0];
  }
}

// every face of the launch's face range (blockIdx.y = direction), one per
// thread: the form for smaller group counts (the cooperative kernels serve
// G >= 5)
template <int G>
__global__ void __launch_bounds__(NT) k_nod_pass2(NodArgs a) {
  const BlkView v = blk_view(a.lv);
  const Geo &q = *v.q;
  Scal *sc = v.sc;
  if (a.sched && !nodal_outer(sc)) return;   // uniform over the core's blocks
  const int u = blockIdx.y;
  const int nu = u == 0 ? q.nxx : (u == 1 ? q.nyy : q.nzz);
  const int e1 = u == 0 ? q.nyy : q.nxx;
  const int e2 = u == 2 ? q.nyy : q.nzz;
  const long nf = (long)(nu + 1) * e1 * e2;
  const long t = v.b.lo + threadIdx.x;
  double best = -1.0, bkey = 0.0, bloc = 0.0;
  if (t < nf && !sc->done) face_update<G>(a, q, sc, u, t, best, bkey, bloc);
  face_argmax(best, bkey, bloc, a.part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 3);
}

// ---------------------------------------------------------------------------
// LU_solve (:829-897) of an M x M system held one row per lane: each item
// (node or face) of a wavefront has M (or more) lanes, one per row.  The
// elimination broadcasts each pivot row to the item's lanes and the
// substitutions each solved unknown, so every entry sees the operations of
// the per-thread lu_solve in the same order and no lane holds more than its
// own row.  On return every lane of the item holds the solution in xs.
// Lanes with r >= M take part in the broadcasts only.
//
// The broadcasts are DPP moves within 16-lane rows of the wavefront (no LDS
// round trip in the elimination's dependent chain).  Two lane layouts (L):
//   LaneRow16: lane 16 f + r, item f is one row; row p of it is broadcast
//              by row_newbcast:p;
//   LaneRow8:  lane 8 f + r, items 2m and 2m + 1 share a row; row_newbcast:p
//              and row_newbcast:8+p, each lane keeping its own item's.
// An item's lanes are active or inactive together, so a lane reads only
// lanes of its own item.
// ---------------------------------------------------------------------------

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&fn) {
  if constexpr (B < E) {
    fn(std::integral_constant<int, B>{});
    static_for<B + 1, E>(fn);
  }
}

struct LaneRow16 {
  template <int P>
  static __device__ __forceinline__ float bcast(float v, int) {
    static_assert(P >= 0 && P < 16, "row lane");
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + P, 0xf, 0xf, false));
  }
  template <int P>
  static __device__ __forceinline__ double bcast(double v, int) {
    static_assert(P >= 0 && P < 16, "row lane");
    // one 64-bit move (v_mov_b64_dpp) rather than two 32-bit halves
    return __longlong_as_double(__builtin_amdgcn_mov_dpp(__double_as_longlong(v), 0x150 + P, 0xf, 0xf, false));
  }
};

struct LaneRow8 {
  template <int P, typename T>
  static __device__ __forceinline__ T bcast(T v, int f) {
    static_assert(P >= 0 && P < 8, "half-row lane");
    const T lo = LaneRow16::template bcast<P>(v, f);
    const T hi = LaneRow16::template bcast<8 + P>(v, f);
    return (f & 1) ? hi : lo;
  }
};

// the reference's input-diagonal test for this lane's row
template <int M>
__device__ __forceinline__ bool coop_lu_diag_small(const double (&R)[M], int r) {
  double dg = 0.0;
#pragma unroll
  for (int c2 = 0; c2 < M; ++c2)
    if (c2 == r) dg = R[c2];
  return r < M && fabs(dg) < (double)10e-5f;
}

// the elimination: R becomes L (below the diagonal) and U (on and above it)
template <int M, typename L, typename T>
__device__ __forceinline__ void coop_lu_factor(T (&R)[M], int r, int f) {
#pragma clang fp contract(off)
  const bool rowlane = r < M;
  static_for<0, M>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    // pivot row p (its entries p..M-1, final after the steps before p; lane
    // p's row does not change in this step), each entry broadcast where it
    // is used; the broadcasts run on every lane, the updates on rows r > p
    const bool upd = r > p && rowlane;
    const T piv = R[p] / L::template bcast<p>(R[p], f);
#pragma unroll
    for (int c2 = p + 1; c2 < M; ++c2) {
      const T e = L::template bcast<p>(R[c2], f);
      const T nv = R[c2] - piv * e;
      R[c2] = upd ? nv : R[c2];
    }
    R[p] = upd ? piv : R[p];
  });
}

// the substitutions with those factors, for right-hand side s (this lane's row)
template <int M, typename L, typename T>
__device__ __forceinline__ void coop_lu_subst(const T (&R)[M], T s, int r, int f, T (&xs)[M]) {
#pragma clang fp contract(off)
  // forward substitution y(i) = b(i) - sum_k<i L(i,k) y(k), k ascending
  T isum = (T)0, y = (T)0;
  static_for<0, M>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    if (r == p) y = s - isum;
    const T yp = L::template bcast<p>(y, f);
    if (r > p) isum = isum + R[p] * yp;
  });
  // back substitution x(i) = (y(i) - sum_k>i U(i,k) x(k)) / U(i,i), k ascending
  static_for<0, M>([&](auto qc) {
    constexpr int p = M - 1 - decltype(qc)::value;
    T xv;
    if constexpr (p == M - 1) {
      xv = y / R[p];
    } else {
      T bs = (T)0;
#pragma unroll
      for (int c2 = p + 1; c2 < M; ++c2) bs = bs + R[c2] * xs[c2];
      xv = (y - bs) / R[p];
    }
    xs[p] = L::template bcast<p>(xv, f);
  });
}

template <int M, typename L, typename T>
__device__ __forceinline__ void coop_lu(T (&R)[M], T s, int r, int f, T (&xs)[M]) {
  coop_lu_factor<M, L, T>(R, r, f);
  coop_lu_subst<M, L, T>(R, s, r, f, xs);
}

// The FP32+refine form's nodal solves: the elimination and the solve in
// FP32, then one FP64 defect correction -- the residual s - R x of the FP64
// system formed in FP64, its correction solved with the FP32 factors and
// added in FP64.  R (this lane's FP64 row) is left as it was.
template <int M, typename L>
__device__ __forceinline__ void coop_lu_refine(const double (&R)[M], double s, int r, int f, double (&xs)[M]) {
#pragma clang fp contract(off)
  float Rf[M], xf[M];
#pragma unroll
  for (int c2 = 0; c2 < M; ++c2) Rf[c2] = (float)R[c2];
  coop_lu_factor<M, L, float>(Rf, r, f);
  coop_lu_subst<M, L, float>(Rf, (float)s, r, f, xf);
  double res = s;
#pragma unroll
  for (int c2 = 0; c2 < M; ++c2) {
    xs[c2] = (double)xf[c2];
    res = res - R[c2] * xs[c2];
  }
  coop_lu_subst<M, L, float>(Rf, (float)res, r, f, xf);
#pragma unroll
  for (int c2 = 0; c2 < M; ++c2) xs[c2] = xs[c2] + (double)xf[c2];
}

// Direction-interleaved grid of the cooperative passes: block B of a 1D grid
// of 3 nbp blocks (nbp: the table's blocks rounded up to a multiple of 8)
// serves table block (B / 24) 8 + B % 8 in direction (B / 8) % 3.  Blocks go
// to the 8 XCDs round-robin, so B, B + 8 and B + 16 -- the three directions
// of one node or face range -- run at nearly the same time on one XCD, and
// the node terms all three read come from that XCD's L2 after the first.
struct DirBlk {
  int u, bx;
};
__device__ __forceinline__ DirBlk dir_blk() {
  const int B = (int)blockIdx.x;
  return DirBlk{(B >> 3) % 3, (B / 24) * 8 + (B & 7)};
}
__device__ __forceinline__ BlkView blk_view_at(const LV &lv, int bx) {
  BlkView v;
  v.b = lv.blk[bx];
  v.q = lv.geo + v.b.core;
  v.sc = lv.sc + v.b.core;
  return v;
}

// ---------------------------------------------------------------------------
// pass 1, cooperative form for larger group counts: eight lanes per (node,
// direction), lane 8 f + r owning row r of node f's G x G system of
// get_a2matvec (eight neighbouring nodes per wavefront, LaneRow8).  Same arithmetic,
// in the same order, as k_nod_pass1.
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(NT) k_nod_pass1c(NodArgs a, int nbx) {
#pragma clang fp contract(off)
  static_assert(G <= 8, "eight lanes per node");
  const DirBlk db = dir_blk();
  if (db.bx >= nbx) return;                                  // the grid's padding blocks
  const BlkView v = blk_view_at(a.lv, db.bx);
  const Geo &q = *v.q;
  Scal *sc = v.sc;
  if (sc->done || (a.sched && !nodal_outer(sc))) return;   // uniform over the core's blocks
  const int u = db.u;
  const long NC = q.NC;
  const long nxy = (long)q.nxx * q.nyy;
  const int lane = threadIdx.x & 63, r = lane & 7, f = lane >> 3;
  const int slot32 = (int)(threadIdx.x >> 6) * 8 + f;
  const int g = r < G ? r : 0;   // lanes past the last row repeat row 0's loads
  for (int it = 0; it < NT / 32; ++it) {
    const long t = v.b.lo + (long)it * 32 + slot32;
    if (!(t < v.b.hi && t < nxy * q.nzz)) continue;   // uniform over the node's eight lanes
    const int k = (int)(t / nxy);
    const long rr = t - (long)k * nxy;
    const int j = (int)(rr / q.nxx), i = (int)(rr - (long)j * q.nxx);
    const long c = cell_of(q, i, j, k);
    if (!(q.flag[c] & F_ACTIVE)) continue;
    const Line L = line_at(q, u, i, j, k, c);
    const int bclo = q.bc[2 * u + 1], bchi = q.bc[2 * u];
    if (L.first && L.last) {   // single-node line: the reference reads an undefined neighbour
      if (r == 0) atomicMin(&sc->lu_key, lu_key(0, i, j, k));
      continue;
    }
    const double h = L.hn;
    double B[G], R[G];
    nod_B_row_pre<G>(a, NC, c, h, g, B);
    const Coef6 cf = nod_coef(a.kern, a.xs.sigr[(size_t)g * NC + c], a.xs.D[(size_t)g * NC + c], h);
    if (r < G) {   // A, F, G, H of this node and direction for the interface pass
      double *cq = a.cf4 + (((size_t)u * NC + c) * G + g) * 4;
      cq[0] = cf.A;
      cq[1] = cf.F;
      cq[2] = cf.G;
      cq[3] = cf.H;
    }
    const double *Su = a.S + (size_t)u * G * NC;
    const double D = a.xs.D[(size_t)g * NC + c];
    const double S = a.cmode == 2 ? 0.25 * (h * h) / D * Su[(size_t)g * NC + c]
                                  : 0.25 * (h * h) / D * (Su[(size_t)g * NC + c] - a.xs.exsrc[(size_t)g * NC + c]);
    double Bf = 0.0;
#pragma unroll
    for (int hh = 0; hh < G; ++hh) {
      R[hh] = (hh == g) ? B[hh] * cf.E + 3.0 : B[hh] * cf.E;
      Bf = Bf + B[hh] * a.f0[(size_t)hh * NC + c];
    }
    const double Lm2 = 0.25 * (h * h) / D * tl2(L, Su + (size_t)g * NC, c, bclo, bchi);
    const double rhs = Bf - cf.E * Lm2 + S;
    if (coop_lu_diag_small<G>(R, r)) {
      const int fi = u == 0 ? i : (u == 1 ? j : k);
      const int o1 = u == 0 ? j : i, o2 = u == 2 ? j : k;
      atomicMin(&sc->lu_key, lu_key(face_order(q, u, fi, o1, o2), i, j, k));
    }
    double xs[G];
    coop_lu<G, LaneRow8>(R, rhs, r, f, xs);
    if (r < G) {
      double Bx = 0.0;
#pragma unroll
      for (int hh = 0; hh < G; ++hh) Bx = Bx + B[hh] * xs[hh];
      const size_t o = ((size_t)u * G + g) * NC + c;
      a.a2[o] = xs[g];
      a.a4[o] = cf.B * (Bx + Lm2);
      a.L1[o] = 0.25 * (h * h) / D * tl1(L, Su + (size_t)g * NC, c, bclo, bchi);
    }
  }
}

// ---------------------------------------------------------------------------
// pass 2, boundary faces (one node active), cooperative form: the faces
// listed in bfl ((direction << 40) | face index, per core, NT per block),
// eight lanes per face, lane 8 f + r owning row r of face f's G x G system of
// get_a1matvec_first (the active node above the face, :375-407) or
// get_a1matvec_last (below it, :411-441), solved by coop_lu; then get_a3
// and nodal_coup_upd's boundary forms (:330-368).  Same arithmetic, in the
// same order, as the per-thread face_update.  The argmax partials go after
// the 3 x nfb slots of the face launches.
// ---------------------------------------------------------------------------
template <int G>
__global__ void __launch_bounds__(NT) k_nod_pass2b(NodArgs a, const long *bfl, int nfb) {
#pragma clang fp contract(off)
  static_assert(G <= 8, "eight lanes per face");
  const BlkView v = blk_view(a.lv);
  const Geo &q = *v.q;
  Scal *sc = v.sc;
  if (a.sched && !nodal_outer(sc)) return;   // uniform over the core's blocks
  const long NC = q.NC;
  const int lane = threadIdx.x & 63, r = lane & 7, f = lane >> 3;
  const int slot32 = (int)(threadIdx.x >> 6) * 8 + f;
  const int g = r < G ? r : 0;   // lanes past the last row repeat row 0's loads
  double best = -1.0, bkey = 0.0, bloc = 0.0;
  for (int it = 0; it < NT / 32; ++it) {
    const long t = v.b.lo + (long)it * 32 + slot32;
    if (!(t < v.b.hi) || sc->done) continue;   // uniform over the face's eight lanes
    const long e = bfl[t];
    const int u = (int)(e >> 40);
    const long tf = e & ((1L << 40) - 1);
    const int nu = u == 0 ? q.nxx : (u == 1 ? q.nyy : q.nzz);
    const int e1 = u == 0 ? q.nyy : q.nxx;
    const int fi = (int)(tf % (nu + 1));
    const long rr = tf / (nu + 1);
    const int o1 = (int)(rr % e1), o2 = (int)(rr / e1);
    int i, j, k;
    if (u == 0) { i = fi; j = o1; k = o2; }
    else if (u == 1) { i = o1; j = fi; k = o2; }
    else { i = o1; j = o2; k = fi; }
    const long st = u == 0 ? 1 : (u == 1 ? q.sy : q.sz);
    const long cp = cell_of(q, i, j, k);
    const long cn = cp - st;
    const bool hasp = (q.flag[cp] & F_ACTIVE) != 0;   // else the node below (n) is the active one
    const int sf = 2 * u;
    const unsigned long long ord = face_order(q, u, fi, o1, o2);
    const size_t ub = (size_t)u * G * NC;
    int li = i, lj = j, lk = k;
    if (!hasp) { if (u == 0) li = i - 1; else if (u == 1) lj = j - 1; else lk = k - 1; }
    // the active node, its size along u, its face on this side and its boundary code
    const long cs = hasp ? cp : cn;
    const double hs = hasp ? hdir(q, u, i, j, k) : hdir(q, u, li, lj, lk);
    const int face = hasp ? sf + 1 : sf;
    const int bc = hasp ? q.bc[2 * u + 1] : q.bc[2 * u];
    double B[G];
    nod_B_row_pre<G>(a, NC, cs, hs, g, B);
    const double *cq = a.cf4 + (((size_t)u * NC + cs) * G + g) * 4;
    const double A = cq[0], F = cq[1], Gc = cq[2], H = cq[3];
    const double L1v = a.L1[ub + (size_t)g * NC + cs];
    const double a2v = a.a2[ub + (size_t)g * NC + cs], a4v = a.a4[ub + (size_t)g * NC + cs];
    const double P = 2.0 * a.xs.D[(size_t)g * NC + cs] / hs;
    const double dc = a.xs.dc[((size_t)face * G + g) * NC + cs];
    const double f0v = a.f0[(size_t)g * NC + cs];
    double R[G], bv;
    if (hasp) {   // get_a1matvec_first
#pragma unroll
      for (int hh = 0; hh < G; ++hh) {
        const double Bv = B[hh];
        if (bc == 2) R[hh] = (hh == g) ? P * (Bv * F + 1.0) : P * Bv * F;
        else if (bc == 1)
          R[hh] = (hh == g) ? -dc * (1.0 + A * Bv) - 2.0 * P * (A * Bv * H + 1.0) : -dc * A * Bv - 2.0 * P * A * Bv * H;
        else R[hh] = (hh == g) ? dc * (1.0 + A * Bv) : dc * A * Bv;
      }
      if (bc == 2) bv = P * (3.0 * a2v + Gc * a4v - F * L1v);
      else if (bc == 1) bv = 2.0 * P * (A * H * L1v - 3.0 * a2v - Gc * a4v) - dc * (a2v + a4v + f0v - A * L1v);
      else bv = dc * (a2v + a4v + f0v - A * L1v);
    } else {      // get_a1matvec_last
#pragma unroll
      for (int hh = 0; hh < G; ++hh) {
        const double Bv = B[hh];
        if (bc == 2) R[hh] = (hh == g) ? -P * (Bv * F + 1.0) : -P * Bv * F;
        else if (bc == 1)
          R[hh] = (hh == g) ? dc * (1.0 + A * Bv) + 2.0 * P * (A * Bv * H + 1.0) : dc * A * Bv + 2.0 * P * A * Bv * H;
        else R[hh] = (hh == g) ? dc * (1.0 + A * Bv) : dc * A * Bv;
      }
      if (bc == 2) bv = P * (3.0 * a2v + Gc * a4v + F * L1v);
      else if (bc == 1) bv = -2.0 * P * (A * H * L1v + 3.0 * a2v + Gc * a4v) - dc * (a2v + a4v + f0v + A * L1v);
      else bv = -dc * (a2v + a4v + f0v + A * L1v);
    }
    if (coop_lu_diag_small<G>(R, r)) atomicMin(&sc->lu_key, lu_key(ord, li, lj, lk));
    double xs[G];
    coop_lu<G, LaneRow8>(R, bv, r, f, xs);
    if (r < G) {
      // get_a3 and the boundary forms of nodal_coup_upd for group g
      double Bf = 0.0;
#pragma unroll
      for (int hh = 0; hh < G; ++hh) Bf = Bf + B[hh] * xs[hh];
      const double a3 = A * (Bf + L1v);
      const size_t o = ((size_t)face * G + g) * NC + cs;
      const double ndpr = a.dn[o];
      double dv;
      if (hasp) {
        const double jp = -2.0 * a.xs.D[(size_t)g * NC + cs] / hs * (xs[g] - 3.0 * a2v + H * a3 - Gc * a4v);
        dv = -(jp / f0v + a.df[o]);
      } else {
        const double jp = -2.0 * a.xs.D[(size_t)g * NC + cs] / hs * (xs[g] + 3.0 * a2v + H * a3 + Gc * a4v);
        dv = -(jp / f0v - a.df[o]);
      }
      a.dn[o] = dv;
      const double nder = fabs(dv - ndpr), key = (double)(ord * G + g);
      if (nder > best || (nder == best && key < bkey)) {
        best = nder;
        bkey = key;
        bloc = (double)(((li + 1) << 20) | ((lj + 1) << 10) | (lk + 1));
      }
    }
  }
  face_argmax(best, bkey, bloc, a.part + ((size_t)3 * nfb + blockIdx.x) * 3);
}

// ---------------------------------------------------------------------------
// pass 2, interior faces, cooperative form for larger group counts: 16 lanes
// per face (four faces per wavefront), lane 16 f + r owning row r of face f's
// 2G x 2G system of get_a1matvec (:630-698), solved by coop_lu in the
// LaneRow16 layout (L32: by
// coop_lu_refine, FP32 factors and one FP64 defect correction; the system and
// everything after the solve stay FP64).  A block covers the same 256-face
// range as k_nod_pass2 (16 faces at a time); the boundary faces are
// k_nod_pass2b's.
// ---------------------------------------------------------------------------
template <int G, bool L32 = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) k_nod_pass2c(NodArgs a, int nbx) {
#pragma clang fp contract(off)
  constexpr int M2 = 2 * G;
  static_assert(M2 <= 16, "16 lanes per face");
  const DirBlk db = dir_blk();
  if (db.bx >= nbx) return;                  // the grid's padding blocks
  const BlkView v = blk_view_at(a.lv, db.bx);
  const Geo &q = *v.q;
  Scal *sc = v.sc;
  if (a.sched && !nodal_outer(sc)) return;   // uniform over the core's blocks
  const int u = db.u;
  const long NC = q.NC;
  const int nu = u == 0 ? q.nxx : (u == 1 ? q.nyy : q.nzz);
  const int e1 = u == 0 ? q.nyy : q.nxx;
  const int e2 = u == 2 ? q.nyy : q.nzz;
  const long nf = (long)(nu + 1) * e1 * e2;
  // lane = 16 f + r: face f of the wavefront's four neighbouring faces is
  // one 16-lane row (its pivot rows broadcast by DPP), row r of its system
  const int lane = threadIdx.x & 63, r = lane & 15, fw = lane >> 4;
  const int slot16 = (int)(threadIdx.x >> 6) * 4 + fw;
  const int g = r % G;                         // group of this lane's row (rows r and r + G)
  double best = -1.0, bkey = 0.0, bloc = 0.0;
  for (int it = 0; it < NT / 16; ++it) {
    const long t = v.b.lo + (long)it * 16 + slot16;
    if (!(t < v.b.hi && t < nf) || sc->done) continue;   // uniform over the 16 lanes of the face
    // faces enumerated with i fastest in every direction (the argmax key
    // below does not depend on the enumeration)
    int fi, o1, o2;
    if (u == 0) {
      fi = (int)(t % (nu + 1));
      const long rr = t / (nu + 1);
      o1 = (int)(rr % e1); o2 = (int)(rr / e1);
    } else {
      o1 = (int)(t % e1);
      const long rr = t / e1;
      fi = (int)(rr % (nu + 1)); o2 = (int)(rr / (nu + 1));
    }
    int i, j, k;
    if (u == 0) { i = fi; j = o1; k = o2; }
    else if (u == 1) { i = o1; j = fi; k = o2; }
    else { i = o1; j = o2; k = fi; }
    const long st = u == 0 ? 1 : (u == 1 ? q.sy : q.sz);
    const long cp = cell_of(q, i, j, k);
    const long cn = cp - st;
    if (!((q.flag[cp] & F_ACTIVE) && (q.flag[cn] & F_ACTIVE))) continue;   // interior faces only
    const int sf = 2 * u;
    const unsigned long long ord = face_order(q, u, fi, o1, o2);
    const size_t ub = (size_t)u * G * NC;
    int li = i, lj = j, lk = k;
    if (u == 0) li = i - 1; else if (u == 1) lj = j - 1; else lk = k - 1;
    const double hn = hdir(q, u, u == 0 ? i - 1 : i, u == 1 ? j - 1 : j, u == 2 ? k - 1 : k);
    const double hp = hdir(q, u, i, j, k);
    // this lane's group g: rows 0..G-1 make node n's row of B, coefficients
    // and moments, rows G..2G-1 node p's; the partner row (r + G or r - G)
    // swaps its side in, so every lane holds both sides, computed as the
    // per-thread kernel computes them
    const bool nside = r < G;
    const long cs = nside ? cn : cp;
    const double hsd = nside ? hn : hp;
    double Bs[G];
    nod_B_row_pre<G>(a, NC, cs, hsd, g, Bs);
    Coef6 cfs;   // A, F, G, H of this side's node as pass 1 made them
    {
      const double *cq = a.cf4 + (((size_t)u * NC + cs) * G + g) * 4;
      cfs.A = cq[0];
      cfs.B = 0.0;
      cfs.E = 0.0;
      cfs.F = cq[1];
      cfs.G = cq[2];
      cfs.H = cq[3];
    }
    const double Ls = a.L1[ub + (size_t)g * NC + cs];
    const double a2s = a.a2[ub + (size_t)g * NC + cs], a4s = a.a4[ub + (size_t)g * NC + cs];
    const double Ps = 2.0 * a.xs.D[(size_t)g * NC + cs] / hsd;
    const double dcs = a.xs.dc[((size_t)(nside ? sf : sf + 1) * G + g) * NC + cs];
    const double fs = a.f0[(size_t)g * NC + cs];
    // the partner row's value: for G = 8 the partner is lane r ^ 8 of the
    // face's row, a DPP row_ror:8 move; otherwise a lane shuffle
    const int pl = 16 * fw + (nside ? r + G : r - G);
    auto partner = [&](double x) -> double {
      if constexpr (G == 8)
        return __longlong_as_double(__builtin_amdgcn_mov_dpp(__double_as_longlong(x), 0x128, 0xf, 0xf, false));
      else
        return __shfl(x, pl, 64);
    };
    double Bo[G];
#pragma unroll
    for (int hh = 0; hh < G; ++hh) Bo[hh] = partner(Bs[hh]);
    Coef6 cfo;
    cfo.A = partner(cfs.A);
    cfo.B = 0.0;
    cfo.E = 0.0;
    cfo.F = partner(cfs.F);
    cfo.G = partner(cfs.G);
    cfo.H = partner(cfs.H);
    const double Lo = partner(Ls), a2o = partner(a2s), a4o = partner(a4s);
    const double Po = partner(Ps), dco = partner(dcs), fo = partner(fs);
    double Bn[G], Bp[G];
#pragma unroll
    for (int hh = 0; hh < G; ++hh) {
      Bn[hh] = nside ? Bs[hh] : Bo[hh];
      Bp[hh] = nside ? Bo[hh] : Bs[hh];
    }
    const Coef6 cfn = nside ? cfs : cfo, cfp = nside ? cfo : cfs;
    const double Ln1 = nside ? Ls : Lo, Lp1 = nside ? Lo : Ls;
    const double a2n = nside ? a2s : a2o, a4n = nside ? a4s : a4o;
    const double a2p = nside ? a2o : a2s, a4p = nside ? a4o : a4s;
    const double fn0 = nside ? fs : fo, fp0 = nside ? fo : fs;
    // row r of the system
    double R[M2], s;
    if (r < G) {
      const double Pn = Ps, Pp = Po;   // 2 D / h of n and p
#pragma unroll
      for (int hh = 0; hh < G; ++hh) {
        if (hh == g) {
          R[hh] = -Pn * (Bn[hh] * cfn.F + 1.0);
          R[hh + G] = Pp * (Bp[hh] * cfp.F + 1.0);
        } else {
          R[hh] = -Pn * Bn[hh] * cfn.F;
          R[hh + G] = Pp * Bp[hh] * cfp.F;
        }
      }
      s = Pn * (3.0 * a2n + cfn.G * a4n + cfn.F * Ln1) + Pp * (3.0 * a2p + cfp.G * a4p - cfp.F * Lp1);
    } else {
      const double dcn = dco, dcp = dcs;   // dc of n's upper face and of p's lower face
#pragma unroll
      for (int hh = 0; hh < G; ++hh) {
        if (hh == g) {
          R[hh] = dcn * (Bn[hh] * cfn.A + 1.0);
          R[hh + G] = dcp * (Bp[hh] * cfp.A + 1.0);
        } else {
          R[hh] = dcn * Bn[hh] * cfn.A;
          R[hh + G] = dcp * Bp[hh] * cfp.A;
        }
      }
      // the reference pairs An with Ln1 and Ap with Lp1 here (:693-694)
      s = dcp * (a2p + a4p + fp0 - cfn.A * Ln1) - dcn * (a2n + a4n + fn0 + cfp.A * Lp1);
    }
    // LU_solve (:829-897): the input-diagonal test, then the elimination
    if (coop_lu_diag_small<M2>(R, r)) atomicMin(&sc->lu_key, lu_key(ord, li, lj, lk));
    double xs[M2];
    if constexpr (L32) coop_lu_refine<M2, LaneRow16>(R, s, r, fw, xs);
    else coop_lu<M2, LaneRow16>(R, s, r, fw, xs);
    // get_a3 and nodal_coup_upd for group g (lanes 0..G-1); node n's B row
    // read again rather than kept live through the elimination
    if (r < G) {
      double Bn2[G];
      nod_B_row_pre<G>(a, NC, cn, hn, g, Bn2);
      double Bf = 0.0;
#pragma unroll
      for (int hh = 0; hh < G; ++hh) Bf = Bf + Bn2[hh] * xs[hh];
      const double a3 = cfn.A * (Bf + Ln1);
      const double jp = -2.0 * a.xs.D[(size_t)g * NC + cn] / hn * (xs[g] + 3.0 * a2n + cfn.H * a3 + cfn.G * a4n);
      const size_t on = ((size_t)sf * G + g) * NC + cn, op = ((size_t)(sf + 1) * G + g) * NC + cp;
      const double ndpr = a.dn[on];
      const double fn = fn0, fp = fp0;
      const double dv = (a.df[on] * (fn - fp) - jp) / (fn + fp);
      a.dn[on] = dv;
      a.dn[op] = dv;
      const double nder = fabs(dv - ndpr), key = (double)(ord * G + g);
      if (nder > best || (nder == best && key < bkey)) {
        best = nder;
        bkey = key;
        bloc = (double)(((li + 1) << 20) | ((lj + 1) << 10) | (lk + 1));
      }
    }
  }
  // block argmax (largest value, then smallest sweep key), merged with the
  // boundary faces' partial of this block
  __shared__ double sv[NT], sk[NT], sl[NT];
  sv[threadIdx.x] = best;
  sk[threadIdx.x] = bkey;
  sl[threadIdx.x] = bloc;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int o = threadIdx.x + w;
      if (sv[o] > sv[threadIdx.x] || (sv[o] == sv[threadIdx.x] && sk[o] < sk[threadIdx.x])) {
        sv[threadIdx.x] = sv[o]; sk[threadIdx.x] = sk[o]; sl[threadIdx.x] = sl[o];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const size_t slot = ((size_t)u * nbx + db.bx) * 3;   // k_nod_pass2's slot of this (direction, block)
    a.part[slot] = sv[0];
    a.part[slot + 1] = sk[0];
    a.part[slot + 2] = sl[0];
  }
}

// final argmax over the pass-2 partials, one block per core (blockIdx.x =
// core): its face blocks [ffirst, ffirst + fnb) of each direction, nfb face
// blocks per direction in the launch, then its boundary-face blocks
// [bfirst, bfirst + bnb) of k_nod_pass2b
__global__ void __launch_bounds__(NT) k_nod_argmax(LV lv, const double *part, int nfb, int sched) {
  const Geo &q = lv.geo[blockIdx.x];
  Scal *sc = lv.sc + blockIdx.x;
  if (sc->done || (sched && !nodal_outer(sc))) return;
  __shared__ double sv[NT], sk[NT], sl[NT];
  double best = -1.0, bkey = 0.0, bloc = 0.0;
  const int nslots = 3 * q.fnb;
  for (int s = threadIdx.x; s < nslots + q.bnb; s += NT) {
    size_t o;
    if (s < nslots) {
      const int u = s / q.fnb, fb = s - u * q.fnb;
      o = ((size_t)u * nfb + q.ffirst + fb) * 3;
    } else {
      o = ((size_t)3 * nfb + q.bfirst + (s - nslots)) * 3;   // the boundary-face blocks
    }
    const double v = part[o], kk = part[o + 1];
    if (v > best || (v == best && kk < bkey)) { best = v; bkey = kk; bloc = part[o + 2]; }
  }
  sv[threadIdx.x] = best; sk[threadIdx.x] = bkey; sl[threadIdx.x] = bloc;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const int o = threadIdx.x + w;
      if (sv[o] > sv[threadIdx.x] || (sv[o] == sv[threadIdx.x] && sk[o] < sk[threadIdx.x])) {
        sv[threadIdx.x] = sv[o]; sk[threadIdx.x] = sk[o]; sl[threadIdx.x] = sl[o];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double v = sv[0] < 0.0 ? 0.0 : sv[0];
    sc->ndmax = v;
    const int loc = (int)sl[0];
    if (sv[0] > 0.0) {
      sc->im = (loc >> 20) & 1023;
      sc->jm = (loc >> 10) & 1023;
      sc->km = loc & 1023;
    }
  }
}

}  // namespace adg
