/*
 * CPU restatement of ADPRES's static CMFD + SANM/PNM hot path (FP64).
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.
 *
 * Each function cites the reference routine it restates.  Loop order and
 * the order of floating-point operations follow the reference so that the
 * results agree bit for bit with the amdflang build of the reference when
 * this file is compiled with -ffp-contract=off.
 */
#include "adpres_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX2(n, g, N) ((size_t)(n) + (size_t)(N) * (size_t)(g))
#define IDX3(n, g, f, N, G) ((size_t)(n) + (size_t)(N) * ((size_t)(g) + (size_t)(G) * (size_t)(f)))

/* ------------------------------------------------------------------ */
/* geometry helpers                                                    */
/* ------------------------------------------------------------------ */

typedef struct {
  const orc_problem *pb;
  int N, G;
  int *xyz;                 /* dense (nxx,nyy,nzz) -> 0-based node, -1 outside */
  int *col;                 /* (N,7) neighbour per matrix slot, -1 absent */
} geo_t;

static int xyz_at(const geo_t *q, int i, int j, int k) { /* 1-based i,j,k */
  const orc_problem *pb = q->pb;
  return q->xyz[(size_t)(i - 1) + (size_t)pb->nxx * ((size_t)(j - 1) + (size_t)pb->nyy * (size_t)(k - 1))];
}

static int geo_init(geo_t *q, const orc_problem *pb) {
  q->pb = pb;
  q->N = pb->nnod;
  q->G = pb->ng;
  size_t box = (size_t)pb->nxx * pb->nyy * pb->nzz;
  q->xyz = (int *)malloc(box * sizeof(int));
  q->col = (int *)malloc((size_t)q->N * 7 * sizeof(int));
  if (!q->xyz || !q->col) return ORC_E_ARG;
  for (size_t c = 0; c < box; ++c) q->xyz[c] = -1;
  for (int n = 0; n < q->N; ++n) {
    int i = pb->ix[n], j = pb->iy[n], k = pb->iz[n];
    q->xyz[(size_t)(i - 1) + (size_t)pb->nxx * ((size_t)(j - 1) + (size_t)pb->nyy * (size_t)(k - 1))] = n;
  }
  /* set_ind (src/mod_cmfd.f90:141-213): slots z-, y-, x-, diag, x+, y+, z+ */
  for (int n = 0; n < q->N; ++n) {
    int i = pb->ix[n], j = pb->iy[n], k = pb->iz[n];
    int *c = q->col + 7 * (size_t)n;
    c[0] = (k != 1) ? xyz_at(q, i, j, k - 1) : -1;
    c[1] = (j != pb->xstag_min[i - 1]) ? xyz_at(q, i, j - 1, k) : -1;
    c[2] = (i != pb->ystag_min[j - 1]) ? xyz_at(q, i - 1, j, k) : -1;
    c[3] = n;
    c[4] = (i != pb->ystag_max[j - 1]) ? xyz_at(q, i + 1, j, k) : -1;
    c[5] = (j != pb->xstag_max[i - 1]) ? xyz_at(q, i, j + 1, k) : -1;
    c[6] = (k != pb->nzz) ? xyz_at(q, i, j, k + 1) : -1;
  }
  return ORC_OK;
}

static void geo_free(geo_t *q) {
  free(q->xyz);
  free(q->col);
}

/* ------------------------------------------------------------------ */
/* coup_coef (src/mod_cmfd.f90:11-137)                                 */
/* ------------------------------------------------------------------ */

static double df_boundary(int bc, double D, double h) {
  const double alb = 1.e30;
  if (bc == 0) return 2.0 * alb * D / (2.0 * D + alb * h);
  if (bc == 1) return D / (2.0 * D + 0.5 * h);
  return 0.0;
}

static double df_interior(double D, double dnb, double h_nb, double h) {
  return 2.0 * D * dnb / (D * h_nb + dnb * h);
}

static void coup_coef_q(const geo_t *q, double *df) {
  const orc_problem *pb = q->pb;
  int N = q->N, G = q->G;
  for (int g = 0; g < G; ++g) {
    for (int n = 0; n < N; ++n) {
      int i = pb->ix[n], j = pb->iy[n], k = pb->iz[n];
      double D = pb->D[IDX2(n, g, N)];
      const double *xd = pb->xdel, *yd = pb->ydel, *zd = pb->zdel;
      double v;
      /* x+ (east) */
      if (i == pb->ystag_max[j - 1]) v = df_boundary(pb->bc[0], D, xd[i - 1]);
      else v = df_interior(D, pb->D[IDX2(xyz_at(q, i + 1, j, k), g, N)], xd[i], xd[i - 1]);
      df[IDX3(n, g, 0, N, G)] = v;
      /* x- (west) */
      if (i == pb->ystag_min[j - 1]) v = df_boundary(pb->bc[1], D, xd[i - 1]);
      else v = df_interior(D, pb->D[IDX2(xyz_at(q, i - 1, j, k), g, N)], xd[i - 2], xd[i - 1]);
      df[IDX3(n, g, 1, N, G)] = v;
      /* y+ (north) */
      if (j == pb->xstag_max[i - 1]) v = df_boundary(pb->bc[2], D, yd[j - 1]);
      else v = df_interior(D, pb->D[IDX2(xyz_at(q, i, j + 1, k), g, N)], yd[j], yd[j - 1]);
      df[IDX3(n, g, 2, N, G)] = v;
      /* y- (south) */
      if (j == pb->xstag_min[i - 1]) v = df_boundary(pb->bc[3], D, yd[j - 1]);
      else v = df_interior(D, pb->D[IDX2(xyz_at(q, i, j - 1, k), g, N)], yd[j - 2], yd[j - 1]);
      df[IDX3(n, g, 3, N, G)] = v;
      /* z+ (top) */
      if (k == pb->nzz) v = df_boundary(pb->bc[5], D, zd[k - 1]);
      else v = df_interior(D, pb->D[IDX2(xyz_at(q, i, j, k + 1), g, N)], zd[k], zd[k - 1]);
      df[IDX3(n, g, 4, N, G)] = v;
      /* z- (bottom) */
      if (k == 1) v = df_boundary(pb->bc[4], D, zd[k - 1]);
      else v = df_interior(D, pb->D[IDX2(xyz_at(q, i, j, k - 1), g, N)], zd[k - 2], zd[k - 1]);
      df[IDX3(n, g, 5, N, G)] = v;
    }
  }
}

void orc_coup_coef(const orc_problem *pb, orc_state *st) {
  geo_t q;
  if (geo_init(&q, pb) != ORC_OK) { geo_free(&q); return; }
  coup_coef_q(&q, st->df);
  geo_free(&q);
}

/* ------------------------------------------------------------------ */
/* matrix_setup (src/mod_cmfd.f90:217-304): A(N,G,7) with col slots    */
/* ------------------------------------------------------------------ */

static void matrix_setup_q(const geo_t *q, const double *df, const double *dn, double *A) {
  const orc_problem *pb = q->pb;
  int N = q->N, G = q->G;
  for (int g = 0; g < G; ++g) {
    for (int n = 0; n < N; ++n) {
      int i = pb->ix[n], j = pb->iy[n], k = pb->iz[n];
      double hx = pb->xdel[i - 1], hy = pb->ydel[j - 1], hz = pb->zdel[k - 1];
      double f[6], d[6];
      for (int s = 0; s < 6; ++s) {
        f[s] = df[IDX3(n, g, s, N, G)];
        d[s] = dn[IDX3(n, g, s, N, G)];
      }
      double *a = A + 7 * IDX2(n, g, N);
      a[0] = -(f[5] - d[5]) / hz;
      a[1] = -(f[3] - d[3]) / hy;
      a[2] = -(f[1] - d[1]) / hx;
      a[3] = (f[0] + f[1] - d[0] + d[1]) / hx + (f[2] + f[3] - d[2] + d[3]) / hy +
             (f[4] + f[5] - d[4] + d[5]) / hz + pb->sigr[IDX2(n, g, N)];
      a[4] = -(f[0] + d[0]) / hx;
      a[5] = -(f[2] + d[2]) / hy;
      a[6] = -(f[4] + d[4]) / hz;
    }
  }
}

/* sp_matvec (src/mod_cmfd.f90:1247-1268) */
static void matvec_q(const geo_t *q, const double *A, int g, const double *x, double *v) {
  int N = q->N;
  for (int n = 0; n < N; ++n) {
    const int *c = q->col + 7 * (size_t)n;
    const double *a = A + 7 * IDX2(n, g, N);
    double s = 0.0;
    for (int t = 0; t < 7; ++t)
      if (c[t] >= 0) s = s + a[t] * x[c[t]];
    v[n] = s;
  }
}

/* dproduct (src/mod_cmfd.f90:1272-1286) */
static double dprod(const double *a, const double *b, int N) {
  double x = 0.0;
  for (int n = 0; n < N; ++n) x = x + a[n] * b[n];
  return x;
}

/* bicg (src/mod_cmfd.f90:1203-1243): exactly imax iterations, warm start */
static int bicg_q(const geo_t *q, const double *A, int imax, int g, const double *b, double *x) {
  int N = q->N;
  double *w = (double *)malloc((size_t)N * 7 * sizeof(double));
  if (!w) return ORC_E_ARG;
  double *r = w, *rs = w + N, *v = w + 2 * (size_t)N, *p = w + 3 * (size_t)N;
  double *s = w + 4 * (size_t)N, *t = w + 5 * (size_t)N, *ax = w + 6 * (size_t)N;
  matvec_q(q, A, g, x, ax);
  for (int n = 0; n < N; ++n) r[n] = b[n] - ax[n];
  memcpy(rs, r, (size_t)N * sizeof(double));
  double rho = 1.0, alpha = 1.0, omega = 1.0, rho_prev, beta, theta;
  for (int n = 0; n < N; ++n) { v[n] = 0.0; p[n] = 0.0; }
  for (int i = 0; i < imax; ++i) {
    rho_prev = rho;
    rho = dprod(rs, r, N);
    beta = (rho / rho_prev) * (alpha / omega);
    for (int n = 0; n < N; ++n) p[n] = r[n] + beta * (p[n] - omega * v[n]);
    matvec_q(q, A, g, p, v);
    alpha = rho / dprod(rs, v, N);
    for (int n = 0; n < N; ++n) s[n] = r[n] - alpha * v[n];
    matvec_q(q, A, g, s, t);
    theta = dprod(t, t, N);
    omega = dprod(t, s, N) / theta;
    for (int n = 0; n < N; ++n) x[n] = x[n] + alpha * p[n] + omega * s[n];
    for (int n = 0; n < N; ++n) r[n] = s[n] - omega * t[n];
  }
  free(w);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* nodal update (src/mod_nodal.f90)                                    */
/* ------------------------------------------------------------------ */

typedef struct {
  const geo_t *q;
  orc_state *st;
  int cmode;               /* 1 forward, 0 adjoint, 2 transient */
  int G;
  double *S[3];            /* S1, S2, S3 (N,G) */
  double *Bcn, *Bcp;       /* (G,G) row-major [g*G+h] */
  double *An, *Bn, *En, *Fn, *Gn, *Hn;
  double *Ap, *Bp, *Ep, *Fp, *Gp, *Hp;
  double *Lm2, *Ln1, *Lp1;
  double *a1n, *a2n, *a3n, *a4n, *a1p, *a2p, *a3p, *a4p;
  double *M, *rhs, *sol;   /* LU scratch, up to 2G */
  double *L, *U, *y;
  int err;                 /* first error status */
  int err_loc[3];
} nod_t;

static double hsize(const orc_problem *pb, int u, int n) {
  if (u == 1) return pb->xdel[pb->ix[n] - 1];
  if (u == 2) return pb->ydel[pb->iy[n] - 1];
  return pb->zdel[pb->iz[n] - 1];
}

/* Lxyz (src/mod_nodal.f90:901-1005) */
static void lxyz(const nod_t *w, int n, int g, double *L1, double *L2, double *L3) {
  const geo_t *q = w->q;
  const orc_problem *pb = q->pb;
  int N = q->N, G = q->G;
  const double *f0 = w->st->f0, *df = w->st->df, *dn = w->st->dn;
  int i = pb->ix[n], j = pb->iy[n], k = pb->iz[n];
  double fn = f0[IDX2(n, g, N)];
#define DF(s) df[IDX3(n, g, s, N, G)]
#define DN(s) dn[IDX3(n, g, s, N, G)]
  double jp, jm;
  /* x */
  if (i == pb->ystag_max[j - 1]) {
    jp = (pb->bc[0] == 2) ? 0.0 : DF(0) * fn - DN(0) * fn;
  } else {
    double fp = f0[IDX2(xyz_at(q, i + 1, j, k), g, N)];
    jp = -DF(0) * (fp - fn) - DN(0) * (fp + fn);
  }
  if (i == pb->ystag_min[j - 1]) {
    jm = (pb->bc[1] == 2) ? 0.0 : -DF(1) * fn - DN(1) * fn;
  } else {
    double fm = f0[IDX2(xyz_at(q, i - 1, j, k), g, N)];
    jm = -DF(1) * (fn - fm) - DN(1) * (fn + fm);
  }
  *L1 = (jp - jm) / pb->xdel[i - 1];
  /* y */
  if (j == pb->xstag_max[i - 1]) {
    jp = (pb->bc[2] == 2) ? 0.0 : DF(2) * fn - DN(2) * fn;
  } else {
    double fp = f0[IDX2(xyz_at(q, i, j + 1, k), g, N)];
    jp = -DF(2) * (fp - fn) - DN(2) * (fp + fn);
  }
  if (j == pb->xstag_min[i - 1]) {
    jm = (pb->bc[3] == 2) ? 0.0 : -DF(3) * fn - DN(3) * fn;
  } else {
    double fm = f0[IDX2(xyz_at(q, i, j - 1, k), g, N)];
    jm = -DF(3) * (fn - fm) - DN(3) * (fn + fm);
  }
  *L2 = (jp - jm) / pb->ydel[j - 1];
  /* z */
  if (k == pb->nzz) {
    jp = (pb->bc[5] == 2) ? 0.0 : DF(4) * fn - DN(4) * fn;
  } else {
    double fp = f0[IDX2(xyz_at(q, i, j, k + 1), g, N)];
    jp = -DF(4) * (fp - fn) - DN(4) * (fp + fn);
  }
  if (k == 1) {
    jm = (pb->bc[4] == 2) ? 0.0 : -DF(5) * fn - DN(5) * fn;
  } else {
    double fm = f0[IDX2(xyz_at(q, i, j, k - 1), g, N)];
    jm = -DF(5) * (fn - fm) - DN(5) * (fn + fm);
  }
  *L3 = (jp - jm) / pb->zdel[k - 1];
#undef DF
#undef DN
}

/* get_source (src/mod_nodal.f90:1009-1043); cmode 2 subtracts exsrc here */
static void get_source(nod_t *w) {
  int N = w->q->N, G = w->q->G;
  const double *ex = w->q->pb->exsrc;
  for (int g = 0; g < G; ++g)
    for (int n = 0; n < N; ++n) {
      double L1, L2, L3;
      lxyz(w, n, g, &L1, &L2, &L3);
      if (w->cmode == 2) {
        w->S[0][IDX2(n, g, N)] = L2 + L3 - ex[IDX2(n, g, N)];
        w->S[1][IDX2(n, g, N)] = L1 + L3 - ex[IDX2(n, g, N)];
        w->S[2][IDX2(n, g, N)] = L1 + L2 - ex[IDX2(n, g, N)];
      } else {
        w->S[0][IDX2(n, g, N)] = L2 + L3;
        w->S[1][IDX2(n, g, N)] = L1 + L3;
        w->S[2][IDX2(n, g, N)] = L1 + L2;
      }
    }
}

/* line neighbours along direction u: m (lower) and p (upper); -1 when at
 * the line end.  Returns the position flag: 1 first node, 2 last node,
 * 0 interior (the reference tests smin before smax). */
static void line_nb(const nod_t *w, int u, int n, int *m, int *p, int *isfirst, int *islast,
                    double *hm, double *hn, double *hp) {
  const geo_t *q = w->q;
  const orc_problem *pb = q->pb;
  int i = pb->ix[n], j = pb->iy[n], k = pb->iz[n];
  *m = *p = -1;
  if (u == 1) {
    *isfirst = (i == pb->ystag_min[j - 1]);
    *islast = (i == pb->ystag_max[j - 1]);
    if (!*islast) { *p = xyz_at(q, i + 1, j, k); *hp = pb->xdel[i]; }
    if (!*isfirst) { *m = xyz_at(q, i - 1, j, k); *hm = pb->xdel[i - 2]; }
    *hn = pb->xdel[i - 1];
  } else if (u == 2) {
    *isfirst = (j == pb->xstag_min[i - 1]);
    *islast = (j == pb->xstag_max[i - 1]);
    if (!*islast) { *p = xyz_at(q, i, j + 1, k); *hp = pb->ydel[j]; }
    if (!*isfirst) { *m = xyz_at(q, i, j - 1, k); *hm = pb->ydel[j - 2]; }
    *hn = pb->ydel[j - 1];
  } else {
    *isfirst = (k == 1);
    *islast = (k == pb->nzz);
    if (!*islast) { *p = xyz_at(q, i, j, k + 1); *hp = pb->zdel[k]; }
    if (!*isfirst) { *m = xyz_at(q, i, j, k - 1); *hm = pb->zdel[k - 2]; }
    *hn = pb->zdel[k - 1];
  }
}

static int bc_lo(const orc_problem *pb, int u) { return u == 1 ? pb->bc[1] : (u == 2 ? pb->bc[3] : pb->bc[4]); }
static int bc_hi(const orc_problem *pb, int u) { return u == 1 ? pb->bc[0] : (u == 2 ? pb->bc[2] : pb->bc[5]); }

/* TLUpd1 (src/mod_nodal.f90:1047-1199) */
static double tlupd1(nod_t *w, int u, int n, int g) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N;
  const double *S = w->S[u - 1];
  int m, p, isf, isl;
  double hm = 0, hn = 0, hp = 0;
  line_nb(w, u, n, &m, &p, &isf, &isl, &hm, &hn, &hp);
  if (isf && isl) { w->err = w->err ? w->err : ORC_E_ARG; return 0.0; }
  double Sn = S[IDX2(n, g, N)];
  double tm, tp, p1m, p2m, p1p, p2p, hpp, L;
  if (isf) {
    if (bc_lo(pb, u) == 2) {
      tm = 1.0; tp = hp / hn;
      p1m = tm + 1.0; p2m = 2.0 * tm + 1.0;
      p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      L = (p1m * p2m * (S[IDX2(p, g, N)] - Sn)) / hpp;
    } else {
      tp = hp / hn;
      p1p = tp + 1.0;
      L = (S[IDX2(p, g, N)] - Sn) / p1p;
    }
  } else if (isl) {
    if (bc_hi(pb, u) == 2) {
      tm = hm / hn; tp = 1.0;
      p1m = tm + 1.0;
      p1p = tp + 1.0; p2p = 2.0 * tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      L = (p1p * p2p * (Sn - S[IDX2(m, g, N)])) / hpp;
    } else {
      tm = hm / hn;
      p1m = tm + 1.0;
      L = (Sn - S[IDX2(m, g, N)]) / p1m;
    }
  } else {
    tm = hm / hn; tp = hp / hn;
    p1m = tm + 1.0; p2m = 2.0 * tm + 1.0;
    p1p = tp + 1.0; p2p = 2.0 * tp + 1.0;
    hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
    L = (p1m * p2m * (S[IDX2(p, g, N)] - Sn) + p1p * p2p * (Sn - S[IDX2(m, g, N)])) / hpp;
  }
  return 0.25 * (hn * hn) / pb->D[IDX2(n, g, N)] * L;
}

/* TLUpd2 (src/mod_nodal.f90:1203-1341) */
static double tlupd2(nod_t *w, int u, int n, int g) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N;
  const double *S = w->S[u - 1];
  int m, p, isf, isl;
  double hm = 0, hn = 0, hp = 0;
  line_nb(w, u, n, &m, &p, &isf, &isl, &hm, &hn, &hp);
  if (isf && isl) { w->err = w->err ? w->err : ORC_E_ARG; return 0.0; }
  double Sn = S[IDX2(n, g, N)];
  double tm, tp, p1m, p1p, hpp, L;
  if (isf) {
    if (bc_lo(pb, u) == 2) {
      tm = 1.0; tp = hp / hn;
      p1m = tm + 1.0; p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      L = (p1m * (S[IDX2(p, g, N)] - Sn)) / hpp;
    } else {
      L = 0.0;
    }
  } else if (isl) {
    if (bc_hi(pb, u) == 2) {
      tm = hm / hn; tp = 1.0;
      p1m = tm + 1.0; p1p = tp + 1.0;
      hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
      L = (p1p * (S[IDX2(m, g, N)] - Sn)) / hpp;
    } else {
      L = 0.0;
    }
  } else {
    tm = hm / hn; tp = hp / hn;
    p1m = tm + 1.0; p1p = tp + 1.0;
    hpp = 2.0 * p1m * p1p * (tm + tp + 1.0);
    L = (p1m * (S[IDX2(p, g, N)] - Sn) + p1p * (S[IDX2(m, g, N)] - Sn)) / hpp;
  }
  return 0.25 * (hn * hn) / pb->D[IDX2(n, g, N)] * L;
}

/* get_B (src/mod_nodal.f90:1345-1405), cmode 1 forward / 0 adjoint / 2
   transient (chi twice in the diagonal term, as the reference has it) */
static void get_B(const nod_t *w, int u, int n, double *B) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G;
  double h = hsize(pb, u, n);
  double Ke = w->st->Ke;
  int mt = pb->mat[n] - 1;
  for (int g = 0; g < G; ++g)
    for (int hh = 0; hh < G; ++hh) {
      double dum;
      if (w->cmode == 1) {
        if (g == hh) dum = pb->sigr[IDX2(n, g, N)] - pb->chi[IDX2(mt, g, pb->nmat)] * pb->nuf[IDX2(n, hh, N)] / Ke;
        else dum = -pb->sigs[IDX3(n, hh, g, N, G)] - pb->chi[IDX2(mt, g, pb->nmat)] * pb->nuf[IDX2(n, hh, N)] / Ke;
      } else if (w->cmode == 2) {
        double c = pb->chi[IDX2(mt, g, pb->nmat)];
        if (g == hh) dum = pb->sigr[IDX2(n, g, N)] - pb->wtr[n] * c * c * pb->nuf[IDX2(n, hh, N)];
        else dum = -pb->sigs[IDX3(n, hh, g, N, G)] - pb->wtr[n] * c * pb->nuf[IDX2(n, hh, N)];
      } else {
        if (g == hh) dum = pb->sigr[IDX2(n, g, N)] - pb->chi[IDX2(mt, g, pb->nmat)] * pb->nuf[IDX2(n, hh, N)] / Ke;
        else dum = -pb->sigs[IDX3(n, g, hh, N, G)] - pb->chi[IDX2(mt, hh, pb->nmat)] * pb->nuf[IDX2(n, g, N)] / Ke;
      }
      B[g * G + hh] = 0.25 * (h * h) / pb->D[IDX2(n, g, N)] * dum;
    }
}

/* get_ABEFGH (src/mod_nodal.f90:1409-1451) */
static void get_ABEFGH(const nod_t *w, int n, int u, double *Aa, double *Bb, double *Ee,
                       double *Ff, double *Gg, double *Hh) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G;
  double dn = hsize(pb, u, n);
  for (int g = 0; g < G; ++g) {
    double alp = 0.5 * sqrt(pb->sigr[IDX2(n, g, N)] / pb->D[IDX2(n, g, N)]) * dn;
    double alp2 = alp * alp;
    double sh = sinh(alp), ch = cosh(alp);
    double m0c = sh / alp;
    double m1s = 3.0 * (ch / alp - sh / alp2);
    double m2c = 5.0 * (sh / alp - 3.0 * ch / alp2 + 3.0 * sh / (alp * alp * alp));
    Aa[g] = (sh - m1s) / (alp2 * m1s);
    Bb[g] = (ch - m0c - m2c) / (alp2 * m2c);
    Ee[g] = (m0c / m2c - 3.0 / alp2);
    Ff[g] = (alp * ch - m1s) / (alp2 * m1s);
    Gg[g] = (alp * sh - 3.0 * m2c) / (ch - m0c - m2c);
    Hh[g] = (alp * ch - m1s) / (sh - m1s);
  }
}

/* LU_solve (src/mod_nodal.f90:829-897): Doolittle without pivoting.
 * mat row-major (m x m). */
static void lu_solve(nod_t *w, int nt, int m, const double *mat, const double *b, double *x) {
  double *L = w->L, *U = w->U, *y = w->y;
  const float tiny = 10e-5f;   /* the reference compares with a default-real literal */
  memcpy(U, mat, (size_t)m * m * sizeof(double));
  for (int c = 0; c < m * m; ++c) L[c] = 0.0;
  for (int i = 0; i < m; ++i) {
    if (fabs(mat[i * m + i]) < (double)tiny) {
      if (!w->err) {
        const orc_problem *pb = w->q->pb;
        w->err = ORC_E_LU_PIVOT;
        w->err_loc[0] = pb->ix[nt]; w->err_loc[1] = pb->iy[nt]; w->err_loc[2] = pb->iz[nt];
      }
    }
    L[i * m + i] = 1.0;
    for (int j = i + 1; j < m; ++j) {
      double piv = U[j * m + i] / U[i * m + i];
      L[j * m + i] = piv;
      for (int k = i; k < m; ++k) U[j * m + k] = U[j * m + k] - piv * U[i * m + k];
      U[j * m + i] = 0.0;
    }
  }
  y[0] = b[0];
  for (int i = 1; i < m; ++i) {
    double isum = 0.0;
    for (int k = 0; k < i; ++k) isum = isum + L[i * m + k] * y[k];
    y[i] = b[i] - isum;
  }
  x[m - 1] = y[m - 1] / U[(m - 1) * m + (m - 1)];
  for (int i = m - 2; i >= 0; --i) {
    double isum = 0.0;
    for (int k = i + 1; k < m; ++k) isum = isum + U[i * m + k] * x[k];
    x[i] = (y[i] - isum) / U[i * m + i];
  }
}

/* get_a2matvec (src/mod_nodal.f90:769-825) for node n with Bcp/Ep;
 * sets Lm2 (TLUpd2). */
static void get_a2matvec(nod_t *w, int u, int n, double *A, double *b) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G;
  const double *f0 = w->st->f0;
  double h = hsize(pb, u, n);
  double Bf[16];
  for (int g = 0; g < G; ++g) Bf[g] = 0.0;
  for (int g = 0; g < G; ++g) {
    double Su = w->S[u - 1][IDX2(n, g, N)];
    double S = w->cmode == 2 ? 0.25 * (h * h) / pb->D[IDX2(n, g, N)] * Su
                             : 0.25 * (h * h) / pb->D[IDX2(n, g, N)] * (Su - pb->exsrc[IDX2(n, g, N)]);
    for (int hh = 0; hh < G; ++hh) {
      if (hh == g) A[g * G + g] = w->Bcp[g * G + hh] * w->Ep[g] + 3.0;
      else A[g * G + hh] = w->Bcp[g * G + hh] * w->Ep[g];
      Bf[g] = Bf[g] + w->Bcp[g * G + hh] * f0[IDX2(n, hh, N)];
    }
    w->Lm2[g] = tlupd2(w, u, n, g);
    b[g] = Bf[g] - w->Ep[g] * w->Lm2[g] + S;
  }
}

/* get_a4 (src/mod_nodal.f90:740-765) */
static void get_a4(nod_t *w, const double *a2, double *a4) {
  int G = w->q->G;
  for (int g = 0; g < G; ++g) {
    double Bf = 0.0;
    for (int h = 0; h < G; ++h) Bf = Bf + w->Bcp[g * G + h] * a2[h];
    a4[g] = w->Bp[g] * (Bf + w->Lm2[g]);
  }
}

/* get_a3 (src/mod_nodal.f90:702-736) */
static void get_a3(nod_t *w, int cp, const double *a1, const double *Lmn1, double *a3) {
  int G = w->q->G;
  const double *Bc = (cp == 1) ? w->Bcn : w->Bcp;
  const double *Ac = (cp == 1) ? w->An : w->Ap;
  for (int g = 0; g < G; ++g) {
    double Bf = 0.0;
    for (int h = 0; h < G; ++h) Bf = Bf + Bc[g * G + h] * a1[h];
    a3[g] = Ac[g] * (Bf + Lmn1[g]);
  }
}

/* get_a1matvec_first (src/mod_nodal.f90:483-552) */
static void get_a1matvec_first(nod_t *w, int bc, int u, int p, double *A, double *b) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G;
  double dn = hsize(pb, u, p);
  int sf = (u == 1) ? 1 : (u == 2 ? 3 : 5);   /* 0-based face x-, y-, z- */
  const double *f0 = w->st->f0;
  for (int g = 0; g < G; ++g) {
    w->Lp1[g] = tlupd1(w, u, p, g);
    double Pp = 2.0 * pb->D[IDX2(p, g, N)] / dn;
    double dc = pb->dc[IDX3(p, g, sf, N, G)];
    double Ap = w->Ap[g], Fp = w->Fp[g], Gp = w->Gp[g], Hp = w->Hp[g];
    if (bc == 2) {
      for (int h = 0; h < G; ++h) {
        double B = w->Bcp[g * G + h];
        if (h == g) A[g * G + g] = Pp * (B * Fp + 1.0);
        else A[g * G + h] = Pp * B * Fp;
      }
      b[g] = Pp * (3.0 * w->a2p[g] + Gp * w->a4p[g] - Fp * w->Lp1[g]);
    } else if (bc == 1) {
      for (int h = 0; h < G; ++h) {
        double B = w->Bcp[g * G + h];
        if (h == g) A[g * G + g] = -dc * (1.0 + Ap * B) - 2.0 * Pp * (Ap * B * Hp + 1.0);
        else A[g * G + h] = -dc * Ap * B - 2.0 * Pp * Ap * B * Hp;
      }
      b[g] = 2.0 * Pp * (Ap * Hp * w->Lp1[g] - 3.0 * w->a2p[g] - Gp * w->a4p[g]) -
             dc * (w->a2p[g] + w->a4p[g] + f0[IDX2(p, g, N)] - Ap * w->Lp1[g]);
    } else {
      for (int h = 0; h < G; ++h) {
        double B = w->Bcp[g * G + h];
        if (h == g) A[g * G + g] = dc * (1.0 + Ap * B);
        else A[g * G + h] = dc * Ap * B;
      }
      b[g] = dc * (w->a2p[g] + w->a4p[g] + f0[IDX2(p, g, N)] - Ap * w->Lp1[g]);
    }
  }
}

/* get_a1matvec_last (src/mod_nodal.f90:556-626) */
static void get_a1matvec_last(nod_t *w, int bc, int u, int n, double *A, double *b) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G;
  double dn = hsize(pb, u, n);
  int sf = (u == 1) ? 0 : (u == 2 ? 2 : 4);   /* 0-based face x+, y+, z+ */
  const double *f0 = w->st->f0;
  for (int g = 0; g < G; ++g) {
    w->Ln1[g] = w->Lp1[g];
    double Pn = 2.0 * pb->D[IDX2(n, g, N)] / dn;
    double dc = pb->dc[IDX3(n, g, sf, N, G)];
    double An = w->An[g], Fn = w->Fn[g], Gn = w->Gn[g], Hn = w->Hn[g];
    if (bc == 2) {
      for (int h = 0; h < G; ++h) {
        double B = w->Bcn[g * G + h];
        if (h == g) A[g * G + g] = -Pn * (B * Fn + 1.0);
        else A[g * G + h] = -Pn * B * Fn;
      }
      b[g] = Pn * (3.0 * w->a2n[g] + Gn * w->a4n[g] + Fn * w->Ln1[g]);
    } else if (bc == 1) {
      for (int h = 0; h < G; ++h) {
        double B = w->Bcn[g * G + h];
        if (h == g) A[g * G + g] = dc * (1.0 + An * B) + 2.0 * Pn * (An * B * Hn + 1.0);
        else A[g * G + h] = dc * An * B + 2.0 * Pn * An * B * Hn;
      }
      b[g] = -2.0 * Pn * (An * Hn * w->Ln1[g] + 3.0 * w->a2n[g] + Gn * w->a4n[g]) -
             dc * (w->a2n[g] + w->a4n[g] + f0[IDX2(n, g, N)] + An * w->Ln1[g]);
    } else {
      for (int h = 0; h < G; ++h) {
        double B = w->Bcn[g * G + h];
        if (h == g) A[g * G + g] = dc * (1.0 + An * B);
        else A[g * G + h] = dc * An * B;
      }
      b[g] = -dc * (w->a2n[g] + w->a4n[g] + f0[IDX2(n, g, N)] + An * w->Ln1[g]);
    }
  }
}

/* get_a1matvec (src/mod_nodal.f90:630-698): 2G x 2G interface system.
 * The flux-continuity right-hand side pairs An with Ln1 and Ap with Lp1 as
 * the reference does (:693-694). */
static void get_a1matvec(nod_t *w, int u, int n, int p, double *A, double *b) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G, M = 2 * G;
  double hn = hsize(pb, u, n), hp = hsize(pb, u, p);
  int sf = (u == 1) ? 0 : (u == 2 ? 2 : 4);
  const double *f0 = w->st->f0;
  double Pn[16], Pp[16];
  for (int g = 0; g < G; ++g) {
    w->Ln1[g] = w->Lp1[g];
    w->Lp1[g] = tlupd1(w, u, p, g);
    Pn[g] = 2.0 * pb->D[IDX2(n, g, N)] / hn;
    Pp[g] = 2.0 * pb->D[IDX2(p, g, N)] / hp;
    for (int h = 0; h < G; ++h) {
      double Bn_ = w->Bcn[g * G + h], Bp_ = w->Bcp[g * G + h];
      if (h == g) {
        A[g * M + g] = -Pn[g] * (Bn_ * w->Fn[g] + 1.0);
        A[g * M + g + G] = Pp[g] * (Bp_ * w->Fp[g] + 1.0);
      } else {
        A[g * M + h] = -Pn[g] * Bn_ * w->Fn[g];
        A[g * M + h + G] = Pp[g] * Bp_ * w->Fp[g];
      }
    }
    b[g] = Pn[g] * (3.0 * w->a2n[g] + w->Gn[g] * w->a4n[g] + w->Fn[g] * w->Ln1[g]) +
           Pp[g] * (3.0 * w->a2p[g] + w->Gp[g] * w->a4p[g] - w->Fp[g] * w->Lp1[g]);
  }
  for (int g = 0; g < G; ++g) {
    double dcn = pb->dc[IDX3(n, g, sf, N, G)];
    double dcp = pb->dc[IDX3(p, g, sf + 1, N, G)];
    for (int h = 0; h < G; ++h) {
      double Bn_ = w->Bcn[g * G + h], Bp_ = w->Bcp[g * G + h];
      if (h == g) {
        A[(g + G) * M + g] = dcn * (Bn_ * w->An[g] + 1.0);
        A[(g + G) * M + g + G] = dcp * (Bp_ * w->Ap[g] + 1.0);
      } else {
        A[(g + G) * M + h] = dcn * Bn_ * w->An[g];
        A[(g + G) * M + h + G] = dcp * Bp_ * w->Ap[g];
      }
    }
    b[g + G] = dcp * (w->a2p[g] + w->a4p[g] + f0[IDX2(p, g, N)] - w->An[g] * w->Ln1[g]) -
               dcn * (w->a2n[g] + w->a4n[g] + f0[IDX2(n, g, N)] + w->Ap[g] * w->Lp1[g]);
  }
}

/* nodal_coup_upd (src/mod_nodal.f90:282-371); n or p may be -1 (absent) */
static void nodal_coup_upd(nod_t *w, int u, const double *a1, const double *a2, const double *a3,
                           const double *a4, int n, int p) {
  const orc_problem *pb = w->q->pb;
  int N = w->q->N, G = w->q->G;
  orc_state *st = w->st;
  double dh = hsize(pb, u, n >= 0 ? n : p);
  int sf = (u == 1) ? 0 : (u == 2 ? 2 : 4);
  const double *f0 = st->f0;
  for (int g = 0; g < G; ++g) {
    double jp, ndpr, nder;
    if (n >= 0 && p >= 0) {
      jp = -2.0 * pb->D[IDX2(n, g, N)] / dh * (a1[g] + 3.0 * a2[g] + w->Hn[g] * a3[g] + w->Gn[g] * a4[g]);
      ndpr = st->dn[IDX3(n, g, sf, N, G)];
      double fn = f0[IDX2(n, g, N)], fp = f0[IDX2(p, g, N)];
      double v = (st->df[IDX3(n, g, sf, N, G)] * (fn - fp) - jp) / (fn + fp);
      st->dn[IDX3(n, g, sf, N, G)] = v;
      st->dn[IDX3(p, g, sf + 1, N, G)] = v;
      nder = fabs(v - ndpr);
      if (nder > st->ndmax) { st->ndmax = nder; st->im = pb->ix[n]; st->jm = pb->iy[n]; st->km = pb->iz[n]; }
    } else if (p >= 0) {
      jp = -2.0 * pb->D[IDX2(p, g, N)] / dh * (a1[g] - 3.0 * a2[g] + w->Hp[g] * a3[g] - w->Gp[g] * a4[g]);
      ndpr = st->dn[IDX3(p, g, sf + 1, N, G)];
      double v = -(jp / f0[IDX2(p, g, N)] + st->df[IDX3(p, g, sf + 1, N, G)]);
      st->dn[IDX3(p, g, sf + 1, N, G)] = v;
      nder = fabs(v - ndpr);
      if (nder > st->ndmax) { st->ndmax = nder; st->im = pb->ix[p]; st->jm = pb->iy[p]; st->km = pb->iz[p]; }
    } else {
      jp = -2.0 * pb->D[IDX2(n, g, N)] / dh * (a1[g] + 3.0 * a2[g] + w->Hn[g] * a3[g] + w->Gn[g] * a4[g]);
      ndpr = st->dn[IDX3(n, g, sf, N, G)];
      double v = -(jp / f0[IDX2(n, g, N)] - st->df[IDX3(n, g, sf, N, G)]);
      st->dn[IDX3(n, g, sf, N, G)] = v;
      nder = fabs(v - ndpr);
      if (nder > st->ndmax) { st->ndmax = nder; st->im = pb->ix[n]; st->jm = pb->iy[n]; st->km = pb->iz[n]; }
    }
  }
}

static void set_params(nod_t *w, int kern, int u, int node, int right) {
  double *A = right ? w->Ap : w->An, *B = right ? w->Bp : w->Bn, *E = right ? w->Ep : w->En;
  double *F = right ? w->Fp : w->Fn, *Gg = right ? w->Gp : w->Gn, *H = right ? w->Hp : w->Hn;
  if (kern == 2) {
    get_ABEFGH(w, node, u, A, B, E, F, Gg, H);
  } else {
    for (int g = 0; g < w->G; ++g) {
      A[g] = 1.0 / 15.0; B[g] = 1.0 / 35.0; E[g] = 2.0 / 7.0;
      F[g] = 2.0 / 5.0; Gg[g] = 10.0; H[g] = 6.0;
    }
  }
}

static void shift_np(nod_t *w) {
  int G = w->G;
  memcpy(w->Bcn, w->Bcp, (size_t)G * G * sizeof(double));
  memcpy(w->An, w->Ap, G * sizeof(double)); memcpy(w->Bn, w->Bp, G * sizeof(double));
  memcpy(w->En, w->Ep, G * sizeof(double)); memcpy(w->Fn, w->Fp, G * sizeof(double));
  memcpy(w->Gn, w->Gp, G * sizeof(double)); memcpy(w->Hn, w->Hp, G * sizeof(double));
}

/* get_coefs_first (:375-407), get_coefs (:445-479), get_coefs_last (:411-441) */
static void coefs_first(nod_t *w, int bc, int u, int p) {
  int G = w->G;
  get_a2matvec(w, u, p, w->M, w->rhs);
  lu_solve(w, p, G, w->M, w->rhs, w->a2p);
  get_a4(w, w->a2p, w->a4p);
  get_a1matvec_first(w, bc, u, p, w->M, w->rhs);
  lu_solve(w, p, G, w->M, w->rhs, w->a1p);
  get_a3(w, 2, w->a1p, w->Lp1, w->a3p);
}

static void coefs_mid(nod_t *w, int u, int n, int p) {
  int G = w->G;
  memcpy(w->a2n, w->a2p, G * sizeof(double));
  memcpy(w->a4n, w->a4p, G * sizeof(double));
  get_a2matvec(w, u, p, w->M, w->rhs);
  lu_solve(w, p, G, w->M, w->rhs, w->a2p);
  get_a4(w, w->a2p, w->a4p);
  get_a1matvec(w, u, n, p, w->M, w->rhs);
  lu_solve(w, n, 2 * G, w->M, w->rhs, w->sol);
  memcpy(w->a1n, w->sol, G * sizeof(double));
  get_a3(w, 1, w->a1n, w->Ln1, w->a3n);
}

static void coefs_last(nod_t *w, int bc, int u, int n) {
  int G = w->G;
  memcpy(w->a2n, w->a2p, G * sizeof(double));
  memcpy(w->a4n, w->a4p, G * sizeof(double));
  get_a1matvec_last(w, bc, u, n, w->M, w->rhs);
  lu_solve(w, n, G, w->M, w->rhs, w->a1n);
  get_a3(w, 1, w->a1n, w->Ln1, w->a3n);
}

/* one line sweep: nodes[0..len-1] in increasing coordinate along u */
static void sweep_line(nod_t *w, int kern, int u, const int *nodes, int len) {
  const orc_problem *pb = w->q->pb;
  int p = nodes[0];
  get_B(w, u, p, w->Bcp);
  set_params(w, kern, u, p, 1);
  coefs_first(w, bc_lo(pb, u), u, p);
  nodal_coup_upd(w, u, w->a1p, w->a2p, w->a3p, w->a4p, -1, p);
  for (int t = 0; t + 1 < len; ++t) {
    int n = nodes[t];
    p = nodes[t + 1];
    shift_np(w);
    get_B(w, u, p, w->Bcp);
    set_params(w, kern, u, p, 1);
    coefs_mid(w, u, n, p);
    nodal_coup_upd(w, u, w->a1n, w->a2n, w->a3n, w->a4n, n, p);
  }
  int n = nodes[len - 1];
  shift_np(w);
  coefs_last(w, bc_hi(pb, u), u, n);
  nodal_coup_upd(w, u, w->a1n, w->a2n, w->a3n, w->a4n, n, -1);
}

static int nodal_update_q(const geo_t *q, orc_state *st, int kern, int cmode) {
  const orc_problem *pb = q->pb;
  int N = q->N, G = q->G;
  if (G > 16) return ORC_E_ARG;
  nod_t w;
  memset(&w, 0, sizeof(w));
  w.q = q; w.st = st; w.cmode = cmode; w.G = G;
  size_t nv = 3 * (size_t)N * G + 2 * (size_t)G * G + 27 * (size_t)G + 3 * 4 * (size_t)G * G + 3 * 2 * (size_t)G;
  double *mem = (double *)calloc(nv, sizeof(double));
  int maxlen = pb->nxx > pb->nyy ? pb->nxx : pb->nyy;
  if (pb->nzz > maxlen) maxlen = pb->nzz;
  int *line = (int *)malloc((size_t)maxlen * sizeof(int));
  if (!mem || !line) { free(mem); free(line); return ORC_E_ARG; }
  double *c = mem;
  for (int s = 0; s < 3; ++s) { w.S[s] = c; c += (size_t)N * G; }
  w.Bcn = c; c += G * G; w.Bcp = c; c += G * G;
  double **vecs[] = {&w.An, &w.Bn, &w.En, &w.Fn, &w.Gn, &w.Hn, &w.Ap, &w.Bp, &w.Ep, &w.Fp, &w.Gp,
                     &w.Hp, &w.Lm2, &w.Ln1, &w.Lp1, &w.a1n, &w.a2n, &w.a3n, &w.a4n, &w.a1p,
                     &w.a2p, &w.a3p, &w.a4p};
  for (size_t v = 0; v < sizeof(vecs) / sizeof(vecs[0]); ++v) { *vecs[v] = c; c += G; }
  w.M = c; c += 4 * G * G; w.L = c; c += 4 * G * G; w.U = c; c += 4 * G * G;
  w.rhs = c; c += 2 * G; w.sol = c; c += 2 * G; w.y = c; c += 2 * G;

  get_source(&w);
  /* x sweeps: k, j, i */
  for (int k = 1; k <= pb->nzz; ++k)
    for (int j = 1; j <= pb->nyy; ++j) {
      int len = 0;
      for (int i = pb->ystag_min[j - 1]; i <= pb->ystag_max[j - 1]; ++i) line[len++] = xyz_at(q, i, j, k);
      if (len > 0) sweep_line(&w, kern, 1, line, len);
    }
  /* y sweeps: k, i, j */
  for (int k = 1; k <= pb->nzz; ++k)
    for (int i = 1; i <= pb->nxx; ++i) {
      int len = 0;
      for (int j = pb->xstag_min[i - 1]; j <= pb->xstag_max[i - 1]; ++j) line[len++] = xyz_at(q, i, j, k);
      if (len > 0) sweep_line(&w, kern, 2, line, len);
    }
  /* z sweeps: j, i, k */
  for (int j = 1; j <= pb->nyy; ++j)
    for (int i = pb->ystag_min[j - 1]; i <= pb->ystag_max[j - 1]; ++i) {
      int len = 0;
      for (int k = 1; k <= pb->nzz; ++k) line[len++] = xyz_at(q, i, j, k);
      sweep_line(&w, kern, 3, line, len);
    }
  int err = w.err;
  if (err == ORC_E_LU_PIVOT) { st->im = w.err_loc[0]; st->jm = w.err_loc[1]; st->km = w.err_loc[2]; }
  free(mem);
  free(line);
  if (err) return err;
  if (st->ndmax > 1.e3) return ORC_E_NODAL_UNSTABLE;
  return ORC_OK;
}

int orc_nodal_update(const orc_problem *pb, orc_state *st, int kern, int cmode) {
  if (cmode == 2 && !pb->wtr) return ORC_E_ARG;
  geo_t q;
  int rc = geo_init(&q, pb);
  if (rc == ORC_OK) {
    coup_coef_q(&q, st->df);
    st->ndmax = 0.0;
    rc = nodal_update_q(&q, st, kern, cmode);
  }
  geo_free(&q);
  return rc;
}

/* Lxyz (src/mod_nodal.f90:901-1005) of every node and group of the state
   (D-tilde recomputed from D, as coup_coef at the start of an outer):
   L(N,G,3) column-major */
int orc_leakage(const orc_problem *pb, orc_state *st, double *L) {
  geo_t q;
  int rc = geo_init(&q, pb);
  if (rc == ORC_OK) {
    int N = pb->nnod, G = pb->ng;
    nod_t w;
    memset(&w, 0, sizeof w);
    w.q = &q; w.st = st; w.G = G;
    coup_coef_q(&q, st->df);
    for (int g = 0; g < G; ++g)
      for (int n = 0; n < N; ++n)
        lxyz(&w, n, g, &L[IDX3(n, g, 0, N, G)], &L[IDX3(n, g, 1, N, G)], &L[IDX3(n, g, 2, N, G)]);
  }
  geo_free(&q);
  return rc;
}

int orc_bicg(const orc_problem *pb, const orc_state *st, int g, int nin, const double *b, double *x) {
  geo_t q;
  int rc = geo_init(&q, pb);
  double *A = NULL;
  if (rc == ORC_OK) {
    A = (double *)malloc((size_t)pb->nnod * pb->ng * 7 * sizeof(double));
    double *df = (double *)malloc((size_t)pb->nnod * pb->ng * 6 * sizeof(double));
    if (!A || !df) rc = ORC_E_ARG;
    else {
      coup_coef_q(&q, df);
      matrix_setup_q(&q, df, st->dn, A);
      rc = bicg_q(&q, A, nin, g, b, x);
    }
    free(df);
  }
  free(A);
  geo_free(&q);
  return rc;
}

int orc_matvec(const orc_problem *pb, const orc_state *st, int g, const double *x, double *y) {
  geo_t q;
  int rc = geo_init(&q, pb);
  double *A = NULL, *df = NULL;
  if (rc == ORC_OK) {
    A = (double *)malloc((size_t)pb->nnod * pb->ng * 7 * sizeof(double));
    df = (double *)malloc((size_t)pb->nnod * pb->ng * 6 * sizeof(double));
    if (!A || !df) rc = ORC_E_ARG;
    else {
      coup_coef_q(&q, df);
      matrix_setup_q(&q, df, st->dn, A);
      matvec_q(&q, A, g, x, y);
    }
  }
  free(A); free(df);
  geo_free(&q);
  return rc;
}

/* ------------------------------------------------------------------ */
/* sources and reductions                                              */
/* ------------------------------------------------------------------ */

/* FSrc (src/mod_cmfd.f90:956-977) / FSrcAd (:981-1002) */
static void fsrc(const orc_problem *pb, const double *f0, double *fs, int adjoint) {
  int N = pb->nnod, G = pb->ng;
  for (int n = 0; n < N; ++n) fs[n] = 0.0;
  for (int g = 0; g < G; ++g)
    for (int n = 0; n < N; ++n) {
      double w = adjoint ? pb->chi[IDX2(pb->mat[n] - 1, g, pb->nmat)] : pb->nuf[IDX2(n, g, N)];
      fs[n] = fs[n] + f0[IDX2(n, g, N)] * w;
    }
}

/* TSrc (src/mod_cmfd.f90:1006-1033) / TSrcAd (:1037-1064) */
/* TSrc / TSrcAd / TSrcTr (src/mod_cmfd.f90:1006-1096): transient = TSrcTr */
static void tsrc(const orc_problem *pb, int g, double Keff, const double *f0, const double *fs0,
                 double *s0g, double *bs, int adjoint, int transient) {
  int N = pb->nnod, G = pb->ng;
  for (int n = 0; n < N; ++n) s0g[n] = 0.0;
  for (int h = 0; h < G; ++h)
    for (int n = 0; n < N; ++n)
      if (g != h) {
        double sg = adjoint ? pb->sigs[IDX3(n, g, h, N, G)] : pb->sigs[IDX3(n, h, g, N, G)];
        s0g[n] = s0g[n] + sg * f0[IDX2(n, h, N)];
      }
  for (int n = 0; n < N; ++n) {
    if (transient) {
      bs[n] = pb->wtr[n] * pb->chi[IDX2(pb->mat[n] - 1, g, pb->nmat)] * fs0[n] + s0g[n] + pb->exsrc[IDX2(n, g, N)];
      continue;
    }
    double w = adjoint ? pb->nuf[IDX2(n, g, N)] : pb->chi[IDX2(pb->mat[n] - 1, g, pb->nmat)];
    bs[n] = w * fs0[n] / Keff + s0g[n] + pb->exsrc[IDX2(n, g, N)];
  }
}

static double integrate(const orc_problem *pb, const double *vdel, const double *s) {
  double intg = 0.0;
  for (int n = 0; n < pb->nnod; ++n) intg = intg + vdel[n] * s[n];
  return intg;
}

static double l2norm(const double *a, int N) {
  double x = 0.0;
  for (int n = 0; n < N; ++n) x = x + a[n] * a[n];
  return sqrt(x);
}

/* RelE (src/mod_cmfd.f90:1143-1168) */
static double rele(const double *nw, const double *od, int N) {
  double rel = 0.0;
  for (int n = 0; n < N; ++n)
    if (fabs(nw[n]) > 1.e-10) {
      double e = fabs(nw[n] - od[n]) / fabs(nw[n]);
      if (e > rel) rel = e;
    }
  return rel;
}

/* RelEg (src/mod_cmfd.f90:1172-1199): n outer, g inner */
static double releg(const double *nw, const double *od, int N, int G) {
  double rel = 0.0;
  for (int n = 0; n < N; ++n)
    for (int g = 0; g < G; ++g) {
      double a = nw[IDX2(n, g, N)];
      if (fabs(a) > 1.e-10) {
        double e = fabs(a - od[IDX2(n, g, N)]) / fabs(a);
        if (e > rel) rel = e;
      }
    }
  return rel;
}

/* outer (src/mod_cmfd.f90:415-509) and outer_ad (:602-699) */
int orc_outer(const orc_problem *pb, orc_state *st, const orc_iter *it, int mode, orc_trace *tr) {
  int N = pb->nnod, G = pb->ng;
  /* mode 0: outer, 1: outer_ad, 2: outer_fs (src/mod_cmfd.f90:513-598: the
     forward outer with k held at its value, no k update), 3: outer_tr
     (:800-880: TSrcTr sources, no k update, nodal update in cmode 2) */
  const int adjoint = mode == 1 || mode == 4, transient = mode == 3, fixed = mode == 2 || mode == 3;
  /* mode 4: outer_ad with popt = 0 (:677-679): no nodal update */
  const int nodal = it->kern != 0 && mode != 4;
  if (transient && !pb->wtr) return ORC_E_ARG;
  geo_t q;
  int rc = geo_init(&q, pb);
  if (rc != ORC_OK) { geo_free(&q); return rc; }
  double *A = (double *)malloc((size_t)N * G * 7 * sizeof(double));
  double *vdel = (double *)malloc((size_t)N * sizeof(double));
  double *fs0c = (double *)malloc((size_t)N * sizeof(double));
  double *bs = (double *)malloc((size_t)N * sizeof(double));
  double *s0g = (double *)malloc((size_t)N * sizeof(double));
  double *errn = (double *)malloc((size_t)N * sizeof(double));
  double *f0c = (double *)malloc((size_t)N * G * sizeof(double));
  if (!A || !vdel || !fs0c || !bs || !s0g || !errn || !f0c) { rc = ORC_E_ARG; goto done; }
  for (int n = 0; n < N; ++n)
    vdel[n] = pb->xdel[pb->ix[n] - 1] * pb->ydel[pb->iy[n] - 1] * pb->zdel[pb->iz[n] - 1];

  /* matrix_setup(1) */
  coup_coef_q(&q, st->df);
  matrix_setup_q(&q, st->df, st->dn, A);
  if (st->first) {
    st->Ke = 1.0;
    for (size_t c = 0; c < (size_t)N * G; ++c) st->f0[c] = 1.0;
    fsrc(pb, st->f0, st->fs0, adjoint);
    st->first = 0;
  }
  double f = integrate(pb, vdel, st->fs0);
  for (int n = 0; n < N; ++n) errn[n] = 1.0;
  double e1 = integrate(pb, vdel, errn), e2;
  if (tr) tr->n = 0;
  int p;
  for (p = 1; p <= it->nout; ++p) {
    double fc = f;
    memcpy(fs0c, st->fs0, (size_t)N * sizeof(double));
    memcpy(f0c, st->f0, (size_t)N * G * sizeof(double));
    double Keo = st->Ke;
    for (int gg = 0; gg < G; ++gg) {
      int g = adjoint ? G - 1 - gg : gg;
      tsrc(pb, g, st->Ke, st->f0, st->fs0, s0g, bs, adjoint, transient);
      bicg_q(&q, A, it->nin, g, bs, st->f0 + (size_t)N * g);
    }
    fsrc(pb, st->f0, st->fs0, adjoint);
    for (int n = 0; n < N; ++n) errn[n] = st->fs0[n] - fs0c[n];
    e2 = l2norm(errn, N);
    int flags = 0;
    if (p % it->nac == 0) {
      /* fiss_extrp (src/mod_cmfd.f90:308-335) */
      double domiR = e2 / e1;
      for (int n = 0; n < N; ++n) st->fs0[n] = st->fs0[n] + domiR / (1.0 - domiR) * errn[n];
      flags |= 1;
    }
    e1 = e2;
    if (!fixed) {
      f = integrate(pb, vdel, st->fs0);
      st->Ke = Keo * f / fc;
    }
    double ser = rele(st->fs0, fs0c, N);
    double fer = releg(st->f0, f0c, N, G);
    if (p % it->nupd == 0 && nodal) {
      /* nodal_upd (src/mod_cmfd.f90:339-383) */
      st->ndmax = 0.0;
      rc = nodal_update_q(&q, st, it->kern, adjoint ? 0 : (transient ? 2 : 1));
      flags |= 2;
      if (rc != ORC_OK) goto done;
      matrix_setup_q(&q, st->df, st->dn, A);
    }
    if (tr && tr->n < tr->cap) {
      int t = tr->n++;
      tr->k[t] = st->Ke; tr->ser[t] = ser; tr->fer[t] = fer; tr->flags[t] = flags;
      tr->ndmax[t] = st->ndmax;
      tr->loc[3 * t] = st->im; tr->loc[3 * t + 1] = st->jm; tr->loc[3 * t + 2] = st->km;
    }
    if (ser < it->serc && fer < it->ferc && st->ndmax < 1.e-2) break;
  }
  if (p > it->nout) rc = ORC_E_MAX_OUTER;
done:
  free(A); free(vdel); free(fs0c); free(bs); free(s0g); free(errn); free(f0c);
  geo_free(&q);
  return rc;
}

/* PowDis (src/mod_cmfd.f90:1290-1333) */
int orc_powdis(const orc_problem *pb, const double *f0, double *p) {
  int N = pb->nnod, G = pb->ng;
  for (int n = 0; n < N; ++n) p[n] = 0.0;
  for (int g = 0; g < G; ++g)
    for (int n = 0; n < N; ++n) {
      double vd = pb->xdel[pb->ix[n] - 1] * pb->ydel[pb->iy[n] - 1] * pb->zdel[pb->iz[n] - 1];
      double pw = f0[IDX2(n, g, N)] * pb->sigf[IDX2(n, g, N)] * vd;
      if (pw < 0.0) pw = 0.0;
      p[n] = p[n] + pw;
    }
  double tpow = 0.0;
  for (int n = 0; n < N; ++n) tpow = tpow + p[n];
  if (tpow <= 0.0) return ORC_E_POWER_NONPOS;
  for (int n = 0; n < N; ++n) p[n] = p[n] / tpow;
  return ORC_OK;
}

void orc_sanm_coef(double sigr, double D, double h, double *out6) {
  double alp = 0.5 * sqrt(sigr / D) * h;
  double alp2 = alp * alp;
  double sh = sinh(alp), ch = cosh(alp);
  double m0c = sh / alp;
  double m1s = 3.0 * (ch / alp - sh / alp2);
  double m2c = 5.0 * (sh / alp - 3.0 * ch / alp2 + 3.0 * sh / (alp * alp * alp));
  out6[0] = (sh - m1s) / (alp2 * m1s);
  out6[1] = (ch - m0c - m2c) / (alp2 * m2c);
  out6[2] = (m0c / m2c - 3.0 / alp2);
  out6[3] = (alp * ch - m1s) / (alp2 * m1s);
  out6[4] = (alp * sh - 3.0 * m2c) / (ch - m0c - m2c);
  out6[5] = (alp * ch - m1s) / (sh - m1s);
}

/* ---- cross-section update: xsec::XS_updt (src/mod_xsec.f90:11-46) ----- */

int orc_xs_update(const orc_problem *pb, const orc_xs_tables *t, double bcon, const double *ftem,
                  const double *mtem, const double *cden, const double *bpos, double *sigtr,
                  double *siga, double *nuf, double *sigf, double *sigs, double *D, double *sigr) {
  const int N = pb->nnod, G = pb->ng, M = t->nmat;
  double *X[5] = {sigtr, siga, nuf, sigf, sigs};
  for (int k = 0; k < 5; ++k)
    if (!t->tab[0][k]) return ORC_E_ARG;
  for (int n = 0; n < N; ++n)
    if (pb->mat[n] < 1 || pb->mat[n] > M) return ORC_E_ARG;
  if ((t->tab[2][0] && !ftem) || (t->tab[3][0] && !mtem) || (t->tab[4][0] && !cden) ||
      (t->tab[5][0] && (!bpos || !t->fbmap)))
    return ORC_E_ARG;

  /* base_updt (:172-196) */
  for (int n = 0; n < N; ++n) {
    const int m = pb->mat[n] - 1;
    for (int k = 0; k < 4; ++k)
      for (int g = 0; g < G; ++g) X[k][n + (long)N * g] = t->tab[0][k][m + (long)M * g];
    for (int h = 0; h < G; ++h)
      for (int g = 0; g < G; ++g) sigs[n + (long)N * (g + G * h)] = t->tab[0][4][m + (long)M * (g + G * h)];
  }
  /* bcon_updt, ftem_updt, mtem_updt, cden_updt (:336-436): one pass each,
     X(n,g) = X(n,g) + table(mat(n),g) * (parameter - reference value) */
  for (int kind = 1; kind <= 4; ++kind) {
    if (!t->tab[kind][0]) continue;
    for (int n = 0; n < N; ++n) {
      const int m = pb->mat[n] - 1;
      double dv;
      if (kind == 1) dv = bcon - t->rbcon;
      else if (kind == 2) dv = sqrt(ftem[n]) - sqrt(t->rftem);
      else if (kind == 3) dv = mtem[n] - t->rmtem;
      else dv = cden[n] - t->rcden;
      for (int g = 0; g < G; ++g) {
        for (int k = 0; k < 4; ++k)
          X[k][n + (long)N * g] = X[k][n + (long)N * g] + t->tab[kind][k][m + (long)M * g] * dv;
        for (int h = 0; h < G; ++h)
          sigs[n + (long)N * (g + G * h)] =
              sigs[n + (long)N * (g + G * h)] + t->tab[kind][4][m + (long)M * (g + G * h)] * dv;
      }
    }
  }
  /* crod_updt (:230-287): from the top of each rodded column down, whole
     nodes get the rod increment until the node holding the rod tip, which
     gets it weighted by its rodded fraction; then negative siga, nuf, sigf
     and sigs of the column are set to zero */
  if (t->tab[5][0]) {
    const int nx = pb->nxx, ny = pb->nyy, nz = pb->nzz;
    int *xyz = (int *)calloc((size_t)nx * ny * nz, sizeof(int));
    if (!xyz) return ORC_E_ARG;
    for (int n = 0; n < N; ++n)
      xyz[(pb->ix[n] - 1) + (long)nx * ((pb->iy[n] - 1) + (long)ny * (pb->iz[n] - 1))] = n + 1;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int b = t->fbmap[i + nx * j];
        if (b <= 0) continue;
        if (b > t->nb) { free(xyz); return ORC_E_ARG; }
        for (int k = 0; k < nz; ++k)
          if (!xyz[i + (long)nx * (j + (long)ny * k)]) { free(xyz); return ORC_E_ARG; }
        const double rodh = t->coreh - t->pos0 - bpos[b - 1] * t->ssize;
        double dum = 0.0;
        for (int k = nz - 1; k >= 0; --k) {
          const int n = xyz[i + (long)nx * (j + (long)ny * k)] - 1, m = pb->mat[n] - 1;
          const int part = rodh >= dum && rodh <= dum + pb->zdel[k];
          const double vf = part ? (rodh - dum) / pb->zdel[k] : 0.0;
          for (int g = 0; g < G; ++g) {
            for (int kk = 0; kk < 4; ++kk) {
              const double d = t->tab[5][kk][m + (long)M * g];
              X[kk][n + (long)N * g] = X[kk][n + (long)N * g] + (part ? vf * d : d);
            }
            for (int h = 0; h < G; ++h) {
              const double d = t->tab[5][4][m + (long)M * (g + G * h)];
              sigs[n + (long)N * (g + G * h)] = sigs[n + (long)N * (g + G * h)] + (part ? vf * d : d);
            }
          }
          if (part) break;
          dum = dum + pb->zdel[k];
        }
        for (int k = nz - 1; k >= 0; --k) {
          const int n = xyz[i + (long)nx * (j + (long)ny * k)] - 1;
          for (int g = 0; g < G; ++g) {
            for (int kk = 1; kk < 4; ++kk)
              if (X[kk][n + (long)N * g] < 0.0) X[kk][n + (long)N * g] = 0.0;
            for (int h = 0; h < G; ++h)
              if (sigs[n + (long)N * (g + G * h)] < 0.0) sigs[n + (long)N * (g + G * h)] = 0.0;
          }
        }
      }
    free(xyz);
  }
  /* Dsigr_updt (:199-226) and check_xs (:86-150) */
  for (int n = 0; n < N; ++n)
    for (int g = 0; g < G; ++g) {
      /* the thresholds are default-real (single precision) literals there */
      if (sigtr[n + (long)N * g] < (double)1.e-5f) return ORC_E_XS;
      D[n + (long)N * g] = 1.0 / (3.0 * sigtr[n + (long)N * g]);
      double dum = 0.0;
      for (int h = 0; h < G; ++h)
        if (g != h) dum = dum + sigs[n + (long)N * (g + G * h)];
      sigr[n + (long)N * g] = siga[n + (long)N * g] + dum;
    }
  for (long i = 0; i < (long)N * G; ++i)
    if (D[i] < (double)1.e-20f || sigr[i] < 0.0 || nuf[i] < 0.0) return ORC_E_XS;
  for (long i = 0; i < (long)N * G * G; ++i)
    if (sigs[i] < 0.0) return ORC_E_XS;
  return ORC_OK;
}

/* ---- branch-table update: xsec::XStab_updt (src/mod_xsec.f90:48-84) ---- */

/* the two branch points around x (brInterp's bracket search, :553-660):
   inside the range the first interval holding x; outside, the end interval
   when x is within 20 % (rel) or 100 ppm (boron) of the end point, both
   default-real literals there; otherwise the reference stops */
static int br_bracket(const double *p, int n, double x, int rel, int *i1, int *i2) {
  *i1 = 0;
  *i2 = 0;
  if (n <= 1) return 0;
  if (x >= p[0] && x <= p[n - 1]) {
    for (int s = 1; s < n; ++s)
      if (x >= p[s - 1] && x <= p[s]) {
        *i1 = s - 1;
        *i2 = s;
        break;
      }
  } else if (x < p[0] && (rel ? (p[0] - x) / p[0] < (double)0.2f : (p[0] - x) < (double)100.f)) {
    *i1 = 0;
    *i2 = 1;
  } else if (x > p[n - 1] && (rel ? (x - p[n - 1]) / p[n - 1] < (double)0.2f : (x - p[n - 1]) < (double)100.f)) {
    *i1 = n - 2;
    *i2 = n - 1;
  } else {
    return ORC_E_XS;
  }
  return 0;
}

/* brInterp (:520-788) for one material: tab is its (rodded or unrodded)
   record block, P its points; out (E) */
static int br_interp(const int *dm, const double *P, const double *tab, int E, double cd, double bc, double ft,
                     double mt, double *xs /* 8 E scratch */, double *out) {
  const int nd = dm[0], nb = dm[1], nf = dm[2], nm = dm[3];
  const double *pd = P, *pbr = P + nd, *pf = P + nd + nb, *pm = P + nd + nb + nf;
  int s1, s2, t1, t2, u1, u2, v1, v2;
  if (br_bracket(pd, nd, cd, 1, &s1, &s2) || br_bracket(pbr, nb, bc, 0, &t1, &t2) ||
      br_bracket(pf, nf, ft, 1, &u1, &u2) || br_bracket(pm, nm, mt, 1, &v1, &v2))
    return ORC_E_XS;
#define REC(s, t, u, v) (tab + (long)E * ((s) + (long)nd * ((t) + (long)nb * ((u) + (long)nf * (v)))))
  const int cs[8] = {s1, s1, s1, s1, s2, s2, s2, s2}, ct[8] = {t1, t1, t2, t2, t1, t1, t2, t2};
  const int cu[8] = {u1, u2, u1, u2, u1, u2, u1, u2};
  /* moderator temperature */
  if (nm > 1) {
    const double radx = (mt - pm[v1]) / (pm[v2] - pm[v1]);
    for (int c = 0; c < 8; ++c) {
      const double *A = REC(cs[c], ct[c], cu[c], v1), *B = REC(cs[c], ct[c], cu[c], v2);
      for (int e = 0; e < E; ++e) xs[c * E + e] = A[e] + radx * (B[e] - A[e]);
    }
  } else {
    for (int c = 0; c < 8; ++c) memcpy(xs + c * E, REC(cs[c], ct[c], cu[c], v1), sizeof(double) * E);
  }
#undef REC
  /* fuel temperature: pairs (1,2) (3,4) (5,6) (7,8) */
  if (nf > 1) {
    const double radx = (ft - pf[u1]) / (pf[u2] - pf[u1]);
    for (int c = 0; c < 8; c += 2)
      for (int e = 0; e < E; ++e) xs[c * E + e] = xs[c * E + e] + radx * (xs[(c + 1) * E + e] - xs[c * E + e]);
  }
  /* boron: pairs (1,3) (5,7) */
  if (nb > 1) {
    const double radx = (bc - pbr[t1]) / (pbr[t2] - pbr[t1]);
    for (int c = 0; c < 8; c += 4)
      for (int e = 0; e < E; ++e) xs[c * E + e] = xs[c * E + e] + radx * (xs[(c + 2) * E + e] - xs[c * E + e]);
  }
  /* coolant density: (1,5) */
  if (nd > 1) {
    const double r = (cd - pd[s1]) / (pd[s2] - pd[s1]);
    for (int e = 0; e < E; ++e) xs[e] = xs[e] + r * (xs[4 * E + e] - xs[e]);
  }
  memcpy(out, xs, sizeof(double) * E);
  return 0;
}

int orc_xtab_update(const orc_problem *pb, const orc_xtab_tables *t, double bcon, const double *ftem,
                    const double *mtem, const double *cden, const double *bpos, double *sigtr,
                    double *siga, double *nuf, double *sigf, double *sigs, double *D, double *sigr,
                    double *dc) {
  const int N = pb->nnod, G = pb->ng, M = t->nmat, E = 4 * G + G * G + 6 * G;
  if (!t->dims || !t->points || !t->xsec || !ftem || !mtem || !cden) return ORC_E_ARG;
  for (int n = 0; n < N; ++n)
    if (pb->mat[n] < 1 || pb->mat[n] > M) return ORC_E_ARG;
  long *poff = (long *)calloc((size_t)M, sizeof(long)), *xoff = (long *)calloc((size_t)M, sizeof(long));
  long *roff = (long *)calloc((size_t)M, sizeof(long));
  double *xs = (double *)malloc(sizeof(double) * 8 * E), *rec = (double *)malloc(sizeof(double) * E);
  double *rr = (double *)malloc(sizeof(double) * E);
  int *xyz = NULL, rc = ORC_OK;
  if (!poff || !xoff || !roff || !xs || !rec || !rr) { rc = ORC_E_ARG; goto done; }
  {
    long p = 0, x = 0, r = 0;
    for (int m = 0; m < M; ++m) {
      const int *dm = t->dims + 5 * m;
      const long nr = (long)dm[0] * dm[1] * dm[2] * dm[3];
      poff[m] = p; xoff[m] = x; roff[m] = dm[4] == 1 ? r : -1;
      p += dm[0] + dm[1] + dm[2] + dm[3];
      x += nr * E;
      if (dm[4] == 1) r += nr * E;
    }
  }
  /* the unrodded interpolation of every node (XStab_updt :70-74) */
  for (int n = 0; n < N; ++n) {
    const int m = pb->mat[n] - 1;
    if (br_interp(t->dims + 5 * m, t->points + poff[m], t->xsec + xoff[m], E, cden[n], bcon, ftem[n], mtem[n],
                  xs, rec)) { rc = ORC_E_XS; goto done; }
    for (int g = 0; g < G; ++g) {
      sigtr[n + (long)N * g] = rec[g];
      siga[n + (long)N * g] = rec[G + g];
      nuf[n + (long)N * g] = rec[2 * G + g];
      sigf[n + (long)N * g] = rec[3 * G + g];
    }
    for (int h = 0; h < G; ++h)
      for (int g = 0; g < G; ++g) sigs[n + (long)N * (g + G * h)] = rec[4 * G + g + G * h];
    for (int f = 0; f < 6; ++f)
      for (int g = 0; g < G; ++g) dc[n + (long)N * (g + G * f)] = rec[4 * G + G * G + g + G * f];
  }
  /* crod_tab_updt (:291-376): rodded columns from the top down to the tip,
     (1 - vfrac) X + vfrac R at the tip, R above it; then the column's
     negative siga, nuf, sigf, sigs and dc set to zero */
  if (t->fbmap) {
    const int nx = pb->nxx, ny = pb->nyy, nz = pb->nzz;
    if (!bpos || !t->rxsec) { rc = ORC_E_ARG; goto done; }
    xyz = (int *)calloc((size_t)nx * ny * nz, sizeof(int));
    if (!xyz) { rc = ORC_E_ARG; goto done; }
    for (int n = 0; n < N; ++n)
      xyz[(pb->ix[n] - 1) + (long)nx * ((pb->iy[n] - 1) + (long)ny * (pb->iz[n] - 1))] = n + 1;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int b = t->fbmap[i + nx * j];
        if (b <= 0) continue;
        if (b > t->nb) { rc = ORC_E_ARG; goto done; }
        for (int k = 0; k < nz; ++k)
          if (!xyz[i + (long)nx * (j + (long)ny * k)]) { rc = ORC_E_ARG; goto done; }
        const double rodh = t->coreh - t->pos0 - bpos[b - 1] * t->ssize;
        double dum = 0.0;
        for (int k = nz - 1; k >= 0; --k) {
          const int n = xyz[i + (long)nx * (j + (long)ny * k)] - 1, m = pb->mat[n] - 1;
          if (roff[m] < 0) { rc = ORC_E_XS; goto done; }
          if (br_interp(t->dims + 5 * m, t->points + poff[m], t->rxsec + roff[m], E, cden[n], bcon, ftem[n],
                        mtem[n], xs, rr)) { rc = ORC_E_XS; goto done; }
          const int part = rodh >= dum && rodh <= dum + pb->zdel[k];
          const double vf = part ? (rodh - dum) / pb->zdel[k] : 0.0;
          double *X[4] = {sigtr, siga, nuf, sigf};
          for (int q = 0; q < 4; ++q)
            for (int g = 0; g < G; ++g) {
              double *x = &X[q][n + (long)N * g];
              *x = part ? (1.0 - vf) * *x + vf * rr[q * G + g] : rr[q * G + g];
            }
          for (int h = 0; h < G; ++h)
            for (int g = 0; g < G; ++g) {
              double *x = &sigs[n + (long)N * (g + G * h)];
              *x = part ? (1.0 - vf) * *x + vf * rr[4 * G + g + G * h] : rr[4 * G + g + G * h];
            }
          for (int f = 0; f < 6; ++f)
            for (int g = 0; g < G; ++g) {
              double *x = &dc[n + (long)N * (g + G * f)];
              const double r = rr[4 * G + G * G + g + G * f];
              *x = part ? (1.0 - vf) * *x + vf * r : r;
            }
          if (part) break;
          dum = dum + pb->zdel[k];
        }
        for (int k = nz - 1; k >= 0; --k) {
          const int n = xyz[i + (long)nx * (j + (long)ny * k)] - 1;
          for (int g = 0; g < G; ++g) {
            if (siga[n + (long)N * g] < 0.0) siga[n + (long)N * g] = 0.0;
            if (nuf[n + (long)N * g] < 0.0) nuf[n + (long)N * g] = 0.0;
            if (sigf[n + (long)N * g] < 0.0) sigf[n + (long)N * g] = 0.0;
            for (int h = 0; h < G; ++h)
              if (sigs[n + (long)N * (g + G * h)] < 0.0) sigs[n + (long)N * (g + G * h)] = 0.0;
            for (int f = 0; f < 6; ++f)
              if (dc[n + (long)N * (g + G * f)] < 0.0) dc[n + (long)N * (g + G * f)] = 0.0;
          }
        }
      }
  }
  /* Dsigr_updt (:199-226) and check_xs (:86-150) */
  for (int n = 0; n < N; ++n)
    for (int g = 0; g < G; ++g) {
      if (sigtr[n + (long)N * g] < (double)1.e-5f) { rc = ORC_E_XS; goto done; }
      D[n + (long)N * g] = 1.0 / (3.0 * sigtr[n + (long)N * g]);
      double dum = 0.0;
      for (int h = 0; h < G; ++h)
        if (g != h) dum = dum + sigs[n + (long)N * (g + G * h)];
      sigr[n + (long)N * g] = siga[n + (long)N * g] + dum;
    }
  for (long i = 0; i < (long)N * G; ++i)
    if (D[i] < (double)1.e-20f || sigr[i] < 0.0 || nuf[i] < 0.0) { rc = ORC_E_XS; goto done; }
  for (long i = 0; i < (long)N * G * G; ++i)
    if (sigs[i] < 0.0) { rc = ORC_E_XS; goto done; }
done:
  free(poff); free(xoff); free(roff); free(xs); free(rec); free(rr); free(xyz);
  return rc;
}

/* ---- output maps: io::AsmPow (src/mod_io.f90:3267-3403), AxiPow (:3407) -- */

int orc_power_maps(const orc_problem *pb, int nx, int ny, int nz, const int *xdiv, const int *ydiv,
                   const int *zdiv, const double *pow, double *fasm, int *pmax, double *faxi, int *amax) {
  const int N = pb->nnod, NX = pb->nxx, NY = pb->nyy, NZ = pb->nzz;
  int sx = 0, sy = 0, sz = 0;
  for (int i = 0; i < nx; ++i) sx += xdiv[i];
  for (int j = 0; j < ny; ++j) sy += ydiv[j];
  for (int k = 0; k < nz; ++k) sz += zdiv[k];
  if (sx != NX || sy != NY || sz != NZ) return ORC_E_ARG;
  double *fx = (double *)calloc((size_t)NX * NY * NZ, sizeof(double));
  double *vd = (double *)calloc((size_t)NX * NY * NZ, sizeof(double));
  double *fnode = (double *)calloc((size_t)NX * NY, sizeof(double));
  if (!fx || !vd || !fnode) { free(fx); free(vd); free(fnode); return ORC_E_ARG; }
#define AT(i, j, k) ((i) + (long)NX * ((j) + (long)NY * (k)))
  for (int n = 0; n < N; ++n) {
    const int i = pb->ix[n] - 1, j = pb->iy[n] - 1, k = pb->iz[n] - 1;
    fx[AT(i, j, k)] = pow[n];
    vd[AT(i, j, k)] = pb->xdel[i] * pb->ydel[j] * pb->zdel[k];   /* vdel, src/mod_io.f90:1343 */
  }
  /* AsmPow: axially averaged node power, then area-weighted assembly power */
  for (int j = 0; j < NY; ++j)
    for (int i = pb->ystag_min[j] - 1; i <= pb->ystag_max[j] - 1; ++i) {
      double summ = 0.0, vsumm = 0.0;
      for (int k = 0; k < NZ; ++k) {
        summ = summ + fx[AT(i, j, k)] * pb->zdel[k];
        vsumm = vsumm + pb->zdel[k];
      }
      fnode[i + (long)NX * j] = summ / vsumm;
    }
  int nfuel = 0;
  double totp = 0.0;
  int ys = 0, yf = 0;
  for (int j = 0; j < ny; ++j) {
    yf += ydiv[j];
    int xs = 0, xf = 0;
    for (int i = 0; i < nx; ++i) {
      xf += xdiv[i];
      double summ = 0.0, vsumm = 0.0;
      for (int ly = ys; ly < yf; ++ly)
        for (int lx = xs; lx < xf; ++lx) {
          summ = summ + fnode[lx + (long)NX * ly] * pb->xdel[lx] * pb->ydel[ly];
          vsumm = vsumm + pb->xdel[lx] * pb->ydel[ly];
        }
      fasm[i + nx * j] = summ / vsumm;
      xs += xdiv[i];
      if (fasm[i + nx * j] > 0.0) { nfuel++; totp = totp + fasm[i + nx * j]; }
    }
    ys += ydiv[j];
  }
  pmax[0] = 1; pmax[1] = 1;
  double fmax = 0.0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      if (totp > 0.0) fasm[i + nx * j] = (double)nfuel / totp * fasm[i + nx * j];
      if (fasm[i + nx * j] > fmax) { pmax[0] = i + 1; pmax[1] = j + 1; fmax = fasm[i + nx * j]; }
    }
  /* AxiPow: power over volume per axial plane (the node volumes of every
     footprint position, zero where a plane has no node there) */
  nfuel = 0;
  totp = 0.0;
  int zt = 0;
  for (int k = 0; k < nz; ++k) {
    double summ = 0.0, vsumm = 0.0;
    for (int lz = 0; lz < zdiv[k]; ++lz) {
      for (int j = 0; j < NY; ++j)
        for (int i = pb->ystag_min[j] - 1; i <= pb->ystag_max[j] - 1; ++i) {
          summ = summ + fx[AT(i, j, zt)];
          vsumm = vsumm + vd[AT(i, j, zt)];
        }
      zt++;
    }
    faxi[k] = summ / vsumm;
    if (faxi[k] > 0.0) { nfuel++; totp = totp + faxi[k]; }
  }
  *amax = 1;
  fmax = 0.0;
  for (int k = 0; k < nz; ++k) {
    faxi[k] = (double)nfuel / totp * faxi[k];
    if (faxi[k] > fmax) { *amax = k + 1; fmax = faxi[k]; }
  }
#undef AT
  free(fx); free(vd); free(fnode);
  return ORC_OK;
}

/* io::AsmFlux (src/mod_io.f90:3502-3644) */
int orc_flux_map(const orc_problem *pb, int nx, int ny, const int *xdiv, const int *ydiv, const double *f0,
                 double norm, double *fasm, int *negf) {
  const int N = pb->nnod, G = pb->ng, NX = pb->nxx, NY = pb->nyy, NZ = pb->nzz;
  int sx = 0, sy = 0;
  for (int i = 0; i < nx; ++i) sx += xdiv[i];
  for (int j = 0; j < ny; ++j) sy += ydiv[j];
  if (sx != NX || sy != NY) return ORC_E_ARG;
  double *fx = (double *)calloc((size_t)NX * NY * NZ * G, sizeof(double));
  double *fnode = (double *)calloc((size_t)NX * NY * G, sizeof(double));
  double *totp = (double *)calloc((size_t)G, sizeof(double));
  if (!fx || !fnode || !totp) { free(fx); free(fnode); free(totp); return ORC_E_ARG; }
  for (int g = 0; g < G; ++g)
    for (int n = 0; n < N; ++n)
      fx[(pb->ix[n] - 1) + (long)NX * ((pb->iy[n] - 1) + (long)NY * ((pb->iz[n] - 1) + (long)NZ * g))] =
          f0[n + (long)N * g];
  for (int g = 0; g < G; ++g)
    for (int j = 0; j < NY; ++j)
      for (int i = pb->ystag_min[j] - 1; i <= pb->ystag_max[j] - 1; ++i) {
        double summ = 0.0, vsumm = 0.0;
        for (int k = 0; k < NZ; ++k) {
          summ = summ + fx[i + (long)NX * (j + (long)NY * (k + (long)NZ * g))] * pb->xdel[i] * pb->ydel[j] *
                            pb->zdel[k];
          vsumm = vsumm + pb->xdel[i] * pb->ydel[j] * pb->zdel[k];
        }
        fnode[i + (long)NX * (j + (long)NY * g)] = summ / vsumm;
      }
  *negf = 0;
  for (int g = 0; g < G; ++g) {
    totp[g] = 0.0;
    int ys = 0, yf = 0;
    for (int j = 0; j < ny; ++j) {
      yf += ydiv[j];
      int xs = 0, xf = 0;
      for (int i = 0; i < nx; ++i) {
        xf += xdiv[i];
        double summ = 0.0, vsumm = 0.0;
        for (int ly = ys; ly < yf; ++ly)
          for (int lx = xs; lx < xf; ++lx) {
            summ = summ + fnode[lx + (long)NX * (ly + (long)NY * g)] * pb->xdel[lx] * pb->ydel[ly];
            vsumm = vsumm + pb->xdel[lx] * pb->ydel[ly];
          }
        double *f = &fasm[i + (long)nx * (j + (long)ny * g)];
        *f = summ / vsumm;
        xs += xdiv[i];
        if (*f > 0.0) totp[g] = totp[g] + *f;
        if (*f < 0.0) *negf = 1;
      }
      ys += ydiv[j];
    }
  }
  if (norm > 0.0)
    for (int g = 0; g < G; ++g)
      for (int t = 0; t < nx * ny; ++t) fasm[t + (long)nx * ny * g] = norm / totp[g] * fasm[t + (long)nx * ny * g] * norm;
  free(fx); free(fnode); free(totp);
  return ORC_OK;
}
