// Host side of the gfx950 static k-eigenvalue engine: context, data
// movement between the reference's node list and the padded device box, the
// outer-iteration driver and the C-ABI of include/adpres_gpu.h.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/adpres_gpu.h"
#include "adg_cmfd.h"
#include "adg_common.h"
#include "adg_nodal.h"

using namespace adg;

namespace {

struct HipError {
  hipError_t e;
  const char *what;
};

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t _e = (x);                       \
    if (_e != hipSuccess) throw HipError{_e, #x}; \
  } while (0)

template <typename T>
T *dalloc(size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  HIPCHK(hipMemset(p, 0, n * sizeof(T)));
  return (T *)p;
}

}  // namespace

struct adg_ctx {
  int N = 0, G = 0, nmat = 0, nxx = 0, nyy = 0, nzz = 0, precision = 0;
  int NXP = 0, NYP = 0, NZP = 0;
  long sy = 0, sz = 0, NC = 0;
  int bcface[6] = {0, 0, 0, 0, 0, 0};
  std::vector<long> cell;          // node n -> padded cell
  std::vector<double> hx, hy, hz;
  hipStream_t s = nullptr;
  // device
  double *d_hx = nullptr, *d_hy = nullptr, *d_hz = nullptr;
  unsigned char *flag = nullptr;
  double *D = nullptr, *sigr = nullptr, *nuf = nullptr, *sigf = nullptr, *chi = nullptr, *sigs = nullptr;
  double *dc = nullptr, *exsrc = nullptr, *vdel = nullptr;
  double *f0 = nullptr, *f0c = nullptr, *fs0 = nullptr, *fs0c = nullptr, *errn = nullptr;
  double *dn = nullptr, *df = nullptr, *A64 = nullptr, *b = nullptr, *pw = nullptr, *tmp = nullptr;
  // BiCGSTAB vectors: FP64 form
  double *rs64 = nullptr, *t64 = nullptr, *rr64[2] = {nullptr, nullptr}, *pp64[2] = {nullptr, nullptr};
  double *vv64[2] = {nullptr, nullptr};
  // FP32 form
  float *A32 = nullptr, *rs32 = nullptr, *t32 = nullptr, *d32 = nullptr;
  float *rr32[2] = {nullptr, nullptr}, *pp32[2] = {nullptr, nullptr}, *vv32[2] = {nullptr, nullptr};
  // nodal
  double *S = nullptr, *a2 = nullptr, *a4 = nullptr, *L1 = nullptr, *npart = nullptr;
  int npart_blocks = 0;
  // reductions / scalars
  double *part = nullptr;
  unsigned *cnt = nullptr;
  Scal *sc = nullptr;
  Scal *hsc = nullptr;   // pinned host mirror
  int nblk = 1;
  bool xs_set = false, first = true;
  bool graphs = true;     // replay the outer iteration as a captured hipGraph
  int check_every = 8;    // outers between host polls of the done flag
  double Ke_host = 1.0;
  double timing[4] = {0, 0, 0, 0};
  std::string err;
  // profiling: event pairs recorded on the context stream during adg_outer
  struct ProfRec {
    int type, e0, e1;   // type 0 K1 kernel, 1 group solve, 2 nodal update
  };
  bool prof = false;
  std::vector<hipEvent_t> evs;
  int ev_next = 0;
  std::vector<ProfRec> recs;
  double prof_out[6] = {0, 0, 0, 0, 0, 0};
  // trace buffers (device) and the captured outer iteration
  double *tk = nullptr, *tser = nullptr, *tfer = nullptr, *tnd = nullptr;
  int *tflags = nullptr, *tloc = nullptr;
  int tcap = 0;
  hipGraphExec_t gexec = nullptr;
  int gkey[4] = {-1, -1, -1, -1};   // adjoint, nin, kern, nupd of the captured iteration

  Geo geo() const {
    Geo q;
    q.nxx = nxx; q.nyy = nyy; q.nzz = nzz; q.G = G;
    q.NXP = NXP; q.NYP = NYP;
    q.sy = sy; q.sz = sz; q.NC = NC;
    q.c0 = sz; q.M = (long)nzz * sz;
    for (int s6 = 0; s6 < 6; ++s6) q.bc[s6] = bcface[s6];
    q.hx = d_hx; q.hy = d_hy; q.hz = d_hz;
    q.flag = flag;
    return q;
  }
  XS xs() const {
    XS x;
    x.D = D; x.sigr = sigr; x.nuf = nuf; x.sigf = sigf; x.chi = chi; x.sigs = sigs;
    x.dc = dc; x.exsrc = exsrc; x.vdel = vdel;
    return x;
  }
};

namespace {

int fail(adg_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(adg_ctx *c, const HipError &e) {
  std::string m = std::string(e.what) + ": " + hipGetErrorString(e.e);
  return fail(c, ADG_E_HIP, m);
}

constexpr int MAX_PROF_EVENTS = 16384;

// returns the index of an event recorded now on the context stream, or -1
int prof_event(adg_ctx *c) {
  if (!c->prof || c->ev_next >= MAX_PROF_EVENTS) return -1;
  if (c->ev_next == (int)c->evs.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->evs.push_back(e);
  }
  const int k = c->ev_next++;
  HIPCHK(hipEventRecord(c->evs[k], c->s));
  return k;
}

// index of an unrecorded event for hipExtLaunchKernelGGL, which records its
// start/stop events at the kernel's own start and end, or -1
int prof_slot(adg_ctx *c) {
  if (!c->prof || c->ev_next >= MAX_PROF_EVENTS) return -1;
  if (c->ev_next == (int)c->evs.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->evs.push_back(e);
  }
  return c->ev_next++;
}

void prof_pair(adg_ctx *c, int type, int e0, int e1) {
  if (e0 >= 0 && e1 >= 0) c->recs.push_back({type, e0, e1});
}

dim3 cells_grid(const adg_ctx *c, int y) {
  long n = (long)c->nxx * c->nyy * c->nzz;
  return dim3((unsigned)((n + NT - 1) / NT), (unsigned)y, 1);
}

// scatter a Fortran (N, K) array into K padded boxes
void scatter(const adg_ctx *c, const double *src, int K, double *dst_dev, std::vector<double> &stage) {
  stage.assign((size_t)K * c->NC, 0.0);
  for (int k = 0; k < K; ++k)
    for (int n = 0; n < c->N; ++n) stage[(size_t)k * c->NC + c->cell[n]] = src[(size_t)k * c->N + n];
  HIPCHK(hipMemcpy(dst_dev, stage.data(), stage.size() * sizeof(double), hipMemcpyHostToDevice));
}

void gather(const adg_ctx *c, const double *src_dev, int K, double *dst, std::vector<double> &stage) {
  stage.resize((size_t)K * c->NC);
  HIPCHK(hipMemcpy(stage.data(), src_dev, stage.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < K; ++k)
    for (int n = 0; n < c->N; ++n) dst[(size_t)k * c->N + n] = stage[(size_t)k * c->NC + c->cell[n]];
}

void launch_matrix(adg_ctx *c, int sched) {
  Geo q = c->geo();
  if (c->precision == ADG_FP32_REFINE)
    hipLaunchKernelGGL(k_matrix<float>, cells_grid(c, c->G), dim3(NT), 0, c->s, q, c->xs(), c->df, c->dn, c->A64, c->A32,
                       c->sc, sched);
  else
    hipLaunchKernelGGL(k_matrix<double>, cells_grid(c, c->G), dim3(NT), 0, c->s, q, c->xs(), c->df, c->dn, c->A64,
                       (double *)nullptr, c->sc, sched);
  HIPCHK(hipGetLastError());
}

// bicg (src/mod_cmfd.f90:1203-1243) of group g: b in c->b, x = flux of g.
// Kernels: setup, then KA(1) KB(1) KA(2) KB(2) ... KB(nin), KT (two stencil
// kernels per iteration, see adg_cmfd.h).  r, p, v ping-pong: KA(it) reads
// set (it-1)&1 and writes set it&1.  Partials: setup -> P0, KA(1) P0 -> P1,
// KB P1 -> P2, KA(it>1) P2 -> P1, KT reads P2.
template <typename TV>
void launch_bicg_t(adg_ctx *c, int g, int nin, const TV *A, TV *const *R, TV *const *P, TV *const *V, TV *t,
                   TV *rs, TV *d) {
  Geo q = c->geo();
  const long NC = c->NC;
  double *x64 = c->f0 + (size_t)g * NC;
  double *P0 = c->part, *P1 = c->part + 4 * MAXBLK, *P2 = c->part + 8 * MAXBLK;
  const double *A64g = c->A64 + (size_t)g * 7 * NC;
  // FP64 form: the iterate is the flux itself; FP32 form: a correction d
  TV *x = d ? d : (TV *)x64;
  const int sample = nin > 1 ? 2 : 0;   // steady-state KA launch that is timed (none if nin = 1)
  int ek0 = -1, ek1 = -1;
  hipLaunchKernelGGL(k_bicg_setup<TV>, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->sc, A64g, c->b, x64, R[0], rs, d, P0);
  for (int it = 1; it <= nin; ++it) {
    const int rd = (it - 1) & 1, wr = it & 1;
    if (it == sample) {
      ek0 = prof_slot(c);
      ek1 = prof_slot(c);
    }
    if (it == sample && ek0 >= 0 && ek1 >= 0)
      hipExtLaunchKernelGGL((k_bicgA<TV, false>), dim3(c->nblk), dim3(NTF), 0, c->s, c->evs[ek0], c->evs[ek1], 0, q,
                            c->sc, A, it, R[rd], P[rd], V[rd], t, rs, x, R[wr], P[wr], V[wr], P2, P1);
    else if (it == 1)
      hipLaunchKernelGGL((k_bicgA<TV, true>), dim3(c->nblk), dim3(NTF), 0, c->s, q, c->sc, A, it, R[rd], P[rd], V[rd],
                         t, rs, x, R[wr], P[wr], V[wr], P0, P1);
    else
      hipLaunchKernelGGL((k_bicgA<TV, false>), dim3(c->nblk), dim3(NTF), 0, c->s, q, c->sc, A, it, R[rd], P[rd], V[rd],
                         t, rs, x, R[wr], P[wr], V[wr], P2, P1);
    hipLaunchKernelGGL(k_bicgB<TV>, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->sc, A, it, R[wr], V[wr], t, rs, P1, P2);
  }
  const int wr = nin & 1;
  hipLaunchKernelGGL(k_bicgT<TV>, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->sc, nin, R[wr], P[wr], V[wr], x,
                     d ? x64 : (double *)nullptr, P2);
  HIPCHK(hipGetLastError());
  prof_pair(c, 0, ek0, ek1);
}

void launch_bicg(adg_ctx *c, int g, int nin) {
  const long NC = c->NC;
  if (c->precision == ADG_FP32_REFINE)
    launch_bicg_t<float>(c, g, nin, c->A32 + (size_t)g * 7 * NC, c->rr32, c->pp32, c->vv32, c->t32, c->rs32, c->d32);
  else
    launch_bicg_t<double>(c, g, nin, c->A64 + (size_t)g * 7 * NC, c->rr64, c->pp64, c->vv64, c->t64, c->rs64,
                          (double *)nullptr);
}

// one group solve: TSrc + bicg (src/mod_cmfd.f90:472-478)
void group_solve(adg_ctx *c, int g, int adjoint, int first_group, int nin) {
  Geo q = c->geo();
  const int eg0 = prof_event(c);
  hipLaunchKernelGGL(k_src, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->xs(), c->sc, g, adjoint, first_group, c->f0, c->f0c,
                     c->fs0, c->fs0c, c->b);
  launch_bicg(c, g, nin);
  prof_pair(c, 1, eg0, prof_event(c));
}

// nodal::nodal_update(_pnm) + matrix rebuild (src/mod_cmfd.f90:339-383)
void nodal_update(adg_ctx *c, int kern, int cmode, int sched) {
  NodArgs a;
  a.q = c->geo(); a.xs = c->xs(); a.sc = c->sc;
  a.f0 = c->f0; a.df = c->df; a.dn = c->dn;
  a.S = c->S; a.a2 = c->a2; a.a4 = c->a4; a.L1 = c->L1;
  a.kern = kern; a.cmode = cmode; a.sched = sched;
  a.part = c->npart; a.cnt = c->cnt + 4;
  // pass 2 runs as a (blocks, 3) grid; blocks cover the largest direction
  const dim3 g2(c->npart_blocks, 3, 1);
  switch (c->G) {
#define NOD_CASE(GG)                                                                 \
  case GG:                                                                           \
    hipLaunchKernelGGL(k_nod_src, cells_grid(c, GG), dim3(NT), 0, c->s, a);          \
    hipLaunchKernelGGL(k_nod_pass1<GG>, cells_grid(c, 3), dim3(NT), 0, c->s, a);     \
    hipLaunchKernelGGL(k_nod_pass2<GG>, g2, dim3(NT), 0, c->s, a);                   \
    break;
    NOD_CASE(1) NOD_CASE(2) NOD_CASE(3) NOD_CASE(4) NOD_CASE(5) NOD_CASE(6) NOD_CASE(7) NOD_CASE(8)
#undef NOD_CASE
  }
  hipLaunchKernelGGL(k_nod_argmax, dim3(1), dim3(NT), 0, c->s, c->sc, c->npart, (int)(3 * c->npart_blocks), sched);
  HIPCHK(hipGetLastError());
  launch_matrix(c, sched);
}

void set_scal_start(adg_ctx *c, const adg_iter *it) {
  // host-side initialisation of the device scalars for an outer call
  Scal h;
  memset(&h, 0, sizeof(h));
  HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scal), hipMemcpyDeviceToHost));
  h.done = 0;
  h.status = 0;
  h.p = 0;
  h.flags = 0;
  h.lu_key = ~0ull;
  h.Ke = c->Ke_host;
  h.nac = it ? it->nac : 1;
  h.nupd = it ? it->nupd : 1;
  h.nodal_on = (it && it->kern != ADG_KERN_FDM) ? 1 : 0;
  h.serc = it ? it->serc : 0.0;
  h.ferc = it ? it->ferc : 0.0;
  HIPCHK(hipMemcpy(c->sc, &h, sizeof(Scal), hipMemcpyHostToDevice));
}

void ensure_trace(adg_ctx *c, int nout) {
  if (nout <= c->tcap) return;
  void *ptrs[] = {c->tk, c->tser, c->tfer, c->tnd, c->tflags, c->tloc};
  for (void *pp : ptrs)
    if (pp) (void)hipFree(pp);
  const int cap = std::max(nout, 2 * c->tcap);
  c->tk = dalloc<double>(cap); c->tser = dalloc<double>(cap); c->tfer = dalloc<double>(cap);
  c->tnd = dalloc<double>(cap); c->tflags = dalloc<int>(cap); c->tloc = dalloc<int>(3 * (size_t)cap);
  c->tcap = cap;
  if (c->gexec) {
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
  }
}

// One outer iteration (src/mod_cmfd.f90:465-496) as a fixed kernel sequence:
// G group solves, FSrc + error norms, extrapolation + k update, the nodal
// update (its kernels return at once unless this outer is a nodal-update
// outer -- the schedule lives on the device), and the end-of-outer kernel.
// p_host is the outer index as seen by the host (used for profiling only).
void enqueue_iteration(adg_ctx *c, int adjoint, int nin, int kern, int p_host, int nupd) {
  Geo q = c->geo();
  XS x = c->xs();
  for (int gg = 0; gg < c->G; ++gg) {
    const int g = adjoint ? c->G - 1 - gg : gg;
    group_solve(c, g, adjoint, gg == 0 ? 1 : 0, nin);
  }
  double *P1 = c->part + 4 * MAXBLK, *P2 = c->part + 8 * MAXBLK;
  hipLaunchKernelGGL(k_fsrc, dim3(c->nblk), dim3(NTF), 0, c->s, q, x, c->sc, adjoint, c->f0, c->f0c, c->fs0, c->fs0c,
                     c->errn, P1);
  hipLaunchKernelGGL(k_post, dim3(c->nblk), dim3(NTF), 0, c->s, q, x, c->sc, c->fs0, c->fs0c, c->errn, P1, P2);
  hipLaunchKernelGGL(k_fin_outer, dim3(1), dim3(NTF), 0, c->s, c->sc, P2, c->nblk);
  if (kern != ADG_KERN_FDM) {
    const bool timed = p_host > 0 && (p_host % nupd) == 0;
    const int en0 = timed ? prof_event(c) : -1;
    nodal_update(c, kern, adjoint ? 0 : 1, 1);
    if (timed) prof_pair(c, 2, en0, prof_event(c));
  }
  hipLaunchKernelGGL(k_end, dim3(1), dim3(64), 0, c->s, c->sc, c->tk, c->tser, c->tfer, c->tnd, c->tflags, c->tloc);
  HIPCHK(hipGetLastError());
}

}  // namespace

extern "C" {

const char *adg_status_string(int status) {
  switch (status) {
    case ADG_OK: return "ok";
    case ADG_E_MAX_OUTER: return "MAXIMUM NUMBER OF OUTER ITERATION IS REACHED";
    case ADG_E_NODAL_UNSTABLE: return "The two-node nonlinear iteration seems not stable";
    case ADG_E_LU_PIVOT: return "ERROR IN MATRIX DECOMP: DIAGONAL ELEMENTS CLOSE TO ZERO";
    case ADG_E_POWER_NONPOS: return "TOTAL NODES POWER IS ZERO OR LESS";
    case ADG_E_ARG: return "invalid argument";
    case ADG_E_HIP: return "HIP runtime error";
    case ADG_E_NOT_SET: return "cross sections not set";
  }
  return "unknown status";
}

const char *adg_last_error(adg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int adg_create(const adg_geom *g, adg_ctx **out) {
  if (!g || !out) return ADG_E_ARG;
  *out = nullptr;
  if (g->nnod <= 0 || g->ng <= 0 || g->ng > MAXG || g->nxx <= 0 || g->nyy <= 0 || g->nzz <= 0 || !g->ix || !g->iy ||
      !g->iz || !g->xdel || !g->ydel || !g->zdel)
    return ADG_E_ARG;
  if (g->precision != ADG_FP64 && g->precision != ADG_FP32_REFINE) return ADG_E_ARG;
  adg_ctx *c = new adg_ctx();
  try {
    c->N = g->nnod; c->G = g->ng; c->nmat = g->nmat;
    if (const char *ev = getenv("ADG_GRAPH")) c->graphs = atoi(ev) != 0;
    if (const char *ev = getenv("ADG_CHECK")) c->check_every = std::max(1, atoi(ev));
    c->nxx = g->nxx; c->nyy = g->nyy; c->nzz = g->nzz; c->precision = g->precision;
    c->NXP = c->nxx + 2; c->NYP = c->nyy + 2; c->NZP = c->nzz + 2;
    c->sy = c->NXP; c->sz = (long)c->NXP * c->NYP; c->NC = c->sz * c->NZP;
    // faces x+, x-, y+, y-, z+, z- <- xeast, xwest, ynorth, ysouth, ztop, zbott
    const int map[6] = {0, 1, 2, 3, 5, 4};
    for (int s6 = 0; s6 < 6; ++s6) {
      c->bcface[s6] = g->bc[map[s6]];
      if (c->bcface[s6] < 0 || c->bcface[s6] > 2) { delete c; return ADG_E_ARG; }
    }
    c->hx.assign(g->xdel, g->xdel + c->nxx);
    c->hy.assign(g->ydel, g->ydel + c->nyy);
    c->hz.assign(g->zdel, g->zdel + c->nzz);
    // node -> cell, active flags, footprint checks
    std::vector<unsigned char> fl(c->NC, 0);
    c->cell.resize(c->N);
    for (int n = 0; n < c->N; ++n) {
      int i = g->ix[n] - 1, j = g->iy[n] - 1, k = g->iz[n] - 1;
      if (i < 0 || i >= c->nxx || j < 0 || j >= c->nyy || k < 0 || k >= c->nzz) {
        c->err = "node coordinate outside the box";
        delete c;
        return ADG_E_ARG;
      }
      long cc = ((long)(k + 1) * c->NYP + (j + 1)) * c->NXP + (i + 1);
      if (fl[cc]) { c->err = "duplicate node coordinate"; delete c; return ADG_E_ARG; }
      fl[cc] = F_ACTIVE;
      c->cell[n] = cc;
    }
    const long st[3] = {1, c->sy, c->sz};
    for (int n = 0; n < c->N; ++n) {
      long cc = c->cell[n];
      for (int s6 = 0; s6 < 6; ++s6) {
        long nb = (s6 & 1) ? cc - st[s6 >> 1] : cc + st[s6 >> 1];
        if (fl[nb] & F_ACTIVE) fl[cc] |= f_nb(s6);
      }
    }
    // every line must be contiguous (the reference's footprint rule) and
    // hold at least two nodes (TLUpd1/2 read both neighbours of an end node)
    std::vector<double> vd(c->NC, 0.0);
    for (int n = 0; n < c->N; ++n) {
      long cc = c->cell[n];
      vd[cc] = c->hx[g->ix[n] - 1] * c->hy[g->iy[n] - 1] * c->hz[g->iz[n] - 1];
    }
    HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    const size_t NC = c->NC, G = c->G;
    c->d_hx = dalloc<double>(c->nxx); c->d_hy = dalloc<double>(c->nyy); c->d_hz = dalloc<double>(c->nzz);
    HIPCHK(hipMemcpy(c->d_hx, c->hx.data(), c->nxx * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_hy, c->hy.data(), c->nyy * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_hz, c->hz.data(), c->nzz * sizeof(double), hipMemcpyHostToDevice));
    c->flag = dalloc<unsigned char>(NC);
    HIPCHK(hipMemcpy(c->flag, fl.data(), NC, hipMemcpyHostToDevice));
    c->vdel = dalloc<double>(NC);
    HIPCHK(hipMemcpy(c->vdel, vd.data(), NC * sizeof(double), hipMemcpyHostToDevice));
    c->D = dalloc<double>(G * NC); c->sigr = dalloc<double>(G * NC); c->nuf = dalloc<double>(G * NC);
    c->sigf = dalloc<double>(G * NC); c->chi = dalloc<double>(G * NC); c->sigs = dalloc<double>(G * G * NC);
    c->dc = dalloc<double>(6 * G * NC); c->exsrc = dalloc<double>(G * NC);
    c->f0 = dalloc<double>(G * NC); c->f0c = dalloc<double>(G * NC); c->fs0 = dalloc<double>(NC);
    c->fs0c = dalloc<double>(NC); c->errn = dalloc<double>(NC);
    c->dn = dalloc<double>(6 * G * NC); c->df = dalloc<double>(6 * G * NC);
    c->A64 = dalloc<double>(7 * G * NC); c->b = dalloc<double>(NC); c->pw = dalloc<double>(NC);
    c->tmp = dalloc<double>(NC);
    if (c->precision == ADG_FP32_REFINE) {
      c->A32 = dalloc<float>(7 * G * NC);
      c->rs32 = dalloc<float>(NC); c->t32 = dalloc<float>(NC); c->d32 = dalloc<float>(NC);
      for (int k = 0; k < 2; ++k) {
        c->rr32[k] = dalloc<float>(NC); c->pp32[k] = dalloc<float>(NC); c->vv32[k] = dalloc<float>(NC);
      }
    } else {
      c->rs64 = dalloc<double>(NC); c->t64 = dalloc<double>(NC);
      for (int k = 0; k < 2; ++k) {
        c->rr64[k] = dalloc<double>(NC); c->pp64[k] = dalloc<double>(NC); c->vv64[k] = dalloc<double>(NC);
      }
    }
    c->S = dalloc<double>(3 * G * NC); c->a2 = dalloc<double>(3 * G * NC); c->a4 = dalloc<double>(3 * G * NC);
    c->L1 = dalloc<double>(3 * G * NC);
    long maxf = std::max({(long)(c->nxx + 1) * c->nyy * c->nzz, (long)(c->nyy + 1) * c->nxx * c->nzz,
                          (long)(c->nzz + 1) * c->nxx * c->nyy});
    c->npart_blocks = (int)((maxf + NT - 1) / NT);
    c->npart = dalloc<double>((size_t)3 * 3 * c->npart_blocks);
    c->part = dalloc<double>((size_t)MAXBLK * 12);   // three rotating buffers of 4 x MAXBLK
    c->cnt = dalloc<unsigned>(16);
    c->sc = dalloc<Scal>(1);
    HIPCHK(hipHostMalloc((void **)&c->hsc, sizeof(Scal), hipHostMallocDefault));
    long M = (long)c->nzz * c->sz;
    // grid of the flat kernels: one NTF-cell chunk per block up to MAXBLK,
    // rounded up to a multiple of 8 for the XCD-aware block mapping
    // (the FP32 form streams half the bytes and is best at half the grid)
    const long maxblk = c->precision == ADG_FP32_REFINE ? MAXBLK / 2 : MAXBLK;
    c->nblk = (int)std::min<long>((M + NTF - 1) / NTF, maxblk);
    c->nblk = (c->nblk + 7) / 8 * 8;
    // geometry checks that need the flags: single-node lines
    for (int n = 0; n < c->N; ++n) {
      unsigned char f = fl[c->cell[n]];
      for (int u = 0; u < 3; ++u)
        if (!(f & f_nb(2 * u)) && !(f & f_nb(2 * u + 1))) {
          c->err = "single-node line (the reference's transverse-leakage fit needs two nodes)";
          throw 1;
        }
    }
  } catch (const HipError &e) {
    int rc = hip_fail(c, e);
    fprintf(stderr, "adg_create: %s\n", c->err.c_str());
    adg_destroy(c);
    return rc;
  } catch (int) {
    fprintf(stderr, "adg_create: %s\n", c->err.c_str());
    adg_destroy(c);
    return ADG_E_ARG;
  }
  *out = c;
  return ADG_OK;
}

void adg_destroy(adg_ctx *c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  void *ptrs[] = {c->d_hx, c->d_hy, c->d_hz, c->flag, c->D, c->sigr, c->nuf, c->sigf, c->chi, c->sigs, c->dc,
                  c->exsrc, c->vdel, c->f0, c->f0c, c->fs0, c->fs0c, c->errn, c->dn, c->df, c->A64, c->b, c->pw,
                  c->tmp, c->rr64[0], c->rr64[1], c->rs64, c->t64, c->pp64[0], c->pp64[1], c->vv64[0], c->vv64[1],
                  c->A32, c->rr32[0], c->rr32[1], c->rs32, c->t32, c->d32, c->pp32[0], c->pp32[1], c->vv32[0],
                  c->vv32[1], c->S, c->a2, c->a4,
                  c->L1, c->npart, c->part, c->cnt, c->sc};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->hsc) (void)hipHostFree(c->hsc);
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  void *tptr[] = {c->tk, c->tser, c->tfer, c->tnd, c->tflags, c->tloc};
  for (void *p : tptr)
    if (p) (void)hipFree(p);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

int adg_set_xs(adg_ctx *c, const double *D, const double *sigr, const double *sigs, const double *nuf,
               const double *sigf, const double *chi, const int *mat, const double *dc, const double *exsrc) {
  if (!c) return ADG_E_ARG;
  if (!D || !sigr || !sigs || !nuf || !sigf || !chi || !mat || !dc) return fail(c, ADG_E_ARG, "null XS array");
  try {
    std::vector<double> st;
    const int N = c->N, G = c->G;
    for (size_t n = 0; n < (size_t)N * G; ++n)
      if (!(D[n] > 0.0)) return fail(c, ADG_E_ARG, "non-positive diffusion coefficient");
    for (int n = 0; n < N; ++n)
      if (mat[n] < 1 || mat[n] > c->nmat) return fail(c, ADG_E_ARG, "material index out of range");
    scatter(c, D, G, c->D, st);
    scatter(c, sigr, G, c->sigr, st);
    scatter(c, nuf, G, c->nuf, st);
    scatter(c, sigf, G, c->sigf, st);
    // sigs(n, from, to) -> device [to][from][cell]
    std::vector<double> tr((size_t)G * G * N);
    for (int f = 0; f < G; ++f)
      for (int t = 0; t < G; ++t)
        for (int n = 0; n < N; ++n) tr[((size_t)t * G + f) * N + n] = sigs[((size_t)t * G + f) * N + n];
    scatter(c, tr.data(), G * G, c->sigs, st);
    // chi(mat(n), g) expanded per node
    std::vector<double> ch((size_t)G * N);
    for (int g = 0; g < G; ++g)
      for (int n = 0; n < N; ++n) ch[(size_t)g * N + n] = chi[(size_t)g * c->nmat + (mat[n] - 1)];
    scatter(c, ch.data(), G, c->chi, st);
    scatter(c, dc, 6 * G, c->dc, st);
    if (exsrc) scatter(c, exsrc, G, c->exsrc, st);
    else HIPCHK(hipMemset(c->exsrc, 0, (size_t)G * c->NC * sizeof(double)));
    c->xs_set = true;
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_set_state(adg_ctx *c, const double *f0, const double *fs0, double Ke, const double *dn, int first) {
  if (!c) return ADG_E_ARG;
  try {
    std::vector<double> st;
    if (first) {
      c->first = true;
      HIPCHK(hipMemset(c->dn, 0, (size_t)6 * c->G * c->NC * sizeof(double)));
      return ADG_OK;
    }
    if (!f0 || !fs0 || !dn) return fail(c, ADG_E_ARG, "null state array");
    scatter(c, f0, c->G, c->f0, st);
    scatter(c, fs0, 1, c->fs0, st);
    scatter(c, dn, 6 * c->G, c->dn, st);
    c->Ke_host = Ke;
    c->first = false;
    Scal h;
    HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scal), hipMemcpyDeviceToHost));
    h.Ke = Ke;
    HIPCHK(hipMemcpy(c->sc, &h, sizeof(Scal), hipMemcpyHostToDevice));
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_get_state(adg_ctx *c, double *f0, double *fs0, double *Ke, double *dn, double *df) {
  if (!c) return ADG_E_ARG;
  try {
    std::vector<double> st;
    HIPCHK(hipStreamSynchronize(c->s));
    if (f0) gather(c, c->f0, c->G, f0, st);
    if (fs0) gather(c, c->fs0, 1, fs0, st);
    if (dn) gather(c, c->dn, 6 * c->G, dn, st);
    if (df) gather(c, c->df, 6 * c->G, df, st);
    if (Ke) {
      Scal h;
      HIPCHK(hipMemcpy(&h, c->sc, sizeof(Scal), hipMemcpyDeviceToHost));
      *Ke = h.Ke;
    }
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_outer(adg_ctx *c, int mode, const adg_iter *it, adg_trace *tr) {
  if (!c || !it) return ADG_E_ARG;
  if (!c->xs_set) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  if (it->nout < 1 || it->nin < 1 || it->nac < 1 || it->nupd < 1 || it->kern < 0 || it->kern > 2)
    return fail(c, ADG_E_ARG, "bad iteration control");
  if (mode != ADG_FORWARD && mode != ADG_ADJOINT) return fail(c, ADG_E_ARG, "bad mode");
  const int adjoint = mode == ADG_ADJOINT;
  const int nout = it->nout;
  int rc = ADG_OK;
  try {
    Geo q = c->geo();
    XS x = c->xs();
    ensure_trace(c, nout);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    if (c->first) {
      // f0 = 1, Ke = 1, fs0 = FSrc(f0) on the reference's first call
      std::vector<double> ones((size_t)c->N * c->G, 1.0), st;
      scatter(c, ones.data(), c->G, c->f0, st);
      c->Ke_host = 1.0;
    }
    set_scal_start(c, it);
    c->ev_next = 0;
    c->recs.clear();
    HIPCHK(hipEventRecord(e0, c->s));
    // matrix_setup(1): D-tilde and the matrix from the current D-hat
    hipLaunchKernelGGL(k_coupcoef, cells_grid(c, c->G), dim3(NT), 0, c->s, q, x, c->df);
    launch_matrix(c, 0);
    if (c->first) {
      hipLaunchKernelGGL(k_fsrc_init, dim3(c->nblk), dim3(NTF), 0, c->s, q, x, adjoint, c->f0, c->fs0);
      c->first = false;
    }
    hipLaunchKernelGGL(k_integrals, dim3(c->nblk), dim3(NTF), 0, c->s, q, x, c->fs0, c->part);
    hipLaunchKernelGGL(k_fin_integrals, dim3(1), dim3(NTF), 0, c->s, c->sc, c->part, c->nblk);
    HIPCHK(hipGetLastError());
    // replay: a captured graph of one outer (or direct launches when
    // profiling, whose events need the host-side sequence); the done flag
    // written by k_end turns the remaining kernels into no-ops, and the host
    // polls it every `check` outers
    const bool use_graph = !c->prof && c->graphs;
    if (use_graph) {
      const int key[4] = {adjoint, it->nin, it->kern, it->nupd};
      if (!c->gexec || memcmp(key, c->gkey, sizeof key) != 0) {
        if (c->gexec) {
          (void)hipGraphExecDestroy(c->gexec);
          c->gexec = nullptr;
        }
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(c, adjoint, it->nin, it->kern, 0, it->nupd);
        HIPCHK(hipStreamEndCapture(c->s, &graph));
        HIPCHK(hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        memcpy(c->gkey, key, sizeof key);
      }
    }
    const int check = c->check_every;
    int launched = 0;
    while (launched < nout) {
      const int n = std::min(check, nout - launched);
      for (int k = 0; k < n; ++k) {
        ++launched;
        if (use_graph) HIPCHK(hipGraphLaunch(c->gexec, c->s));
        else enqueue_iteration(c, adjoint, it->nin, it->kern, launched, it->nupd);
      }
      HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      if (c->hsc->done) break;
    }
    HIPCHK(hipEventRecord(e1, c->s));
    HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    c->timing[0] = ms;
    {
      double sum[3] = {0, 0, 0};
      int cnt[3] = {0, 0, 0};
      for (const auto &r : c->recs) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->evs[r.e0], c->evs[r.e1]));
        sum[r.type] += t;
        cnt[r.type] += 1;
      }
      for (int k = 0; k < 3; ++k) {
        c->prof_out[2 * k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
        c->prof_out[2 * k + 1] = cnt[k];
      }
      c->timing[1] = sum[1];
      c->timing[2] = sum[2];
      c->timing[3] = c->prof ? ms - sum[1] - sum[2] : 0.0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const Scal &h = *c->hsc;
    c->Ke_host = h.Ke;
    const int nrec = h.p;
    if (tr) {
      const int ncp = std::min(nrec, tr->cap);
      tr->n = ncp;
      if (ncp > 0) {
        if (tr->k) HIPCHK(hipMemcpy(tr->k, c->tk, ncp * sizeof(double), hipMemcpyDeviceToHost));
        if (tr->ser) HIPCHK(hipMemcpy(tr->ser, c->tser, ncp * sizeof(double), hipMemcpyDeviceToHost));
        if (tr->fer) HIPCHK(hipMemcpy(tr->fer, c->tfer, ncp * sizeof(double), hipMemcpyDeviceToHost));
        if (tr->ndmax) HIPCHK(hipMemcpy(tr->ndmax, c->tnd, ncp * sizeof(double), hipMemcpyDeviceToHost));
        if (tr->flags) HIPCHK(hipMemcpy(tr->flags, c->tflags, ncp * sizeof(int), hipMemcpyDeviceToHost));
        if (tr->loc) HIPCHK(hipMemcpy(tr->loc, c->tloc, 3 * ncp * sizeof(int), hipMemcpyDeviceToHost));
      }
    }
    if (h.status != 0) {
      rc = h.status;
      if (rc == ADG_E_LU_PIVOT) {
        unsigned long long key = h.lu_key;
        char buf[160];
        snprintf(buf, sizeof buf, "LU pivot below 1e-4 at node I = %d, J = %d, K = %d", (int)((key >> 20) & 1023),
                 (int)((key >> 10) & 1023), (int)(key & 1023));
        c->err = buf;
        if (((key >> 20) & 1023) == 0) c->err = "single-node line";
      } else if (rc == ADG_E_NODAL_UNSTABLE) {
        char buf[160];
        snprintf(buf, sizeof buf, "Max. change in nodal coupling coefficient = %.1f", h.ndmax);
        c->err = buf;
      }
    } else if (!h.done) {
      rc = ADG_E_MAX_OUTER;
      c->err = "maximum number of outer iterations reached";
    }
  } catch (const HipError &e) {
    rc = hip_fail(c, e);
  }
  return rc;
}

int adg_powdis(adg_ctx *c, double *p) {
  if (!c || !p) return ADG_E_ARG;
  try {
    Geo q = c->geo();
    hipLaunchKernelGGL(k_powdis, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->xs(), c->f0, c->pw, c->part);
    hipLaunchKernelGGL(k_scale, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->sc, c->pw, c->part);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (!(c->hsc->tpow > 0.0)) return fail(c, ADG_E_POWER_NONPOS, "total node power is zero or less");
    std::vector<double> st;
    gather(c, c->pw, 1, p, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_nodal_update(adg_ctx *c, int kern, int cmode, double *ndmax, int *loc3) {
  if (!c) return ADG_E_ARG;
  if (!c->xs_set) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  if (kern != ADG_KERN_PNM && kern != ADG_KERN_SANM) return fail(c, ADG_E_ARG, "nodal kernel must be PNM or SANM");
  try {
    set_scal_start(c, nullptr);
    Geo q = c->geo();
    hipLaunchKernelGGL(k_coupcoef, cells_grid(c, c->G), dim3(NT), 0, c->s, q, c->xs(), c->df);
    HIPCHK(hipGetLastError());
    nodal_update(c, kern, cmode, 0);
    HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (ndmax) *ndmax = c->hsc->ndmax;
    if (loc3) { loc3[0] = c->hsc->im; loc3[1] = c->hsc->jm; loc3[2] = c->hsc->km; }
    if (c->hsc->lu_key != ~0ull) return fail(c, ADG_E_LU_PIVOT, "LU pivot below threshold");
    if (c->hsc->ndmax > 1.e3) return fail(c, ADG_E_NODAL_UNSTABLE, "nodal update unstable");
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_bicg(adg_ctx *c, int g, int nin, const double *b, double *x) {
  if (!c || !b || !x || g < 0 || g >= c->G || nin < 1) return ADG_E_ARG;
  if (!c->xs_set) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  try {
    set_scal_start(c, nullptr);
    Geo q = c->geo();
    hipLaunchKernelGGL(k_coupcoef, cells_grid(c, c->G), dim3(NT), 0, c->s, q, c->xs(), c->df);
    launch_matrix(c, 0);
    std::vector<double> st;
    // right-hand side goes to c->b, the initial guess to the flux of group g
    scatter(c, b, 1, c->b, st);
    std::vector<double> full;
    full.assign((size_t)c->G * c->N, 0.0);
    gather(c, c->f0, c->G, full.data(), st);
    memcpy(full.data() + (size_t)g * c->N, x, c->N * sizeof(double));
    scatter(c, full.data(), c->G, c->f0, st);
    launch_bicg(c, g, nin);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    gather(c, c->f0, c->G, full.data(), st);
    memcpy(x, full.data() + (size_t)g * c->N, c->N * sizeof(double));
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_matvec(adg_ctx *c, int g, const double *x, double *y) {
  if (!c || !x || !y || g < 0 || g >= c->G) return ADG_E_ARG;
  if (!c->xs_set) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  try {
    Geo q = c->geo();
    hipLaunchKernelGGL(k_coupcoef, cells_grid(c, c->G), dim3(NT), 0, c->s, q, c->xs(), c->df);
    launch_matrix(c, 0);
    std::vector<double> st;
    scatter(c, x, 1, c->b, st);
    HIPCHK(hipMemsetAsync(c->tmp, 0, c->NC * sizeof(double), c->s));
    hipLaunchKernelGGL(k_matvec, dim3(c->nblk), dim3(NTF), 0, c->s, q, c->A64, g, c->b, c->tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    gather(c, c->tmp, 1, y, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_set_profiling(adg_ctx *c, int on) {
  if (!c) return ADG_E_ARG;
  c->prof = on != 0;
  return ADG_OK;
}

int adg_read_profile(adg_ctx *c, double *out6) {
  if (!c || !out6) return ADG_E_ARG;
  for (int k = 0; k < 6; ++k) out6[k] = c->prof_out[k];
  return ADG_OK;
}

int adg_last_timing(adg_ctx *c, double *ms4) {
  if (!c || !ms4) return ADG_E_ARG;
  for (int k = 0; k < 4; ++k) ms4[k] = c->timing[k];
  return ADG_OK;
}

}  // extern "C"
