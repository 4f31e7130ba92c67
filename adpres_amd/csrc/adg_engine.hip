// Host side of the gfx950 static k-eigenvalue engine: context, data
// movement between the reference's node list and the padded device boxes,
// the launch tables, the outer-iteration driver and the C-ABI of
// include/adpres_gpu.h.
//
// A context solves a batch of B >= 1 cores with the same group count.  Each
// core keeps its own padded box; the boxes are concatenated in every
// node-wise device array, and every launch covers the whole batch through a
// block table (flat launches over cells, per-node launches, per-face
// launches).  One core is the batch of one, so the drop-in path and the
// batched path are the same kernels.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/adpres_gpu.h"
#include "adg_adjoint.h"
#include "adg_bicgp.h"
#include "adg_cmfd.h"
#include "adg_common.h"
#include "adg_nodal.h"
#include "adg_nodal_vjp.h"
#include "adg_output.h"
#include "adg_xs.h"

using namespace adg;

// Scratch of the allocation-free output-map entry point
// (adg_power_maps_noalloc): the map kernels' integer tables (footprint rows,
// assembly bounds, the argmax) in the first SCR_INTS ints, their
// intermediate sums (axially averaged node powers, assembly and plane powers)
// in the next SCR_DBLS doubles.  One copy per GPU, held until the process
// ends.  In host code `scratch` only names the symbol; its device address
// on each GPU comes from hipGetSymbolAddress (scratch_init) and only that
// address is passed to kernels or copies.
constexpr long SCR_INTS = 8192;     // 32 KiB
constexpr long SCR_DBLS = 131072;   // 1 MiB
static __device__ __align__(256) unsigned char scratch[SCR_INTS * sizeof(int) + SCR_DBLS * sizeof(double)];

namespace {

constexpr int SCR_MAXDEV = 64;
std::mutex scr_lock;
void *scr_ptr[SCR_MAXDEV] = {};          // device address of `scratch` per device
hipError_t scr_err[SCR_MAXDEV] = {};     // result of the lookup (hipSuccess once resolved)
bool scr_done[SCR_MAXDEV] = {};

// resolve the device address of `scratch` on device dev (current), once
void scratch_init(int dev) {
  if (dev < 0 || dev >= SCR_MAXDEV) return;
  std::lock_guard<std::mutex> lk(scr_lock);
  if (scr_done[dev]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  scr_err[dev] = e;
  scr_ptr[dev] = e == hipSuccess ? p : nullptr;
  scr_done[dev] = true;
}

struct HipError {
  hipError_t e;
  const char *what;
};

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t _e = (x);                       \
    if (_e != hipSuccess) throw HipError{_e, #x}; \
  } while (0)

// Every fill and copy of a context goes through its own stream (never the
// null stream) and is waited for.  The context stream is non-blocking, so a
// null-stream fill (hipMemset returns before the fill lands) could otherwise
// land after a later upload on the context stream; contexts driven from
// several host threads also must not synchronise with each other or with
// another thread's graph capture.
void scopy(hipStream_t s, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, s));
  HIPCHK(hipStreamSynchronize(s));
}

template <typename T>
T *dalloc(hipStream_t s, size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T), s));
  HIPCHK(hipStreamSynchronize(s));
  return (T *)p;
}

// host-side description of one core of the batch
struct CoreHost {
  int N = 0, nmat = 0, nxx = 0, nyy = 0, nzz = 0;
  int NXP = 0, NYP = 0, NZP = 0;
  long sy = 0, sz = 0, box = 0;    // strides and cells of its padded box
  long base = 0;                   // first cell of its box in the batch arrays
  int bcface[6] = {0, 0, 0, 0, 0, 0};
  std::vector<long> cell;          // node n -> cell of the batch arrays
  long coff = 0;                   // its node list's offset in the device cell list
  std::vector<double> hx, hy, hz;
  long hoff[3] = {0, 0, 0};        // offsets of hx, hy, hz in the device arrays
  bool xs_set = false, first = true;
  double Ke_host = 1.0;
  int pfirst = 0, pn = 0, ffirst = 0, fnb = 0;
  int bfirst = 0, bnb = 0;         // its blocks of the boundary-face list launch
  int status = 0, niter = 0;
  // feedback tables of adg_set_xs_tables (device)
  double *xtab = nullptr;          // [6 kinds][4 G M + G G M], then zdum(nzz), zdel(nzz)
  int *xnode = nullptr;            // (2N) material - 1, then bank | (level << 16) (0: no rod)
  int xon = 0, xnb = 0, xnmat = 0;
  double xref[7] = {0, 0, 0, 0, 0, 0, 0};   // rbcon rftem rmtem rcden coreh pos0 ssize
  // branch tables of adg_set_xtab_tables (xkind 2; 1: feedback tables)
  int xkind = 0;
  BrMat *xmeta = nullptr;
  long xpts = 0, xrec = 0, xrrec = 0;   // doubles of points, records, rodded records in xtab
};

}  // namespace

struct adg_ctx {
  int B = 0, G = 0, precision = 0;
  int dev = 0;                     // the device current at adg_create (all of the context's memory is there)
  long NC = 0;                     // cells of all boxes = group stride of the device arrays
  std::vector<CoreHost> core;
  hipStream_t s = nullptr;
  // device
  Geo *geo = nullptr;
  Blk *tflat = nullptr, *tcell = nullptr, *tface = nullptr;
  int nflat = 0, ncell = 0, nface = 0;
  // boundary faces (one node active) of the cooperative nodal form: list of
  // (direction << 40 | face index) per core, NT entries per block
  Blk *tbface = nullptr;
  long *bfl = nullptr;
  int nbface = 0;
  int flat_cap = 0;                // allocated entries of tflat (and 12 x of part)
  long flat_param = 0;             // launch shape of the flat kernels (see build_flat)
  double *d_hx = nullptr, *d_hy = nullptr, *d_hz = nullptr;
  unsigned char *flag = nullptr;
  double *D = nullptr, *sigr = nullptr, *nuf = nullptr, *sigf = nullptr, *chi = nullptr, *sigs = nullptr;
  double *dc = nullptr, *exsrc = nullptr, *vdel = nullptr;
  double *wtr = nullptr;   // transient fission weight per cell (adg_set_transient)
  long *d_cell = nullptr;  // node -> cell of every core (offset coff per core)
  double *dstage = nullptr;   // node-ordered staging of host arrays
  int *xerr = nullptr;        // adg_update_xs's stop flag
  hipEvent_t xev[2] = {nullptr, nullptr};   // bracket adg_update_xs's kernel
  size_t stage_cap = 0;
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
  double *Bd = nullptr, *cf4 = nullptr;   // cooperative nodal passes: get_B node terms, A F G H per direction
  // per core, direction (forward, adjoint) and group: the in-scatter source
  // groups with a nonzero sigma_s anywhere in the core (bits 0..G-1) and a
  // nonzero external source (SMASK_EX); the source kernel leaves out the
  // other terms, which add exactly +0.  Device copy and host copy.
  unsigned *smask = nullptr;
  std::vector<unsigned> hmask;
  // reductions / scalars
  double *part = nullptr;          // three rotating partial buffers of 4 x nflat
  Scal *sc = nullptr;
  Scal *hsc = nullptr;             // pinned host mirror (B)
  bool graphs = true;              // replay the outer iteration as a captured hipGraph
  int check_every = 8;             // outers between host polls of the done flags
  bool nod_coop = true;            // cooperative pass 2 for G >= 5 (ADG_NOD_COOP=0: per-thread)
  bool nod_lu32 = false;           // FP32+refine: interface-pass solves in FP32 + FP64 correction (ADG_NOD_LU32=1)
  bool src_full = true;            // source kernel on one chunk per block (ADG_SRC_FULL=0: flat shape)
  bool xs_lds = true;              // cross-section tables staged in LDS when they fit (ADG_XS_LDS=0: caches)
  double timing[4] = {0, 0, 0, 0};
  std::string err;
  // persistent BiCGSTAB (single core): the brick plan, its kernel instance
  // (cells per thread, 0: no plan), whether group solves use it, and the
  // zeroed synchronisation area
  Brick pb{};
  int pkc = 0;
  bool persist = false;
  std::vector<Brick> pcand;        // checked brick plans, cheapest by the cost model first (the autotuner times them)
  std::vector<int> pcand_kc;
  double *psync = nullptr;
  long pspin = PSPIN;              // polls before a wait of the persistent kernel gives up (ADG_PSPIN)
  int adj_inner = 16;              // generalized-adjoint outers per full-model adjoint step (ADG_ADJ_INNER; 0: to convergence)
  // profiling: event pairs recorded on the context stream during adg_outer
  struct ProfRec {
    int type, e0, e1;   // type 0 KA kernel, 1 group solve, 2 nodal update
  };
  bool prof = false;
  std::vector<hipEvent_t> evs;
  int ev_next = 0;
  // without profiling: event pairs around the replays of nodal-update outers
  // only, from which the nodal share of adg_last_timing is estimated
  std::vector<hipEvent_t> nevs;
  std::vector<ProfRec> recs;
  double prof_out[6] = {0, 0, 0, 0, 0, 0};
  // trace buffers (device, cap outers per core) and the captured outer iteration
  double *tk = nullptr, *tser = nullptr, *tfer = nullptr, *tnd = nullptr;
  int *tflags = nullptr, *tloc = nullptr;
  int tcap = 0;
  hipGraphExec_t gexec = nullptr;         // one outer, nodal-update kernels included
  hipGraphExec_t gexec_plain = nullptr;   // one outer without them (outers the schedule leaves out)
  int gkey[6] = {-1, -1, -1, -1, -1, -1};   // source mode, nin, kern, nupd, variant, transpose of the captured outer
  // differentiation (single core): saved forward state, adjoint, generalized
  // adjoint and gradients
  double *psi = nullptr, *phis = nullptr, *fs0s = nullptr, *exs = nullptr, *gam = nullptr, *dJdP = nullptr;
  double *grad = nullptr;
  bool transpose = false, have_psi = false;
  int variant = 0;                 // 0 eigenvalue outer, 1 generalized-adjoint outer (k fixed, projection)
  double psifphi = 0.0, kfwd = 0.0;
  // transpose of the nodal update and the full-model adjoint: one block,
  // the scratch of one transpose first (zeroed before each), then the
  // adjoint iterate, its successor, the CMFD D-hat gradient and the
  // response's adjoint source
  double *vjp = nullptr;
  size_t vjp_scratch = 0;
  VjpArgs va{};
  double *vw = nullptr, *vwn = nullptr, *vgdn = nullptr, *vsst = nullptr;

  long s_c0 = 0, s_end = 0, s_chunk = 0;   // single core: flat blocks computed in the kernels
  LV lv(const Blk *t) const {
    LV v;
    memset(&v, 0, sizeof v);
    v.geo = geo; v.sc = sc; v.blk = t; v.NC = NC; v.G = G;
    if (t == tflat && B == 1) {
      v.blk = nullptr;
      v.c0 = s_c0; v.end = s_end; v.chunk = s_chunk; v.nb = nflat;
      v.sy = core[0].sy; v.sz = core[0].sz;
    }
    return v;
  }
  XS xs() const {
    XS x;
    x.D = D; x.sigr = sigr; x.nuf = nuf; x.sigf = sigf; x.chi = chi; x.sigs = sigs;
    x.dc = dc; x.exsrc = exsrc; x.vdel = vdel; x.wtr = wtr;
    return x;
  }
  double *P(int k) const { return part + (size_t)4 * nflat * k; }
  bool all_xs() const {
    for (const auto &cc : core)
      if (!cc.xs_set) return false;
    return true;
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

// scatter a Fortran (N, K) array of core b into K group/face slots of its box
// Host arrays in the reference's node order (K components of N nodes) move
// in one copy to a node-ordered device staging buffer, and a kernel places
// them into the padded boxes (and back): the cells outside the footprint
// keep the zeros they were allocated with.
__global__ void k_scatter(const double *src, const long *cell, long N, long total, long NC, double *dst) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long k = i / N, n = i - k * N;
    dst[k * NC + cell[n]] = src[i];
  }
}

__global__ void k_gather(const double *src, const long *cell, long N, long total, long NC, double *dst) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long k = i / N, n = i - k * N;
    dst[i] = src[k * NC + cell[n]];
  }
}

void ensure_stage(adg_ctx *c, size_t n) {
  if (n <= c->stage_cap) return;
  if (c->dstage) HIPCHK(hipFree(c->dstage));
  c->dstage = nullptr;
  c->stage_cap = 0;
  c->dstage = dalloc<double>(c->s, n);
  c->stage_cap = n;
}

void scatter(adg_ctx *c, int b, const double *src, int K, double *dst_dev, std::vector<double> &) {
  const CoreHost &h = c->core[b];
  const long total = (long)K * h.N;
  if (total == 0) return;
  ensure_stage(c, (size_t)total);
  HIPCHK(hipMemcpyAsync(c->dstage, src, (size_t)total * sizeof(double), hipMemcpyHostToDevice, c->s));
  const int nb = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(256), 0, c->s, c->dstage, c->d_cell + h.coff, (long)h.N, total,
                     c->NC, dst_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->s));
}

void gather(adg_ctx *c, int b, const double *src_dev, int K, double *dst, std::vector<double> &) {
  const CoreHost &h = c->core[b];
  const long total = (long)K * h.N;
  if (total == 0) return;
  ensure_stage(c, (size_t)total);
  const int nb = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_gather, dim3(nb), dim3(256), 0, c->s, src_dev, c->d_cell + h.coff, (long)h.N, total, c->NC,
                     c->dstage);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(dst, c->dstage, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
}

// ---------------------------------------------------------------------------
// launch tables
// ---------------------------------------------------------------------------

// Flat launches.  One core (param = largest grid): NTF-multiple chunks, and
// the hardware's block b -> XCD b % 8 dispatch turned into contiguous slabs
// per XCD (logical block (b % 8) * (grid / 8) + b / 8), so a stencil's
// z-neighbour planes are read through the XCD's own L2.  A batch (param =
// cells per block, a multiple of NTF): cores assigned whole to XCDs (largest
// first, to the least loaded XCD), each XCD's block list interleaved at
// stride 8; lists are padded with empty blocks whose partial slots no core
// reads.
void build_flat(adg_ctx *c, std::vector<Blk> &t, long param) {
  t.clear();
  if (c->B == 1) {
    CoreHost &h = c->core[0];
    const long M = (long)h.nzz * h.sz;
    int nb = (int)std::min<long>((M + NTF - 1) / NTF, param);
    nb = (nb + 7) / 8 * 8;
    long chunk = (M + nb - 1) / nb;
    chunk = (chunk + NTF - 1) / NTF * NTF;
    const long c0 = h.base + h.sz, end = c0 + M;
    t.resize(nb);
    for (int b = 0; b < nb; ++b) {   // the same blocks blk_view derives from the launch arguments
      const int lb = (b & 7) * (nb >> 3) + (b >> 3);
      long lo = c0 + (long)lb * chunk, hi = lo + chunk;
      lo = std::min(lo, end);
      hi = std::min(hi, end);
      t[b] = Blk{0, b, 0, nb, lo, hi, h.sy, h.sz};
    }
    h.pfirst = 0;
    h.pn = nb;
    c->s_c0 = c0;
    c->s_end = end;
    c->s_chunk = chunk;
    return;
  }
  const long chunk = param;
  std::vector<int> nbk(c->B);
  for (int b = 0; b < c->B; ++b) {
    const CoreHost &h = c->core[b];
    nbk[b] = (int)(((long)h.nzz * h.sz + chunk - 1) / chunk);
  }
  int slot = 0;
  for (int b = 0; b < c->B; ++b) {
    c->core[b].pfirst = slot;
    c->core[b].pn = nbk[b];
    slot += nbk[b];
  }
  std::vector<int> order(c->B);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nbk[a] > nbk[b]; });
  std::vector<std::vector<Blk>> xl(8);
  std::vector<long> load(8, 0);
  for (int b : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    const CoreHost &h = c->core[b];
    const long c0 = h.base + h.sz, end = c0 + (long)h.nzz * h.sz;
    for (int j = 0; j < nbk[b]; ++j) {
      const long lo = c0 + (long)j * chunk;
      xl[x].push_back(Blk{b, h.pfirst + j, h.pfirst, h.pn, lo, std::min(lo + chunk, end), h.sy, h.sz});
    }
    load[x] += nbk[b];
  }
  size_t L = 0;
  for (auto &v : xl) L = std::max(L, v.size());
  int pad = slot;
  t.resize(8 * L);
  for (size_t j = 0; j < L; ++j)
    for (int x = 0; x < 8; ++x)
      t[8 * j + x] = j < xl[x].size() ? xl[x][j] : Blk{0, pad++, c->core[0].pfirst, c->core[0].pn, 0, 0, 1, 1};
}

// the captured outers (re-captured at the next outer call)
void drop_graphs(adg_ctx *c) {
  for (hipGraphExec_t *g : {&c->gexec, &c->gexec_plain})
    if (*g) {
      (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
}

void upload_geo(adg_ctx *c) {
  std::vector<Geo> gq(c->B);
  for (int b = 0; b < c->B; ++b) {
    const CoreHost &h = c->core[b];
    Geo &q = gq[b];
    memset(&q, 0, sizeof(Geo));
    q.nxx = h.nxx; q.nyy = h.nyy; q.nzz = h.nzz; q.G = c->G;
    q.NXP = h.NXP; q.NYP = h.NYP;
    q.sy = h.sy; q.sz = h.sz; q.NC = c->NC;
    q.base = h.base;
    q.c0 = h.base + h.sz; q.M = (long)h.nzz * h.sz;
    for (int s6 = 0; s6 < 6; ++s6) q.bc[s6] = h.bcface[s6];
    q.pfirst = h.pfirst; q.pn = h.pn; q.ffirst = h.ffirst; q.fnb = h.fnb;
    q.bfirst = h.bfirst; q.bnb = h.bnb;
    q.hx = c->d_hx + h.hoff[0]; q.hy = c->d_hy + h.hoff[1]; q.hz = c->d_hz + h.hoff[2];
    q.flag = c->flag;
  }
  if (!c->geo) c->geo = dalloc<Geo>(c->s, c->B);
  scopy(c->s, c->geo, gq.data(), gq.size() * sizeof(Geo), hipMemcpyHostToDevice);
}

// default launch shape of the flat kernels: one core up to MAXBLK blocks
// (the FP32 form, streaming half the bytes, up to MAXBLK / 2), a batch one
// NTF-cell chunk per block
long default_flat_param(const adg_ctx *c) {
  if (c->B > 1) return NTF;
  return c->precision == ADG_FP32_REFINE ? MAXBLK / 2 : MAXBLK;
}

// (re)build the flat launch table, its partial buffers and the per-core slot
// ranges; a captured outer holds the old grid, so it is dropped
void install_flat(adg_ctx *c, long param) {
  std::vector<Blk> tf;
  build_flat(c, tf, param);
  if ((int)tf.size() > c->flat_cap) {
    if (c->tflat) HIPCHK(hipFree(c->tflat));
    if (c->part) HIPCHK(hipFree(c->part));
    c->tflat = nullptr;
    c->part = nullptr;
    c->flat_cap = (int)tf.size();
    c->tflat = dalloc<Blk>(c->s, tf.size());
    c->part = dalloc<double>(c->s, (size_t)12 * tf.size());   // three rotating buffers of 4 x grid
  }
  c->nflat = (int)tf.size();
  scopy(c->s, c->tflat, tf.data(), tf.size() * sizeof(Blk), hipMemcpyHostToDevice);
  c->flat_param = param;
  upload_geo(c);
  drop_graphs(c);
}

void plan_persist(adg_ctx *c);

// the cooperative nodal kernels (one lane per row of the node and face
// systems) serve the larger group counts
bool nod_coop(const adg_ctx *c) { return c->G >= 5 && c->nod_coop; }

// fls: each core's cell flags (padded box, F_ACTIVE and neighbour bits)
void build_tables(adg_ctx *c, const std::vector<std::vector<unsigned char>> &fls) {
  std::vector<Blk> tc, tfc, tbf;
  std::vector<long> bl;
  for (int b = 0; b < c->B; ++b) {
    CoreHost &h = c->core[b];
    const long nn = (long)h.nxx * h.nyy * h.nzz;
    for (long t0 = 0; t0 < nn; t0 += NT) tc.push_back(Blk{b, 0, 0, 0, t0, nn, h.sy, h.sz});
    const long maxf = std::max({(long)(h.nxx + 1) * h.nyy * h.nzz, (long)(h.nyy + 1) * h.nxx * h.nzz,
                                (long)(h.nzz + 1) * h.nxx * h.nyy});
    h.ffirst = (int)tfc.size();
    for (long t0 = 0; t0 < maxf; t0 += NT) tfc.push_back(Blk{b, 0, 0, 0, t0, maxf, h.sy, h.sz});
    h.fnb = (int)tfc.size() - h.ffirst;
    h.bfirst = (int)tbf.size();
    h.bnb = 0;
    if (!nod_coop(c)) continue;
    // boundary faces in k_nod_pass2's enumeration: direction u, face index
    // t = fi + (nu + 1) (o1 + e1 o2), one of its two nodes active
    const std::vector<unsigned char> &fl = fls[b];
    const long l0 = (long)bl.size();
    for (int u = 0; u < 3; ++u) {
      const int nu = u == 0 ? h.nxx : (u == 1 ? h.nyy : h.nzz);
      const int e1 = u == 0 ? h.nyy : h.nxx, e2 = u == 2 ? h.nyy : h.nzz;
      const long st = u == 0 ? 1 : (u == 1 ? h.sy : h.sz);
      for (long t = 0; t < (long)(nu + 1) * e1 * e2; ++t) {
        const int fi = (int)(t % (nu + 1));
        const long r = t / (nu + 1);
        const int o1 = (int)(r % e1), o2 = (int)(r / e1);
        const int i = u == 0 ? fi : o1, j = u == 0 ? o1 : (u == 1 ? fi : o2), k = u == 2 ? fi : o2;
        const long cp = ((long)(k + 1) * h.NYP + (j + 1)) * h.NXP + (i + 1);
        const bool hp = (fl[cp] & F_ACTIVE) != 0, hn = (fl[cp - st] & F_ACTIVE) != 0;
        if (hp != hn) bl.push_back(((long)u << 40) | t);
      }
    }
    const long l1 = (long)bl.size();
    for (long t0 = l0; t0 < l1; t0 += NT) tbf.push_back(Blk{b, 0, 0, 0, t0, l1, h.sy, h.sz});
    h.bnb = (int)tbf.size() - h.bfirst;
  }
  c->ncell = (int)tc.size();
  c->nface = (int)tfc.size();
  c->nbface = (int)tbf.size();
  c->tcell = dalloc<Blk>(c->s, tc.size());
  c->tface = dalloc<Blk>(c->s, tfc.size());
  scopy(c->s, c->tcell, tc.data(), tc.size() * sizeof(Blk), hipMemcpyHostToDevice);
  scopy(c->s, c->tface, tfc.data(), tfc.size() * sizeof(Blk), hipMemcpyHostToDevice);
  if (!tbf.empty()) {
    c->tbface = dalloc<Blk>(c->s, tbf.size());
    c->bfl = dalloc<long>(c->s, bl.size());
    scopy(c->s, c->tbface, tbf.data(), tbf.size() * sizeof(Blk), hipMemcpyHostToDevice);
    scopy(c->s, c->bfl, bl.data(), bl.size() * sizeof(long), hipMemcpyHostToDevice);
  }
  install_flat(c, default_flat_param(c));
  plan_persist(c);
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// persistent BiCGSTAB: plan, check, launch
// ---------------------------------------------------------------------------

// one persistent solve at a time in the process: the blocks of two such
// launches on one GPU could hold each other's CUs while waiting
std::mutex persist_mu;

template <typename TV>
const void *persist_fn(int kc) {
  switch (kc) {
    case 1: return (const void *)k_bicg_persist<TV, 1>;
    case 2: return (const void *)k_bicg_persist<TV, 2>;
    case 4: return (const void *)k_bicg_persist<TV, 4>;
    case 8: return (const void *)k_bicg_persist<TV, 8>;
  }
  return nullptr;
}

// The brick plan of a single core (plan_bricks), checked on the host
// (verify_plan), kept only when its blocks can all be resident on the device.
void plan_persist(adg_ctx *c) {
  c->pkc = 0;
  c->persist = false;
  c->pcand.clear();
  c->pcand_kc.clear();
  if (c->B != 1) return;
  if (const char *ev = getenv("ADG_PERSIST"))
    if (atoi(ev) == 0) return;
  const CoreHost &h = c->core[0];
  const bool f32 = c->precision == ADG_FP32_REFINE;
  int ncu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
  std::vector<Brick> plans;
  std::vector<int> kcs;
  plan_brick_list(h.nxx, h.nyy, h.nzz, h.base, h.sy, h.sz, c->NC, f32, ncu, plans, kcs, 4);
  // keep the plans whose kernel instance exists, whose blocks are resident
  // together (one per CU at most) and whose every index checks out on the host
  for (size_t q = 0; q < plans.size(); ++q) {
    const Brick &P = plans[q];
    const int kc = kcs[q];
    const void *fn = f32 ? persist_fn<float>(kc) : persist_fn<double>(kc);
    if (!fn) continue;
    int occ = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, PNT, persist_lds(P, f32)));
    if ((long)occ * ncu < P.nb) continue;
    const std::string e = verify_plan(P, kc, h.box);
    if (!e.empty()) {
      fprintf(stderr, "adg: persistent BiCGSTAB plan %dx%dx%d rejected (%s)\n", P.BX, P.BY, P.BZ, e.c_str());
      continue;
    }
    c->pcand.push_back(P);
    c->pcand_kc.push_back(kc);
  }
  if (c->pcand.empty()) return;
  if (!c->psync) c->psync = dalloc<double>(c->s, PSYNC_DOUBLES);
  c->pb = c->pcand[0];
  c->pkc = c->pcand_kc[0];
  c->persist = true;
}

// clear the synchronisation area after a timed-out persistent solve (flags
// and epoch base back to zero)
void reset_psync(adg_ctx *c) {
  if (!c->psync) return;
  HIPCHK(hipMemsetAsync(c->psync, 0, PSYNC_DOUBLES * sizeof(double), c->s));
  HIPCHK(hipStreamSynchronize(c->s));
}

template <typename TV, int KC>
void launch_persist_kc(adg_ctx *c, const PArgs<TV> &a, size_t lds, int ek0, int ek1) {
  if (ek0 >= 0 && ek1 >= 0)
    hipExtLaunchKernelGGL((k_bicg_persist<TV, KC>), dim3(a.P.nb), dim3(PNT), lds, c->s, c->evs[ek0], c->evs[ek1], 0,
                          a);
  else
    hipLaunchKernelGGL((k_bicg_persist<TV, KC>), dim3(a.P.nb), dim3(PNT), lds, c->s, a);
}

// bicg of group g as one persistent launch (the flat form's buffers: S in
// R[0], T in t, p and v by parity in P and V)
template <typename TV>
void launch_bicg_persist(adg_ctx *c, int g, int nin, const TV *A, TV *const *R, TV *const *P, TV *const *V, TV *t) {
  PArgs<TV> a;
  a.P = c->pb;
  a.A64g = c->A64 + (size_t)g * 7 * c->NC;
  a.ATg = A;
  a.b = c->b;
  a.x64 = c->f0 + (size_t)g * c->NC;
  a.S = R[0]; a.T = t; a.P0 = P[0]; a.P1 = P[1]; a.V0 = V[0]; a.V1 = V[1];
  a.sync = c->psync;
  a.sc = c->sc;
  a.nin = nin;
  a.status_timeout = ADG_E_DEVICE_SYNC;
  a.spin = c->pspin;
  const size_t lds = persist_lds(c->pb, sizeof(TV) == 4);
  const int ek0 = prof_slot(c), ek1 = prof_slot(c);
  switch (c->pkc) {
    case 1: launch_persist_kc<TV, 1>(c, a, lds, ek0, ek1); break;
    case 2: launch_persist_kc<TV, 2>(c, a, lds, ek0, ek1); break;
    case 4: launch_persist_kc<TV, 4>(c, a, lds, ek0, ek1); break;
    case 8: launch_persist_kc<TV, 8>(c, a, lds, ek0, ek1); break;
  }
  HIPCHK(hipGetLastError());
  prof_pair(c, 0, ek0, ek1);
}

void launch_coupcoef(adg_ctx *c) {
  hipLaunchKernelGGL(k_coupcoef, dim3(c->ncell, c->G), dim3(NT), 0, c->s, c->lv(c->tcell), c->xs(), c->df);
}

void launch_matrix(adg_ctx *c, int sched) {
  const int tr = c->transpose ? 1 : 0;
  if (c->precision == ADG_FP32_REFINE)
    hipLaunchKernelGGL(k_matrix<float>, dim3(c->ncell, c->G), dim3(NT), 0, c->s, c->lv(c->tcell), c->xs(), c->df,
                       c->dn, c->A64, c->A32, sched, tr);
  else
    hipLaunchKernelGGL(k_matrix<double>, dim3(c->ncell, c->G), dim3(NT), 0, c->s, c->lv(c->tcell), c->xs(), c->df,
                       c->dn, c->A64, (double *)nullptr, sched, tr);
  HIPCHK(hipGetLastError());
}

// bicg (src/mod_cmfd.f90:1203-1243) of group g: b in c->b, x = flux of g.
// Kernels: setup, then KA(1) KB(1) KA(2) KB(2) ... KB(nin), KT (two stencil
// kernels per iteration, see adg_cmfd.h).  r, p, v ping-pong: KA(it) reads
// set (it-1)&1 and writes set it&1.  Partials: setup -> P0, KA(1) P0 -> P1,
// KB P1 -> P2, KA(it>1) P2 -> P1, KT reads P2.
template <typename TV, int TAB>
void launch_bicg_t(adg_ctx *c, int g, int nin, const TV *A, TV *const *R, TV *const *P, TV *const *V, TV *t,
                   TV *rs, TV *d) {
  const LV lv = c->lv(c->tflat);
  const dim3 grid(c->nflat), blk(NTF);
  const long NC = c->NC;
  double *x64 = c->f0 + (size_t)g * NC;
  double *P0 = c->P(0), *P1 = c->P(1), *P2 = c->P(2);
  const double *A64g = c->A64 + (size_t)g * 7 * NC;
  // FP64 form: the iterate is the flux itself; FP32 form: a correction d
  TV *x = d ? d : (TV *)x64;
  const int sample = nin > 1 ? 2 : 0;   // steady-state KA launch that is timed (none if nin = 1)
  int ek0 = -1, ek1 = -1;
  hipLaunchKernelGGL((k_bicg_setup<TV, TAB>), grid, blk, 0, c->s, lv, A64g, c->b, x64, R[0], rs, d, P0);
  for (int it = 1; it <= nin; ++it) {
    const int rd = (it - 1) & 1, wr = it & 1;
    if (it == sample) {
      ek0 = prof_slot(c);
      ek1 = prof_slot(c);
    }
    if (it == sample && ek0 >= 0 && ek1 >= 0)
      hipExtLaunchKernelGGL((k_bicgA<TV, false, TAB>), grid, blk, 0, c->s, c->evs[ek0], c->evs[ek1], 0, lv, A, it,
                            R[rd], P[rd], V[rd], t, rs, x, R[wr], P[wr], V[wr], P2, P1);
    else if (it == 1)
      hipLaunchKernelGGL((k_bicgA<TV, true, TAB>), grid, blk, 0, c->s, lv, A, it, R[rd], P[rd], V[rd], t, rs, x,
                         R[wr], P[wr], V[wr], P0, P1);
    else
      hipLaunchKernelGGL((k_bicgA<TV, false, TAB>), grid, blk, 0, c->s, lv, A, it, R[rd], P[rd], V[rd], t, rs, x,
                         R[wr], P[wr], V[wr], P2, P1);
    hipLaunchKernelGGL((k_bicgB<TV, TAB>), grid, blk, 0, c->s, lv, A, it, R[wr], V[wr], t, rs, P1, P2);
  }
  const int wr = nin & 1;
  hipLaunchKernelGGL((k_bicgT<TV, TAB>), grid, blk, 0, c->s, lv, nin, R[wr], P[wr], V[wr], x,
                     d ? x64 : (double *)nullptr, P2);
  HIPCHK(hipGetLastError());
  prof_pair(c, 0, ek0, ek1);
}

void launch_bicg(adg_ctx *c, int g, int nin) {
  const long NC = c->NC;
  float *A32 = c->A32 + (size_t)g * 7 * NC;
  double *A64 = c->A64 + (size_t)g * 7 * NC;
  if (c->persist) {
    if (c->precision == ADG_FP32_REFINE) launch_bicg_persist<float>(c, g, nin, A32, c->rr32, c->pp32, c->vv32, c->t32);
    else launch_bicg_persist<double>(c, g, nin, A64, c->rr64, c->pp64, c->vv64, c->t64);
    return;
  }
  if (c->precision == ADG_FP32_REFINE) {
    if (c->B > 1) launch_bicg_t<float, 1>(c, g, nin, A32, c->rr32, c->pp32, c->vv32, c->t32, c->rs32, c->d32);
    else launch_bicg_t<float, 0>(c, g, nin, A32, c->rr32, c->pp32, c->vv32, c->t32, c->rs32, c->d32);
  } else {
    if (c->B > 1)
      launch_bicg_t<double, 1>(c, g, nin, A64, c->rr64, c->pp64, c->vv64, c->t64, c->rs64, (double *)nullptr);
    else
      launch_bicg_t<double, 0>(c, g, nin, A64, c->rr64, c->pp64, c->vv64, c->t64, c->rs64, (double *)nullptr);
  }
}

// one group solve: TSrc + bicg (src/mod_cmfd.f90:472-478)
void group_solve(adg_ctx *c, int g, int smode, int first_group, int nin) {
  const int eg0 = prof_event(c);
  // the source kernel has no reduction, so a single core's launch may use
  // one NTF-cell chunk per block (all cells in flight at once) instead of
  // the flat shape that the reductions of the BiCGSTAB kernels favour
  LV lvs = c->lv(c->tflat);
  int nbs = c->nflat;
  if (c->B == 1 && c->src_full) {
    const long cells = c->s_end - c->s_c0;
    nbs = (int)((cells + NTF - 1) / NTF);
    nbs = (nbs + 7) / 8 * 8;
    lvs.chunk = NTF;
    lvs.nb = nbs;
  }
  switch (c->G) {
#define SRC_CASE(GG)                                                                                          \
  case GG:                                                                                                    \
    if (c->B > 1)                                                                                             \
      hipLaunchKernelGGL((k_src<1, GG>), dim3(c->nflat), dim3(NTF), 0, c->s, c->lv(c->tflat), c->xs(), g, smode, \
                         first_group, c->f0, c->f0c, c->fs0, c->fs0c, c->b, (const unsigned *)c->smask);        \
    else                                                                                                      \
      hipLaunchKernelGGL((k_src<0, GG>), dim3(nbs), dim3(NTF), 0, c->s, lvs, c->xs(), g, smode,                \
                         first_group, c->f0, c->f0c, c->fs0, c->fs0c, c->b, (const unsigned *)c->smask);        \
    break;
    SRC_CASE(1) SRC_CASE(2) SRC_CASE(3) SRC_CASE(4) SRC_CASE(5) SRC_CASE(6) SRC_CASE(7) SRC_CASE(8)
#undef SRC_CASE
  }
  launch_bicg(c, g, nin);
  prof_pair(c, 1, eg0, prof_event(c));
}

// nodal::nodal_update(_pnm) + matrix rebuild (src/mod_cmfd.f90:339-383)
void nodal_update(adg_ctx *c, int kern, int cmode, int sched) {
  NodArgs a;
  a.xs = c->xs();
  a.f0 = c->f0; a.df = c->df; a.dn = c->dn;
  a.S = c->S; a.a2 = c->a2; a.a4 = c->a4; a.L1 = c->L1;
  a.kern = kern; a.cmode = cmode; a.sched = sched;
  a.part = c->npart;
  a.coop = 0;
  a.Bd = nod_coop(c) ? c->Bd : nullptr;
  a.cf4 = c->cf4;
  NodArgs an = a, af = a, ab = a;
  an.lv = c->lv(c->tcell);
  af.lv = c->lv(c->tface);
  ab.lv = c->lv(c->tbface);
  // larger group counts: the cooperative passes (one lane per row of the
  // G x G node systems and the 2G x 2G interface systems) and the boundary
  // faces from their list; ADG_NOD_COOP=0 keeps every system per thread.
  // ADG_NOD_LU32=1 has the FP32+refine form solve the interface systems in
  // FP32 with one FP64 defect correction (slower than the FP64 solves since
  // their broadcasts moved to DPP, so off by default)
  const bool coop = nod_coop(c);
  const bool lu32 = coop && c->precision == ADG_FP32_REFINE && c->nod_lu32;
  af.coop = coop ? 1 : 0;
  const dim3 g1(c->ncell, 3, 1), g2(c->nface, 3, 1);
  // the cooperative passes' direction-interleaved 1D grids (dir_blk)
  const dim3 g1c(3 * ((c->ncell + 7) / 8 * 8)), g2c(3 * ((c->nface + 7) / 8 * 8));
  switch (c->G) {
#define NOD_CASE(GG)                                                                                     \
  case GG:                                                                                               \
    hipLaunchKernelGGL(k_nod_src, dim3(c->ncell, GG), dim3(NT), 0, c->s, an);                            \
    if (coop) {                                                                                          \
      hipLaunchKernelGGL(k_nod_bd<GG>, dim3(c->ncell, (NT + NT / GG - 1) / (NT / GG)), dim3(NT), 0, c->s, an); \
      hipLaunchKernelGGL(k_nod_pass1c<GG>, g1c, dim3(NT), 0, c->s, an, (int)c->ncell);                  \
      if (lu32)                                                                                          \
        hipLaunchKernelGGL((k_nod_pass2c<GG, true>), g2c, dim3(NT), 0, c->s, af, (int)c->nface);        \
      else                                                                                               \
        hipLaunchKernelGGL((k_nod_pass2c<GG, false>), g2c, dim3(NT), 0, c->s, af, (int)c->nface);       \
      if (c->nbface)                                                                                     \
        hipLaunchKernelGGL(k_nod_pass2b<GG>, dim3(c->nbface), dim3(NT), 0, c->s, ab, c->bfl, c->nface);  \
    } else {                                                                                             \
      hipLaunchKernelGGL(k_nod_pass1<GG>, g1, dim3(NT), 0, c->s, an);                                    \
      hipLaunchKernelGGL(k_nod_pass2<GG>, g2, dim3(NT), 0, c->s, af);                                    \
    }                                                                                                    \
    break;
    NOD_CASE(1) NOD_CASE(2) NOD_CASE(3) NOD_CASE(4) NOD_CASE(5) NOD_CASE(6) NOD_CASE(7) NOD_CASE(8)
#undef NOD_CASE
  }
  hipLaunchKernelGGL(k_nod_argmax, dim3(c->B), dim3(NT), 0, c->s, c->lv(c->tface), c->npart, c->nface, sched);
  HIPCHK(hipGetLastError());
  launch_matrix(c, sched);
}

// host-side initialisation of the device scalars of every core for an outer
// call; nit = 1: one %ITER for all cores, nit = B: one per core
void set_scal_start(adg_ctx *c, const adg_iter *its, int nit = 1, int fixed_k = 0, int frozen = 0) {
  std::vector<Scal> h(c->B);
  scopy(c->s, h.data(), c->sc, c->B * sizeof(Scal), hipMemcpyDeviceToHost);
  for (int b = 0; b < c->B; ++b) {
    const adg_iter *it = its ? its + (nit == c->B ? b : 0) : nullptr;
    Scal &s = h[b];
    s.done = 0;
    s.status = 0;
    s.p = 0;
    s.flags = 0;
    s.lu_key = ~0ull;
    s.Ke = c->core[b].Ke_host;
    s.init = c->core[b].first ? 1 : 0;
    s.nac = it ? it->nac : 1;
    s.nupd = it ? it->nupd : 1;
    // frozen: the D-hats stay as they are (no nodal update) and ndmax keeps
    // its last value, as in outer_ad called with popt = 0
    s.nodal_on = (it && it->kern != ADG_KERN_FDM && !frozen) ? 1 : 0;
    if (!(it && it->kern != ADG_KERN_FDM)) s.ndmax = 0.0;
    s.serc = it ? it->serc : 0.0;
    s.ferc = it ? it->ferc : 0.0;
    s.aux = c->psifphi;
    s.fixed_k = fixed_k;
  }
  scopy(c->s, c->sc, h.data(), c->B * sizeof(Scal), hipMemcpyHostToDevice);
}

void ensure_trace(adg_ctx *c, int nout) {
  if (nout <= c->tcap) return;
  void *ptrs[] = {c->tk, c->tser, c->tfer, c->tnd, c->tflags, c->tloc};
  for (void *pp : ptrs)
    if (pp) (void)hipFree(pp);
  const int cap = std::max(nout, 2 * c->tcap);
  const size_t n = (size_t)cap * c->B;
  c->tk = dalloc<double>(c->s, n); c->tser = dalloc<double>(c->s, n); c->tfer = dalloc<double>(c->s, n);
  c->tnd = dalloc<double>(c->s, n); c->tflags = dalloc<int>(c->s, n); c->tloc = dalloc<int>(c->s, 3 * n);
  c->tcap = cap;
  drop_graphs(c);   // the captured k_end holds the old trace pointers
}

// One outer iteration (src/mod_cmfd.f90:465-496) as a fixed kernel sequence:
// G group solves, FSrc + error norms, extrapolation + k update, the nodal
// update (its kernels return at once unless this outer is a nodal-update
// outer of the core -- the schedule lives on the device), and the
// end-of-outer kernel.  p_host is the outer index seen by the host
// (profiling only).  smode: 0 forward (outer, outer_fs, outer_th), 1
// adjoint (outer_ad), 2 transient (outer_tr).
void enqueue_iteration(adg_ctx *c, int smode, int nin, int kern, int p_host, int nupd, bool with_nodal = true) {
  const LV lv = c->lv(c->tflat);
  const XS x = c->xs();
  const int adjoint = smode == 1;
  for (int gg = 0; gg < c->G; ++gg) {
    const int g = adjoint ? c->G - 1 - gg : gg;
    group_solve(c, g, smode, gg == 0 ? 1 : 0, nin);
  }
  double *P1 = c->P(1), *P2 = c->P(2);
  if (c->variant == 1) {   // generalized adjoint: change norm, projection of psi, adjoint fission source
    hipLaunchKernelGGL(k_gpt_norm, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->f0c, c->phis, P1);
    hipLaunchKernelGGL(k_gpt_fin, dim3(c->B), dim3(NTF), 0, c->s, lv, P1);
    hipLaunchKernelGGL(k_gpt_proj, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->psi, c->fs0);
    hipLaunchKernelGGL(k_end, dim3((c->B + 63) / 64), dim3(64), 0, c->s, c->sc, c->B, c->tcap, c->tk, c->tser,
                       c->tfer, c->tnd, c->tflags, c->tloc);
    HIPCHK(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_fsrc, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, adjoint, c->f0, c->f0c, c->fs0, c->fs0c,
                     c->errn, P1);
  hipLaunchKernelGGL(k_post, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->fs0, c->fs0c, c->errn, P1, P2);
  hipLaunchKernelGGL(k_fin_outer, dim3(c->B), dim3(NTF), 0, c->s, lv, P2);
  if (kern != ADG_KERN_FDM && with_nodal) {
    const bool timed = p_host > 0 && (p_host % nupd) == 0;
    const int en0 = timed ? prof_event(c) : -1;
    nodal_update(c, kern, adjoint ? 0 : (smode == 2 ? 2 : 1), 1);
    if (timed) prof_pair(c, 2, en0, prof_event(c));
  }
  hipLaunchKernelGGL(k_end, dim3((c->B + 63) / 64), dim3(64), 0, c->s, c->sc, c->B, c->tcap, c->tk, c->tser, c->tfer,
                     c->tnd, c->tflags, c->tloc);
  HIPCHK(hipGetLastError());
}

// status of core b after an outer call, with the message in c->err
int core_status(adg_ctx *c, int b, const Scal &h) {
  int rc = ADG_OK;
  if (h.status != 0) {
    rc = h.status;
    if (rc == ADG_E_LU_PIVOT) {
      const unsigned long long key = h.lu_key;
      char buf[160];
      snprintf(buf, sizeof buf, "LU pivot below 1e-4 at node I = %d, J = %d, K = %d", (int)((key >> 20) & 1023),
               (int)((key >> 10) & 1023), (int)(key & 1023));
      c->err = buf;
      if (((key >> 20) & 1023) == 0) c->err = "single-node line";
    } else if (rc == ADG_E_NODAL_UNSTABLE) {
      char buf[160];
      snprintf(buf, sizeof buf, "Max. change in nodal coupling coefficient = %.1f", h.ndmax);
      c->err = buf;
    } else if (rc == ADG_E_DEVICE_SYNC) {
      c->err = "a device-wide wait of the persistent BiCGSTAB kernel timed out (are its blocks sharing the GPU "
               "with another persistent kernel?); ADG_PERSIST=0 selects the flat kernels";
    }
    if (c->B > 1) c->err += " (core " + std::to_string(b) + ")";
  } else if (!h.done) {
    rc = ADG_E_MAX_OUTER;
    c->err = "maximum number of outer iterations reached";
    if (c->B > 1) c->err += " (core " + std::to_string(b) + ")";
  }
  return rc;
}

// the outer driver for every core of the batch; per-core statuses in c->core.
// nit = 1: one %ITER for all cores; nit = B: one per core (nout, nin and the
// kernel must then agree, the other controls are the core's own)
int run_outer(adg_ctx *c, int mode, const adg_iter *it, int nit = 1) {
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called for every core");
  if (nit != 1 && nit != c->B) return fail(c, ADG_E_ARG, "iteration controls: one, or one per core");
  for (int b = 0; b < nit; ++b) {
    const adg_iter &q = it[b];
    if (q.nout < 1 || q.nin < 1 || q.nac < 1 || q.nupd < 1 || q.kern < 0 || q.kern > 2)
      return fail(c, ADG_E_ARG, "bad iteration control");
    if (q.nout != it->nout || q.nin != it->nin || q.kern != it->kern)
      return fail(c, ADG_E_ARG, "nout, nin and the nodal kernel must be the same for every core");
  }
  if (mode != ADG_FORWARD && mode != ADG_ADJOINT && mode != ADG_FIXED_SOURCE && mode != ADG_TRANSIENT &&
      mode != ADG_ADJOINT_FROZEN)
    return fail(c, ADG_E_ARG, "bad mode");
  const int adjoint = mode == ADG_ADJOINT || mode == ADG_ADJOINT_FROZEN;
  const int smode = adjoint ? 1 : (mode == ADG_TRANSIENT ? 2 : 0);
  const int nout = it->nout;
  std::unique_lock<std::mutex> plk(persist_mu, std::defer_lock);
  if (c->persist) plk.lock();
  const LV lv = c->lv(c->tflat);
  const XS x = c->xs();
  ensure_trace(c, nout);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  bool any_first = false;
  for (int b = 0; b < c->B; ++b) {
    CoreHost &h = c->core[b];
    if (h.first) {
      // f0 = 1, Ke = 1, fs0 = FSrc(f0) on the reference's first call
      std::vector<double> ones((size_t)h.N * c->G, 1.0), st;
      scatter(c, b, ones.data(), c->G, c->f0, st);
      h.Ke_host = 1.0;
      any_first = true;
    }
  }
  set_scal_start(c, it, nit, (mode == ADG_FIXED_SOURCE || mode == ADG_TRANSIENT) ? 1 : 0,
                 mode == ADG_ADJOINT_FROZEN ? 1 : 0);
  c->ev_next = 0;
  c->recs.clear();
  HIPCHK(hipEventRecord(e0, c->s));
  // matrix_setup(1): D-tilde and the matrix from the current D-hat
  launch_coupcoef(c);
  launch_matrix(c, 0);
  if (any_first) {
    hipLaunchKernelGGL(k_fsrc_init, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, adjoint, c->f0, c->fs0, 0);
    for (auto &h : c->core) h.first = false;
  }
  hipLaunchKernelGGL(k_integrals, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->fs0, c->P(0));
  hipLaunchKernelGGL(k_fin_integrals, dim3(c->B), dim3(NTF), 0, c->s, lv, c->P(0));
  HIPCHK(hipGetLastError());
  // replay: a captured graph of one outer (or direct launches when
  // profiling, whose events need the host-side sequence); the done flag of
  // each core written by k_end turns its blocks into no-ops, and the host
  // polls the flags every `check` outers
  const bool use_graph = !c->prof && c->graphs;
  // with one schedule for every core (nit = 1) the host knows which outers
  // update the D-hats (outer q when q % nupd == 0; none when frozen): the
  // others replay a graph without the nodal kernels, which would return at
  // once there.  With one schedule per core every outer replays the full graph.
  const bool split = use_graph && nit == 1 && it->kern != ADG_KERN_FDM;
  const bool frozen = mode == ADG_ADJOINT_FROZEN;
  if (use_graph) {
    const int key[6] = {smode, it->nin, it->kern, it->nupd, c->variant, c->transpose ? 1 : 0};
    if (memcmp(key, c->gkey, sizeof key) != 0) drop_graphs(c);
    if (!c->gexec || (split && !c->gexec_plain)) {
      // one capture at a time in the process (contexts may run on several host threads)
      static std::mutex capture_lock;
      std::lock_guard<std::mutex> lk(capture_lock);
      for (int q = 0; q < 2; ++q) {
        hipGraphExec_t &ge = q == 0 ? c->gexec : c->gexec_plain;
        if (ge || (q == 1 && !split)) continue;
        hipGraph_t graph;
        HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(c, smode, it->nin, it->kern, 0, it->nupd, q == 0);
        HIPCHK(hipStreamEndCapture(c->s, &graph));
        HIPCHK(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
      }
      memcpy(c->gkey, key, sizeof key);
    }
  }
  const int check = c->check_every;
  int launched = 0;
  // replays of the outers the schedule makes nodal-update outers are
  // bracketed by events (one schedule for every core; not while profiling)
  const bool tnod = use_graph && it->kern != ADG_KERN_FDM && nit == 1 && mode != ADG_ADJOINT_FROZEN;
  int nev = 0;
  while (launched < nout) {
    const int n = std::min(check, nout - launched);
    for (int k = 0; k < n; ++k) {
      ++launched;
      const bool bracket = tnod && launched % it->nupd == 0 && nev + 2 <= 4096;
      if (bracket) {
        while ((int)c->nevs.size() < nev + 2) {
          hipEvent_t ev;
          HIPCHK(hipEventCreate(&ev));
          c->nevs.push_back(ev);
        }
        HIPCHK(hipEventRecord(c->nevs[nev], c->s));
      }
      if (use_graph) {
        const bool plain = split && (frozen || launched % it->nupd != 0);
        HIPCHK(hipGraphLaunch(plain ? c->gexec_plain : c->gexec, c->s));
      }
      else enqueue_iteration(c, smode, it->nin, it->kern, launched, it->nupd);
      if (bracket) {
        HIPCHK(hipEventRecord(c->nevs[nev + 1], c->s));
        nev += 2;
      }
    }
    HIPCHK(hipMemcpyAsync(c->hsc, c->sc, c->B * sizeof(Scal), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    bool all = true;
    for (int b = 0; b < c->B && all; ++b) all = c->hsc[b].done != 0;
    if (all) break;
  }
  HIPCHK(hipEventRecord(e1, c->s));
  HIPCHK(hipMemcpyAsync(c->hsc, c->sc, c->B * sizeof(Scal), hipMemcpyDeviceToHost, c->s));
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
    if (!c->prof && nev > 0) {
      // nodal share = time of the nodal-update outers beyond the mean of the others
      double tn = 0.0;
      for (int q = 0; q < nev; q += 2) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->nevs[q], c->nevs[q + 1]));
        tn += t;
      }
      const int nn = nev / 2, no = launched - nn;
      const double other = no > 0 ? (ms - tn) / no : 0.0;
      c->timing[2] = no > 0 ? std::max(0.0, tn - nn * other) : 0.0;
      c->timing[1] = ms - c->timing[2];
      c->timing[3] = 0.0;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int rc = ADG_OK;
  std::string first_err;
  for (int b = 0; b < c->B; ++b) {
    const Scal &h = c->hsc[b];
    c->core[b].Ke_host = h.Ke;
    c->core[b].niter = h.p;
    c->core[b].status = core_status(c, b, h);
    if (c->core[b].status != ADG_OK && rc == ADG_OK) {
      rc = c->core[b].status;
      first_err = c->err;
    }
  }
  if (rc == ADG_E_DEVICE_SYNC) reset_psync(c);
  if (rc != ADG_OK) c->err = first_err;
  return rc;
}

void copy_trace(adg_ctx *c, int b, adg_trace *tr) {
  if (!tr) return;
  const int ncp = std::min(std::min(c->core[b].niter, c->tcap), tr->cap);
  tr->n = ncp;
  if (ncp <= 0) return;
  const size_t o = (size_t)b * c->tcap;
  if (tr->k) scopy(c->s, tr->k, c->tk + o, ncp * sizeof(double), hipMemcpyDeviceToHost);
  if (tr->ser) scopy(c->s, tr->ser, c->tser + o, ncp * sizeof(double), hipMemcpyDeviceToHost);
  if (tr->fer) scopy(c->s, tr->fer, c->tfer + o, ncp * sizeof(double), hipMemcpyDeviceToHost);
  if (tr->ndmax) scopy(c->s, tr->ndmax, c->tnd + o, ncp * sizeof(double), hipMemcpyDeviceToHost);
  if (tr->flags) scopy(c->s, tr->flags, c->tflags + o, ncp * sizeof(int), hipMemcpyDeviceToHost);
  if (tr->loc) scopy(c->s, tr->loc, c->tloc + 3 * o, 3 * ncp * sizeof(int), hipMemcpyDeviceToHost);
}

// ---------------------------------------------------------------------------
// differentiation (single-core contexts)
// ---------------------------------------------------------------------------

int nparam(int G) { return 3 * G + G * (G - 1); }

void diff_alloc(adg_ctx *c) {
  if (c->psi) return;
  const size_t NC = c->NC, G = c->G;
  c->psi = dalloc<double>(c->s, G * NC); c->phis = dalloc<double>(c->s, G * NC); c->gam = dalloc<double>(c->s, G * NC);
  c->exs = dalloc<double>(c->s, G * NC); c->fs0s = dalloc<double>(c->s, NC); c->dJdP = dalloc<double>(c->s, NC);
  c->grad = dalloc<double>(c->s, (size_t)nparam(c->G) * NC);
}

void save_forward(adg_ctx *c) {
  const size_t n = (size_t)c->G * c->NC;
  HIPCHK(hipMemcpyAsync(c->phis, c->f0, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->fs0s, c->fs0, c->NC * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  c->kfwd = c->core[0].Ke_host;
}

void set_device_k(adg_ctx *c, double k) {
  c->core[0].Ke_host = k;
  Scal s;
  scopy(c->s, &s, c->sc, sizeof(Scal), hipMemcpyDeviceToHost);
  s.Ke = k;
  scopy(c->s, c->sc, &s, sizeof(Scal), hipMemcpyHostToDevice);
}

void restore_forward(adg_ctx *c) {
  const size_t n = (size_t)c->G * c->NC;
  HIPCHK(hipMemcpyAsync(c->f0, c->phis, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->fs0, c->fs0s, c->NC * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  c->transpose = false;
  c->variant = 0;
  set_device_k(c, c->kfwd);
}

// sum of the nflat single-value partials in slot order
double host_sum(adg_ctx *c, const double *part) {
  std::vector<double> h(c->nflat);
  HIPCHK(hipMemcpyAsync(h.data(), part, c->nflat * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  double s = 0.0;
  for (double v : h) s += v;
  return s;
}

// sums of the nflat two-value partials
void host_sum2(adg_ctx *c, const double *part, double (&out)[2]) {
  std::vector<double> h(2 * (size_t)c->nflat);
  HIPCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  out[0] = out[1] = 0.0;
  for (int b = 0; b < c->nflat; ++b) {
    out[0] += h[2 * (size_t)b];
    out[1] += h[2 * (size_t)b + 1];
  }
}

// the source kernel's masks of core `core` (host copy -> device)
void push_mask(adg_ctx *c, int core) {
  const size_t o = (size_t)core * 2 * c->G;
  scopy(c->s, c->smask + o, c->hmask.data() + o, 2 * c->G * sizeof(unsigned), hipMemcpyHostToDevice);
}

// every in-scatter group of core `core`, external source as ex (cross
// sections made on the device, or a source written by a kernel)
void mask_all(adg_ctx *c, int core, bool ex) {
  const size_t o = (size_t)core * 2 * c->G;
  const unsigned m = ((1u << c->G) - 1u) | (ex ? SMASK_EX : 0u);
  bool same = true;
  for (int i = 0; i < 2 * c->G; ++i) same = same && c->hmask[o + i] == m;
  if (same) return;
  for (int i = 0; i < 2 * c->G; ++i) c->hmask[o + i] = m;
  push_mask(c, core);
}

// every in-scatter group of core `core`, the external-source bits kept
// (cross sections made on the device)
void mask_scatter_all(adg_ctx *c, int core) {
  const size_t o = (size_t)core * 2 * c->G;
  const unsigned all = (1u << c->G) - 1u;
  bool same = true;
  for (int i = 0; i < 2 * c->G; ++i) same = same && (c->hmask[o + i] & all) == all;
  if (same) return;
  for (int i = 0; i < 2 * c->G; ++i) c->hmask[o + i] |= all;
  push_mask(c, core);
}

// the external-source bit of every group of core `core` on (a kernel
// writes the source); the previous masks are returned for restore_mask
std::vector<unsigned> mask_source_on(adg_ctx *c, int core) {
  std::vector<unsigned> saved = c->hmask;
  const size_t o = (size_t)core * 2 * c->G;
  for (int i = 0; i < 2 * c->G; ++i) c->hmask[o + i] |= SMASK_EX;
  push_mask(c, core);
  return saved;
}

void restore_mask(adg_ctx *c, int core, const std::vector<unsigned> &saved) {
  c->hmask = saved;
  push_mask(c, core);
}

// ---------------------------------------------------------------------------
// transpose of the nodal update and the full-model adjoint (single core)
// ---------------------------------------------------------------------------

void vjp_alloc(adg_ctx *c) {
  if (c->vjp) return;
  const size_t NC = c->NC, G = c->G, PX = vjp_px(c->G), P = nparam(c->G);
  const size_t scratch = (24 * G + 6 * PX + 6 * G + 12 * G + G + PX + 12 * G + G + G + P + 1) * NC;
  const size_t tail = (18 * G + G) * NC;
  c->vjp = dalloc<double>(c->s, scratch + tail);
  c->vjp_scratch = scratch;
  double *p = c->vjp;
  auto take = [&](size_t k) {
    double *r = p;
    p += k * NC;
    return r;
  };
  VjpArgs &v = c->va;
  v.a2b = take(6 * G); v.a4b = take(6 * G); v.l1b = take(6 * G); v.f0b2 = take(6 * G);
  v.xb2 = take(6 * PX); v.dfb2 = take(6 * G);
  v.sdir = take(3 * G); v.y1 = take(3 * G); v.y2 = take(3 * G); v.sb = take(3 * G);
  v.f0b1 = take(G); v.xb1 = take(PX);
  v.dnb = take(6 * G); v.dfb0 = take(6 * G); v.f0b0 = take(G);
  v.phib = take(G); v.grad = take(P); v.kb = take(1);
  c->vw = take(6 * G); c->vwn = take(6 * G); c->vgdn = take(6 * G); c->vsst = take(G);
}

// the forward static nodal passes' arguments at the context's state (per-thread forms)
NodArgs nod_args(adg_ctx *c, int kern) {
  NodArgs a;
  memset(&a, 0, sizeof a);
  a.lv = c->lv(c->tcell);
  a.xs = c->xs();
  a.f0 = c->f0; a.df = c->df; a.dn = c->dn;
  a.S = c->S; a.a2 = c->a2; a.a4 = c->a4; a.L1 = c->L1;
  a.kern = kern; a.cmode = 1;
  a.part = c->npart;
  return a;
}

// df and passes 0 and 1 of the nodal update at the current state (S, a2,
// a4, L1), which the transpose reads
void nodal_fwd(adg_ctx *c, int kern) {
  set_scal_start(c, nullptr);
  launch_coupcoef(c);
  const NodArgs a = nod_args(c, kern);
  switch (c->G) {
#define FWD_CASE(GG)                                                                   \
  case GG:                                                                             \
    hipLaunchKernelGGL(k_nod_src, dim3(c->ncell, GG), dim3(NT), 0, c->s, a);           \
    hipLaunchKernelGGL(k_nod_pass1<GG>, dim3(c->ncell, 3), dim3(NT), 0, c->s, a);      \
    break;
    FWD_CASE(1) FWD_CASE(2) FWD_CASE(3) FWD_CASE(4) FWD_CASE(5) FWD_CASE(6) FWD_CASE(7) FWD_CASE(8)
#undef FWD_CASE
  }
  HIPCHK(hipGetLastError());
}

// transpose of one nodal update at the state nodal_fwd prepared, for the
// cotangent w ([6][G][NC]); exsrc is the external source that update read
// (the generalized-adjoint outers use the context's own array for their
// sources); the flux, old D-hat and parameter cotangents are left in
// c->va, the k cotangent is returned
double nodal_vjp(adg_ctx *c, int kern, const double *w, const double *exsrc) {
  VjpArgs &v = c->va;
  HIPCHK(hipMemsetAsync(c->vjp, 0, c->vjp_scratch * sizeof(double), c->s));
  v.na = nod_args(c, kern);
  v.na.xs.exsrc = exsrc;
  v.w = w;
  VjpArgs vf = v;
  vf.na.lv = c->lv(c->tface);
  const LV lf = c->lv(c->tflat);
  switch (c->G) {
#define VJP_CASE(GG)                                                                               \
  case GG:                                                                                         \
    hipLaunchKernelGGL(k_vjp_face<GG>, dim3(c->nface, 3), dim3(NT), 0, c->s, vf);                   \
    hipLaunchKernelGGL(k_vjp_node<GG>, dim3(c->ncell), dim3(NT), 0, c->s, v);                       \
    hipLaunchKernelGGL(k_vjp_tl, dim3(c->ncell, GG), dim3(NT), 0, c->s, v);                         \
    hipLaunchKernelGGL(k_vjp_src, dim3(c->ncell, GG), dim3(NT), 0, c->s, v);                        \
    hipLaunchKernelGGL(k_vjp_final<GG>, dim3(c->nflat), dim3(NTF), 0, c->s, v, lf);                 \
    break;
    VJP_CASE(1) VJP_CASE(2) VJP_CASE(3) VJP_CASE(4) VJP_CASE(5) VJP_CASE(6) VJP_CASE(7) VJP_CASE(8)
#undef VJP_CASE
  }
  hipLaunchKernelGGL(k_cell_sum, dim3(c->nflat), dim3(NTF), 0, c->s, lf, v.kb, c->P(0));
  HIPCHK(hipGetLastError());
  return host_sum(c, c->P(0));
}

// generalized-adjoint outers (k fixed, the psi component projected out)
// with the source already in exsrc, started from gam (warm) or from zero,
// until its tolerance or its nout; the iterate is left in gam and the
// forward state restored
int gen_adjoint(adg_ctx *c, const adg_iter *it, bool warm) {
  const size_t n = (size_t)c->G * c->NC;
  save_forward(c);
  if (warm) {
    HIPCHK(hipMemcpyAsync(c->f0, c->gam, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    hipLaunchKernelGGL(k_fsrc_init, dim3(c->nflat), dim3(NTF), 0, c->s, c->lv(c->tflat), c->xs(), 1, c->f0, c->fs0,
                       1);
  } else {
    HIPCHK(hipMemsetAsync(c->f0, 0, n * sizeof(double), c->s));
    HIPCHK(hipMemsetAsync(c->fs0, 0, c->NC * sizeof(double), c->s));
  }
  c->transpose = true;
  c->variant = 1;
  adg_iter ig = *it;
  ig.kern = ADG_KERN_FDM;
  const int rc = run_outer(c, ADG_ADJOINT, &ig);
  HIPCHK(hipMemcpyAsync(c->gam, c->f0, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  restore_forward(c);
  return rc;
}

// dJ/dP of the normalised node powers for a response with gradient dJdp
// (PowDis) into c->dJdP; resp = false when it vanishes
int load_response(adg_ctx *c, const double *dJdp, bool &resp) {
  resp = false;
  if (!dJdp) return ADG_OK;
  const LV lv = c->lv(c->tflat);
  const XS x = c->xs();
  const int N = c->core[0].N;
  std::vector<double> st;
  hipLaunchKernelGGL(k_powdis, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->pw, c->P(0));
  hipLaunchKernelGGL(k_scale, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->pw, c->P(0));
  HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  const double T = c->hsc->tpow;
  if (!(T > 0.0)) return fail(c, ADG_E_POWER_NONPOS, "total node power is zero or less");
  std::vector<double> p(N), q(N);
  gather(c, 0, c->pw, 1, p.data(), st);
  double sp = 0.0;
  for (int n = 0; n < N; ++n) sp += dJdp[n] * p[n];
  for (int n = 0; n < N; ++n) {
    q[n] = (dJdp[n] - sp) / T;
    if (q[n] != 0.0) resp = true;
  }
  scatter(c, 0, q.data(), 1, c->dJdP, st);
  return ADG_OK;
}

bool valid_geom(const adg_geom *g) {
  return g && g->nnod > 0 && g->ng > 0 && g->ng <= MAXG && g->nxx > 0 && g->nyy > 0 && g->nzz > 0 && g->ix &&
         g->iy && g->iz && g->xdel && g->ydel && g->zdel &&
         (g->precision == ADG_FP64 || g->precision == ADG_FP32_REFINE) && g->nxx < 1023 && g->nyy < 1023 &&
         g->nzz < 1023;
}

// host description of one core; returns an error message, empty on success
std::string make_core(const adg_geom *g, long base, CoreHost &h, std::vector<unsigned char> &fl,
                      std::vector<double> &vd) {
  h.N = g->nnod; h.nmat = g->nmat;
  h.nxx = g->nxx; h.nyy = g->nyy; h.nzz = g->nzz;
  h.NXP = h.nxx + 2; h.NYP = h.nyy + 2; h.NZP = h.nzz + 2;
  h.sy = h.NXP; h.sz = (long)h.NXP * h.NYP; h.box = h.sz * h.NZP;
  h.base = base;
  // faces x+, x-, y+, y-, z+, z- <- xeast, xwest, ynorth, ysouth, ztop, zbott
  const int map[6] = {0, 1, 2, 3, 5, 4};
  for (int s6 = 0; s6 < 6; ++s6) {
    h.bcface[s6] = g->bc[map[s6]];
    if (h.bcface[s6] < 0 || h.bcface[s6] > 2) return "boundary code outside 0..2";
  }
  h.hx.assign(g->xdel, g->xdel + h.nxx);
  h.hy.assign(g->ydel, g->ydel + h.nyy);
  h.hz.assign(g->zdel, g->zdel + h.nzz);
  // node -> cell, active flags, footprint checks
  fl.assign(h.box, 0);
  vd.assign(h.box, 0.0);
  h.cell.resize(h.N);
  for (int n = 0; n < h.N; ++n) {
    const int i = g->ix[n] - 1, j = g->iy[n] - 1, k = g->iz[n] - 1;
    if (i < 0 || i >= h.nxx || j < 0 || j >= h.nyy || k < 0 || k >= h.nzz) return "node coordinate outside the box";
    const long cc = ((long)(k + 1) * h.NYP + (j + 1)) * h.NXP + (i + 1);
    if (fl[cc]) return "duplicate node coordinate";
    fl[cc] = F_ACTIVE;
    h.cell[n] = base + cc;
  }
  const long st[3] = {1, h.sy, h.sz};
  for (int n = 0; n < h.N; ++n) {
    const long cc = h.cell[n] - base;
    for (int s6 = 0; s6 < 6; ++s6) {
      const long nb = (s6 & 1) ? cc - st[s6 >> 1] : cc + st[s6 >> 1];
      if (fl[nb] & F_ACTIVE) fl[cc] |= f_nb(s6);
    }
  }
  // every line must hold at least two nodes (TLUpd1/2 read both neighbours of an end node)
  for (int n = 0; n < h.N; ++n) {
    const unsigned char f = fl[h.cell[n] - base];
    for (int u = 0; u < 3; ++u)
      if (!(f & f_nb(2 * u)) && !(f & f_nb(2 * u + 1)))
        return "single-node line (the reference's transverse-leakage fit needs two nodes)";
  }
  for (int n = 0; n < h.N; ++n)
    vd[h.cell[n] - base] = h.hx[g->ix[n] - 1] * h.hy[g->iy[n] - 1] * h.hz[g->iz[n] - 1];
  return "";
}

int create_batch(const adg_geom *geoms, int ncore, adg_ctx **out) {
  if (!geoms || !out || ncore < 1) return ADG_E_ARG;
  *out = nullptr;
  for (int b = 0; b < ncore; ++b) {
    if (!valid_geom(&geoms[b])) return ADG_E_ARG;
    if (geoms[b].ng != geoms[0].ng || geoms[b].precision != geoms[0].precision) return ADG_E_ARG;
  }
  adg_ctx *c = new adg_ctx();
  try {
    c->B = ncore;
    HIPCHK(hipGetDevice(&c->dev));
    scratch_init(c->dev);
    c->G = geoms[0].ng;
    c->precision = geoms[0].precision;
    if (const char *ev = getenv("ADG_GRAPH")) c->graphs = atoi(ev) != 0;
    if (const char *ev = getenv("ADG_CHECK")) c->check_every = std::max(1, atoi(ev));
    if (const char *ev = getenv("ADG_NOD_COOP")) c->nod_coop = atoi(ev) != 0;
    if (const char *ev = getenv("ADG_NOD_LU32")) c->nod_lu32 = atoi(ev) != 0;
    if (const char *ev = getenv("ADG_SRC_FULL")) c->src_full = atoi(ev) != 0;
    if (const char *ev = getenv("ADG_XS_LDS")) c->xs_lds = atoi(ev) != 0;
    if (const char *ev = getenv("ADG_PSPIN")) c->pspin = std::max(1L, atol(ev));
    if (const char *ev = getenv("ADG_ADJ_INNER")) c->adj_inner = std::max(0, atoi(ev));
    c->core.resize(ncore);
    std::vector<std::vector<unsigned char>> fls(ncore);
    std::vector<std::vector<double>> vds(ncore);
    long base = 0, hx = 0, hy = 0, hz = 0;
    for (int b = 0; b < ncore; ++b) {
      CoreHost &h = c->core[b];
      const std::string e = make_core(&geoms[b], base, h, fls[b], vds[b]);
      if (!e.empty()) {
        c->err = ncore > 1 ? e + " (core " + std::to_string(b) + ")" : e;
        throw 1;
      }
      h.hoff[0] = hx; h.hoff[1] = hy; h.hoff[2] = hz;
      hx += h.nxx; hy += h.nyy; hz += h.nzz;
      base += h.box;
    }
    c->NC = base;
    HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    {
      std::vector<long> cl;
      for (auto &h : c->core) {
        h.coff = (long)cl.size();
        cl.insert(cl.end(), h.cell.begin(), h.cell.end());
      }
      c->d_cell = dalloc<long>(c->s, cl.size());
      scopy(c->s, c->d_cell, cl.data(), cl.size() * sizeof(long), hipMemcpyHostToDevice);
    }
    const size_t NC = c->NC, G = c->G;
    c->d_hx = dalloc<double>(c->s, hx); c->d_hy = dalloc<double>(c->s, hy); c->d_hz = dalloc<double>(c->s, hz);
    c->flag = dalloc<unsigned char>(c->s, NC);
    c->vdel = dalloc<double>(c->s, NC);
    for (int b = 0; b < ncore; ++b) {
      const CoreHost &h = c->core[b];
      scopy(c->s, c->d_hx + h.hoff[0], h.hx.data(), h.nxx * sizeof(double), hipMemcpyHostToDevice);
      scopy(c->s, c->d_hy + h.hoff[1], h.hy.data(), h.nyy * sizeof(double), hipMemcpyHostToDevice);
      scopy(c->s, c->d_hz + h.hoff[2], h.hz.data(), h.nzz * sizeof(double), hipMemcpyHostToDevice);
      scopy(c->s, c->flag + h.base, fls[b].data(), h.box, hipMemcpyHostToDevice);
      scopy(c->s, c->vdel + h.base, vds[b].data(), h.box * sizeof(double), hipMemcpyHostToDevice);
    }
    c->D = dalloc<double>(c->s, G * NC); c->sigr = dalloc<double>(c->s, G * NC); c->nuf = dalloc<double>(c->s, G * NC);
    c->sigf = dalloc<double>(c->s, G * NC); c->chi = dalloc<double>(c->s, G * NC); c->sigs = dalloc<double>(c->s, G * G * NC);
    c->dc = dalloc<double>(c->s, 6 * G * NC); c->exsrc = dalloc<double>(c->s, G * NC);
    c->f0 = dalloc<double>(c->s, G * NC); c->f0c = dalloc<double>(c->s, G * NC); c->fs0 = dalloc<double>(c->s, NC);
    c->fs0c = dalloc<double>(c->s, NC); c->errn = dalloc<double>(c->s, NC);
    c->dn = dalloc<double>(c->s, 6 * G * NC); c->df = dalloc<double>(c->s, 6 * G * NC);
    c->A64 = dalloc<double>(c->s, 7 * G * NC); c->b = dalloc<double>(c->s, NC); c->pw = dalloc<double>(c->s, NC);
    c->tmp = dalloc<double>(c->s, NC);
    c->wtr = dalloc<double>(c->s, NC);
    if (c->precision == ADG_FP32_REFINE) {
      c->A32 = dalloc<float>(c->s, 7 * G * NC);
      c->rs32 = dalloc<float>(c->s, NC); c->t32 = dalloc<float>(c->s, NC); c->d32 = dalloc<float>(c->s, NC);
      for (int k = 0; k < 2; ++k) {
        c->rr32[k] = dalloc<float>(c->s, NC); c->pp32[k] = dalloc<float>(c->s, NC); c->vv32[k] = dalloc<float>(c->s, NC);
      }
    } else {
      c->rs64 = dalloc<double>(c->s, NC); c->t64 = dalloc<double>(c->s, NC);
      for (int k = 0; k < 2; ++k) {
        c->rr64[k] = dalloc<double>(c->s, NC); c->pp64[k] = dalloc<double>(c->s, NC); c->vv64[k] = dalloc<double>(c->s, NC);
      }
    }
    c->S = dalloc<double>(c->s, 3 * G * NC); c->a2 = dalloc<double>(c->s, 3 * G * NC); c->a4 = dalloc<double>(c->s, 3 * G * NC);
    c->L1 = dalloc<double>(c->s, 3 * G * NC);
    build_tables(c, fls);
    if (nod_coop(c)) {
      c->Bd = dalloc<double>(c->s, G * G * NC);
      c->cf4 = dalloc<double>(c->s, 12 * G * NC);
    }
    c->npart = dalloc<double>(c->s, (size_t)3 * (3 * c->nface + c->nbface));
    c->sc = dalloc<Scal>(c->s, ncore);
    c->smask = dalloc<unsigned>(c->s, (size_t)ncore * 2 * G);
    c->hmask.assign((size_t)ncore * 2 * G, ((1u << G) - 1u) | SMASK_EX);
    scopy(c->s, c->smask, c->hmask.data(), c->hmask.size() * sizeof(unsigned), hipMemcpyHostToDevice);
    HIPCHK(hipHostMalloc((void **)&c->hsc, ncore * sizeof(Scal), hipHostMallocDefault));
  } catch (const HipError &e) {
    const int rc = hip_fail(c, e);
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

// The host table of the map kernels of core h: per node row j its
// footprint [ys0, ys1] (2 nyy ints: all ys0, then all ys1), then for each of
// the nd directions the first node of each of its cnt[d] assemblies and
// the end (cnt[d] + 1 ints).  Returns the table's length, or -1 with the
// message in c->err (at most cap ints are written).
long map_table(adg_ctx *c, const CoreHost &h, int nd, const int *cnt, const int *const *divs, int *ib, long cap) {
  for (int d = 0; d < nd; ++d)
    if (cnt[d] < 1) {
      fail(c, ADG_E_ARG, "empty assembly layout");
      return -1;
    }
  long need = 2L * h.nyy;
  for (int d = 0; d < nd; ++d) need += cnt[d] + 1;
  if (need > cap) {
    fail(c, ADG_E_ARG, "assembly layout table of " + std::to_string(need) + " ints exceeds " + std::to_string(cap));
    return -1;
  }
  int *ys0 = ib, *ys1 = ib + h.nyy;
  for (int j = 0; j < h.nyy; ++j) {
    ys0[j] = h.nxx;
    ys1[j] = -1;
  }
  for (int n = 0; n < h.N; ++n) {
    const long cc = h.cell[n] - h.base;
    const int i = (int)(cc % h.NXP) - 1, j = (int)((cc / h.NXP) % h.NYP) - 1;
    ys0[j] = std::min(ys0[j], i);
    ys1[j] = std::max(ys1[j], i);
  }
  const int tot[3] = {h.nxx, h.nyy, h.nzz};
  long o = 2L * h.nyy;
  for (int d = 0; d < nd; ++d) {
    int s = 0;
    ib[o++] = 0;
    for (int q = 0; q < cnt[d]; ++q) {
      if (divs[d][q] < 1) {
        fail(c, ADG_E_ARG, "assembly division below 1");
        return -1;
      }
      s += divs[d][q];
      ib[o++] = s;
    }
    if (s != tot[d]) {
      fail(c, ADG_E_ARG, "assembly divisions do not add up to the node counts");
      return -1;
    }
  }
  return o;
}

// AsmPow / AxiPow numbers of core `core` (PowDis first) with the device
// buffers di (nib + 3 ints) and dd (nxx nyy + nx ny + nz doubles); every
// element of both that the map kernels read is written by them or by the
// table copy first, so the buffers need no clearing.
int power_maps_run(adg_ctx *c, int core, int nx, int ny, int nz, const int *ib, long nib, int *di, double *dd,
                   double *fasm, int *pmax, double *faxi, int *amax) {
  const CoreHost &h = c->core[core];
  const LV lv = c->lv(c->tflat);
  hipLaunchKernelGGL(k_powdis, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->xs(), c->f0, c->pw, c->P(0));
  hipLaunchKernelGGL(k_scale, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->pw, c->P(0));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->hsc, c->sc, c->B * sizeof(Scal), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (!(c->hsc[core].tpow > 0.0)) return fail(c, ADG_E_POWER_NONPOS, "total node power is zero or less");
  scopy(c->s, di, ib, nib * sizeof(int), hipMemcpyHostToDevice);
  MapArgs a;
  a.pw = c->pw;
  a.hx = c->d_hx + h.hoff[0];
  a.hy = c->d_hy + h.hoff[1];
  a.hz = c->d_hz + h.hoff[2];
  a.vdel = c->vdel;
  a.ys0 = di;
  a.ys1 = di + h.nyy;
  a.xa = di + 2 * h.nyy;
  a.ya = a.xa + nx + 1;
  a.za = a.ya + ny + 1;
  a.imax = di + nib;
  a.fnode = dd;
  a.fasm = dd + (long)h.nxx * h.nyy;
  a.faxi = a.fasm + (long)nx * ny;
  a.base = h.base;
  a.sy = h.sy;
  a.sz = h.sz;
  a.nxx = h.nxx; a.nyy = h.nyy; a.nzz = h.nzz;
  a.nx = nx; a.ny = ny; a.nz = nz;
  hipLaunchKernelGGL(k_map_fnode, dim3((h.nxx * h.nyy + 255) / 256), dim3(256), 0, c->s, a);
  hipLaunchKernelGGL(k_map_sums, dim3((nx * ny + nz + 255) / 256), dim3(256), 0, c->s, a);
  hipLaunchKernelGGL(k_map_final, dim3(1), dim3(64), 0, c->s, a);
  HIPCHK(hipGetLastError());
  int im[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(fasm, a.fasm, sizeof(double) * nx * ny, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(faxi, a.faxi, sizeof(double) * nz, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipMemcpyAsync(im, a.imax, sizeof im, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (pmax) { pmax[0] = im[0]; pmax[1] = im[1]; }
  if (amax) *amax = im[2];
  return ADG_OK;
}

bool core_ok(const adg_ctx *c, int b) { return c && b >= 0 && b < c->B; }

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
    case ADG_E_XS: return "cross section out of range (the reference stops in XS_updt)";
    case ADG_E_DEVICE_SYNC: return "device-wide wait of the persistent BiCGSTAB kernel timed out";
  }
  return "unknown status";
}

const char *adg_last_error(adg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int adg_create(const adg_geom *g, adg_ctx **out) { return create_batch(g, 1, out); }

int adg_create_batch(const adg_geom *geoms, int ncore, adg_ctx **out) { return create_batch(geoms, ncore, out); }

int adg_core_count(adg_ctx *c) { return c ? c->B : 0; }

void adg_destroy(adg_ctx *c) {
  if (!c) return;
  if (c->s) (void)hipStreamSynchronize(c->s);
  void *ptrs[] = {c->geo, c->tflat, c->tcell, c->tface, c->tbface, c->bfl, c->d_hx, c->d_hy, c->d_hz, c->flag, c->D, c->sigr, c->nuf,
                  c->sigf, c->chi, c->sigs, c->dc, c->exsrc, c->vdel, c->f0, c->f0c, c->fs0, c->fs0c, c->errn,
                  c->dn, c->df, c->A64, c->b, c->pw, c->tmp, c->rr64[0], c->rr64[1], c->rs64, c->t64, c->pp64[0],
                  c->pp64[1], c->vv64[0], c->vv64[1], c->A32, c->rr32[0], c->rr32[1], c->rs32, c->t32, c->d32,
                  c->pp32[0], c->pp32[1], c->vv32[0], c->vv32[1], c->S, c->a2, c->a4, c->L1, c->Bd, c->cf4, c->npart, c->part,
                  c->sc, c->psi, c->phis, c->fs0s, c->exs, c->gam, c->dJdP, c->grad, c->wtr, c->d_cell,
                  c->dstage, c->psync, c->vjp, c->smask};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->hsc) (void)hipHostFree(c->hsc);
  for (hipEvent_t e : c->evs) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->nevs) (void)hipEventDestroy(e);
  drop_graphs(c);
  void *tptr[] = {c->tk, c->tser, c->tfer, c->tnd, c->tflags, c->tloc};
  for (void *p : tptr)
    if (p) (void)hipFree(p);
  for (CoreHost &h : c->core) {
    if (h.xtab) (void)hipFree(h.xtab);
    if (h.xnode) (void)hipFree(h.xnode);
    if (h.xmeta) (void)hipFree(h.xmeta);
  }
  if (c->xerr) (void)hipFree(c->xerr);
  for (hipEvent_t e : c->xev)
    if (e) (void)hipEventDestroy(e);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

int adg_set_xs_core(adg_ctx *c, int core, const double *D, const double *sigr, const double *sigs,
                    const double *nuf, const double *sigf, const double *chi, const int *mat, const double *dc,
                    const double *exsrc) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  if (!D || !sigr || !sigs || !nuf || !sigf || !chi || !mat || !dc) return fail(c, ADG_E_ARG, "null XS array");
  try {
    CoreHost &h = c->core[core];
    std::vector<double> st;
    const int N = h.N, G = c->G;
    for (size_t n = 0; n < (size_t)N * G; ++n)
      if (!(D[n] > 0.0)) return fail(c, ADG_E_ARG, "non-positive diffusion coefficient");
    for (int n = 0; n < N; ++n)
      if (mat[n] < 1 || mat[n] > h.nmat) return fail(c, ADG_E_ARG, "material index out of range");
    scatter(c, core, D, G, c->D, st);
    scatter(c, core, sigr, G, c->sigr, st);
    scatter(c, core, nuf, G, c->nuf, st);
    scatter(c, core, sigf, G, c->sigf, st);
    // sigs(n, from, to) -> device [to][from][cell]; which (to, from) pairs
    // are nonzero anywhere, for the source kernel's masks
    std::vector<double> tr((size_t)G * G * N);
    std::vector<unsigned char> nz((size_t)G * G, 0);
    for (int f = 0; f < G; ++f)
      for (int t = 0; t < G; ++t) {
        unsigned char any = 0;
        for (int n = 0; n < N; ++n) {
          const double v = sigs[((size_t)t * G + f) * N + n];
          tr[((size_t)t * G + f) * N + n] = v;
          any |= v != 0.0;
        }
        nz[(size_t)t * G + f] = any;
      }
    scatter(c, core, tr.data(), G * G, c->sigs, st);
    // chi(mat(n), g) expanded per node
    std::vector<double> ch((size_t)G * N);
    for (int g = 0; g < G; ++g)
      for (int n = 0; n < N; ++n) ch[(size_t)g * N + n] = chi[(size_t)g * h.nmat + (mat[n] - 1)];
    scatter(c, core, ch.data(), G, c->chi, st);
    scatter(c, core, dc, 6 * G, c->dc, st);
    std::vector<unsigned char> exnz(G, 0);
    if (exsrc) {
      for (int g = 0; g < G; ++g)
        for (int n = 0; n < N; ++n) exnz[g] |= exsrc[(size_t)g * N + n] != 0.0;
      scatter(c, core, exsrc, G, c->exsrc, st);
    } else {
      std::vector<double> z((size_t)G * N, 0.0);
      scatter(c, core, z.data(), G, c->exsrc, st);
    }
    {
      unsigned *m = c->hmask.data() + (size_t)core * 2 * G;
      for (int g = 0; g < G; ++g) {
        unsigned fw = exnz[g] ? SMASK_EX : 0u, ad = fw;
        for (int hh = 0; hh < G; ++hh) {
          if (nz[(size_t)g * G + hh]) fw |= 1u << hh;    // into g from hh (forward)
          if (nz[(size_t)hh * G + g]) ad |= 1u << hh;    // into hh from g (adjoint)
        }
        m[g] = fw;
        m[G + g] = ad;
      }
      push_mask(c, core);
    }
    h.xs_set = true;
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

// node material map, each node's rod column/level and the levels' running
// depth from the top (crod_updt's and crod_tab_updt's walk), for the table
// forms of the cross-section update; "" or the reason the input is rejected
std::string xs_node_setup(const CoreHost &h, const int *mat, int M, const int *fbmap, int nb,
                          std::vector<int> &node, std::vector<double> &zdz) {
  const int N = h.N;
  const bool rod = fbmap != nullptr;
  for (int n = 0; n < N; ++n)
    if (mat[n] < 1 || mat[n] > M) return "material index out of range";
  if (rod && (nb < 1 || nb > 0xffff || h.nzz > 0x7fff)) return "rod tables need fbmap and 1..65535 banks";
  zdz.assign(2L * h.nzz, 0.0);
  double dum = 0.0;
  for (int k = h.nzz - 1; k >= 0; --k) {
    zdz[k] = dum;
    zdz[h.nzz + k] = h.hz[k];
    dum = dum + h.hz[k];
  }
  node.assign(2L * N, 0);
  std::vector<int> colcount;
  if (rod) colcount.assign((size_t)h.nxx * h.nyy, 0);
  for (int n = 0; n < N; ++n) {
    node[n] = mat[n] - 1;
    if (!rod) continue;
    const long cc = h.cell[n] - h.base;
    const int i = (int)(cc % h.NXP) - 1, j = (int)((cc / h.NXP) % h.NYP) - 1, k = (int)(cc / ((long)h.NXP * h.NYP)) - 1;
    const int b = fbmap[i + h.nxx * j];
    if (b < 0 || b > nb) return "fbmap bank out of range";
    if (b) {
      node[N + n] = b | (k << 16);
      ++colcount[i + h.nxx * j];
    }
  }
  if (rod)
    for (int j = 0; j < h.nyy; ++j)
      for (int i = 0; i < h.nxx; ++i)
        if (fbmap[i + h.nxx * j] > 0 && colcount[i + h.nxx * j] != h.nzz)
          return "rodded column without a node at every level";
  return "";
}

int adg_set_xs_tables_core(adg_ctx *c, int core, const adg_xs_tables *t, const int *mat, const double *chi,
                           const double *dc) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  if (!t || !mat || !chi || !dc) return fail(c, ADG_E_ARG, "null argument");
  CoreHost &h = c->core[core];
  const int N = h.N, G = c->G, M = t->nmat;
  if (M != h.nmat) return fail(c, ADG_E_ARG, "table material count differs from the geometry's");
  for (int q = 0; q < 5; ++q)
    if (!t->tab[ADG_XS_BASE][q]) return fail(c, ADG_E_ARG, "base table missing");
  int on = 0;
  for (int kd = 1; kd < 6; ++kd) {
    if (!t->tab[kd][0]) continue;
    for (int q = 0; q < 5; ++q)
      if (!t->tab[kd][q]) return fail(c, ADG_E_ARG, "incomplete feedback table");
    on |= 1 << kd;
  }
  const bool rod = on & (1 << XK_ROD);
  if (rod && !t->fbmap) return fail(c, ADG_E_ARG, "rod tables need fbmap");
  std::vector<int> node;
  std::vector<double> zdz;
  const std::string e = xs_node_setup(h, mat, M, rod ? t->fbmap : nullptr, t->nb, node, zdz);
  if (!e.empty()) return fail(c, ADG_E_ARG, e);
  try {
    const long K = (4L * G + (long)G * G) * M;
    std::vector<double> tab(6 * K + 2L * h.nzz, 0.0);
    for (int kd = 0; kd < 6; ++kd) {
      if (kd && !(on & (1 << kd))) continue;
      for (int q = 0; q < 4; ++q)
        std::memcpy(&tab[kd * K + (long)q * G * M], t->tab[kd][q], sizeof(double) * G * M);
      std::memcpy(&tab[kd * K + 4L * G * M], t->tab[kd][4], sizeof(double) * G * G * M);
    }
    std::memcpy(&tab[6 * K], zdz.data(), sizeof(double) * zdz.size());
    if (h.xtab) HIPCHK(hipFree(h.xtab));
    h.xtab = nullptr;
    if (h.xnode) HIPCHK(hipFree(h.xnode));
    h.xnode = nullptr;
    if (h.xmeta) HIPCHK(hipFree(h.xmeta));   // left by an earlier branch-table install
    h.xmeta = nullptr;
    h.xtab = dalloc<double>(c->s, tab.size());
    scopy(c->s, h.xtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
    h.xnode = dalloc<int>(c->s, node.size());
    scopy(c->s, h.xnode, node.data(), node.size() * sizeof(int), hipMemcpyHostToDevice);
    if (!c->xerr) c->xerr = dalloc<int>(c->s, 1);
    for (hipEvent_t &e : c->xev)
      if (!e) HIPCHK(hipEventCreate(&e));
    h.xon = on;
    h.xnb = rod ? t->nb : 0;
    h.xnmat = M;
    h.xkind = 1;
    const double refs[7] = {t->rbcon, t->rftem, t->rmtem, t->rcden, t->coreh, t->pos0, t->ssize};
    std::memcpy(h.xref, refs, sizeof refs);
    // what adg_set_xs takes besides the node cross sections
    std::vector<double> st;
    std::vector<double> ch((size_t)G * N);
    for (int g = 0; g < G; ++g)
      for (int n = 0; n < N; ++n) ch[(size_t)g * N + n] = chi[(size_t)g * M + (mat[n] - 1)];
    scatter(c, core, ch.data(), G, c->chi, st);
    scatter(c, core, dc, 6 * G, c->dc, st);
    std::vector<double> z((size_t)G * N, 0.0);
    scatter(c, core, z.data(), G, c->exsrc, st);
    mask_all(c, core, false);
    h.xs_set = false;
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_set_xs_tables(adg_ctx *c, const adg_xs_tables *t, const int *mat, const double *chi, const double *dc) {
  return adg_set_xs_tables_core(c, 0, t, mat, chi, dc);
}

int adg_set_xtab_tables_core(adg_ctx *c, int core, const adg_xtab_tables *t, const int *mat, const double *chi) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  if (!t || !mat || !chi || !t->dims || !t->points || !t->xsec) return fail(c, ADG_E_ARG, "null argument");
  CoreHost &h = c->core[core];
  const int N = h.N, G = c->G, M = t->nmat, E = 4 * G + G * G + 6 * G;
  if (M != h.nmat) return fail(c, ADG_E_ARG, "table material count differs from the geometry's");
  const bool rod = t->fbmap != nullptr;
  std::vector<BrMat> meta(M);
  long np = 0, nx = 0, nr = 0;
  for (int m = 0; m < M; ++m) {
    const int *dm = t->dims + 5 * m;
    for (int q = 0; q < 4; ++q)
      if (dm[q] < 1 || dm[q] > 64) return fail(c, ADG_E_ARG, "branch dimension outside 1..64");
    if (dm[4] != 0 && dm[4] != 1) return fail(c, ADG_E_ARG, "rodded-set flag (dims column 5) outside 0..1");
    const long nrec = (long)dm[0] * dm[1] * dm[2] * dm[3];
    meta[m] = BrMat{dm[0], dm[1], dm[2], dm[3], dm[4], 0, np, nx, dm[4] == 1 ? nr : 0};
    np += dm[0] + dm[1] + dm[2] + dm[3];
    nx += nrec * E;
    if (dm[4] == 1) nr += nrec * E;
  }
  if (nr > 0 && !t->rxsec) return fail(c, ADG_E_ARG, "rodded records missing (a material flags a rodded set)");
  std::vector<int> node;
  std::vector<double> zdz;
  const std::string e = xs_node_setup(h, mat, M, t->fbmap, t->nb, node, zdz);
  if (!e.empty()) return fail(c, ADG_E_ARG, e);
  try {
    // device block: points, records, rodded records, zdum, zdel
    std::vector<double> blk(np + nx + nr + zdz.size());
    std::memcpy(blk.data(), t->points, sizeof(double) * np);
    std::memcpy(blk.data() + np, t->xsec, sizeof(double) * nx);
    if (nr) std::memcpy(blk.data() + np + nx, t->rxsec, sizeof(double) * nr);
    std::memcpy(blk.data() + np + nx + nr, zdz.data(), sizeof(double) * zdz.size());
    if (h.xtab) HIPCHK(hipFree(h.xtab));
    h.xtab = nullptr;
    if (h.xnode) HIPCHK(hipFree(h.xnode));
    h.xnode = nullptr;
    if (h.xmeta) HIPCHK(hipFree(h.xmeta));
    h.xmeta = nullptr;
    h.xtab = dalloc<double>(c->s, blk.size());
    scopy(c->s, h.xtab, blk.data(), blk.size() * sizeof(double), hipMemcpyHostToDevice);
    h.xnode = dalloc<int>(c->s, node.size());
    scopy(c->s, h.xnode, node.data(), node.size() * sizeof(int), hipMemcpyHostToDevice);
    h.xmeta = dalloc<BrMat>(c->s, meta.size());
    scopy(c->s, h.xmeta, meta.data(), meta.size() * sizeof(BrMat), hipMemcpyHostToDevice);
    if (!c->xerr) c->xerr = dalloc<int>(c->s, 1);
    for (hipEvent_t &ev : c->xev)
      if (!ev) HIPCHK(hipEventCreate(&ev));
    h.xkind = 2;
    h.xpts = np;
    h.xrec = nx;
    h.xrrec = nr;
    h.xnb = rod ? t->nb : 0;
    h.xnmat = M;
    h.xon = (1 << XK_FTEM) | (1 << XK_MTEM) | (1 << XK_CDEN) | (rod ? (1 << XK_ROD) : 0);
    h.xref[4] = t->coreh;
    h.xref[5] = t->pos0;
    h.xref[6] = t->ssize;
    std::vector<double> st;
    std::vector<double> ch((size_t)G * N);
    for (int g = 0; g < G; ++g)
      for (int n = 0; n < N; ++n) ch[(size_t)g * N + n] = chi[(size_t)g * M + (mat[n] - 1)];
    scatter(c, core, ch.data(), G, c->chi, st);
    std::vector<double> z((size_t)G * N, 0.0);
    scatter(c, core, z.data(), G, c->exsrc, st);
    mask_all(c, core, false);
    h.xs_set = false;
  } catch (const HipError &ex) {
    return hip_fail(c, ex);
  }
  return ADG_OK;
}

int adg_set_xtab_tables(adg_ctx *c, const adg_xtab_tables *t, const int *mat, const double *chi) {
  return adg_set_xtab_tables_core(c, 0, t, mat, chi);
}

int adg_update_xs_core(adg_ctx *c, int core, double bcon, const double *ftem, const double *mtem,
                       const double *cden, const double *bpos) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  CoreHost &h = c->core[core];
  if (!h.xtab) return fail(c, ADG_E_NOT_SET, "adg_set_xs_tables has not been called");
  const int N = h.N, on = h.xon;
  const double *in[4] = {(on & (1 << XK_FTEM)) ? ftem : nullptr, (on & (1 << XK_MTEM)) ? mtem : nullptr,
                         (on & (1 << XK_CDEN)) ? cden : nullptr, (on & (1 << XK_ROD)) ? bpos : nullptr};
  const int need[4] = {XK_FTEM, XK_MTEM, XK_CDEN, XK_ROD};
  for (int q = 0; q < 4; ++q)
    if ((on & (1 << need[q])) && !in[q]) return fail(c, ADG_E_ARG, "parameter array of an active feedback is null");
  try {
    mask_scatter_all(c, core);
    ensure_stage(c, 3L * N + std::max(h.xnb, 1));
    for (int q = 0; q < 4; ++q)
      if (in[q])
        HIPCHK(hipMemcpyAsync(c->dstage + (long)q * N, in[q], sizeof(double) * (q < 3 ? N : h.xnb),
                              hipMemcpyHostToDevice, c->s));
    HIPCHK(hipMemsetAsync(c->xerr, 0, sizeof(int), c->s));
    if (h.xkind == 2) {
      XtabArgs b;
      b.mats = h.xmeta;
      b.points = h.xtab;
      b.xsec = h.xtab + h.xpts;
      b.rxsec = h.xtab + h.xpts + h.xrec;
      b.zdum = h.xtab + h.xpts + h.xrec + h.xrrec;
      b.zdel = b.zdum + h.nzz;
      b.mat0 = h.xnode;
      b.rodk = h.xnode + N;
      b.ftem = c->dstage;
      b.mtem = c->dstage + N;
      b.cden = c->dstage + 2L * N;
      b.bpos = in[3] ? c->dstage + 3L * N : nullptr;
      b.cell = c->d_cell + h.coff;
      b.D = c->D; b.sigr = c->sigr; b.nuf = c->nuf; b.sigf = c->sigf; b.sigs = c->sigs; b.dc = c->dc;
      b.err = c->xerr;
      b.NC = c->NC;
      b.N = N; b.G = c->G; b.nzz = h.nzz; b.rods = (on >> XK_ROD) & 1;
      b.bcon = bcon;
      b.rod0 = h.xref[4] - h.xref[5];
      b.ssize = h.xref[6];
      const int Er = 4 * c->G + c->G * c->G + 6 * c->G;
      const size_t lds = sizeof(double) * XT_NODES * Er;
      const bool tiled = c->xs_lds && lds <= 48 * 1024;
      const dim3 grid = tiled ? dim3((N + XT_NODES - 1) / XT_NODES) : dim3((N + 255) / 256), blk(tiled ? 512 : 256);
      HIPCHK(hipEventRecord(c->xev[0], c->s));
#define XT_LAUNCH(GG)                                                                  \
  if (tiled)                                                                           \
    hipLaunchKernelGGL(k_xtab_tile<GG>, grid, blk, lds, c->s, b);                      \
  else                                                                                 \
    hipLaunchKernelGGL(k_xtab_update<GG>, grid, blk, 0, c->s, b);
      switch (c->G) {
        case 2: XT_LAUNCH(2) break;
        case 4: XT_LAUNCH(4) break;
        case 8: XT_LAUNCH(8) break;
        default: XT_LAUNCH(0) break;
      }
#undef XT_LAUNCH
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c->xev[1], c->s));
    } else {
    XsArgs a;
    const long K = (4L * c->G + (long)c->G * c->G) * h.xnmat;
    a.tab = h.xtab;
    a.zdum = h.xtab + 6 * K;
    a.zdel = h.xtab + 6 * K + h.nzz;
    a.mat0 = h.xnode;
    a.rodk = h.xnode + N;
    a.ftem = in[0] ? c->dstage : nullptr;
    a.mtem = in[1] ? c->dstage + N : nullptr;
    a.cden = in[2] ? c->dstage + 2L * N : nullptr;
    a.bpos = in[3] ? c->dstage + 3L * N : nullptr;
    a.cell = c->d_cell + h.coff;
    a.D = c->D; a.sigr = c->sigr; a.nuf = c->nuf; a.sigf = c->sigf; a.sigs = c->sigs;
    a.err = c->xerr;
    a.NC = c->NC;
    a.N = N; a.G = c->G; a.M = h.xnmat; a.nzz = h.nzz; a.on = on;
    a.bcon = bcon;
    a.rbcon = h.xref[0]; a.rftem = h.xref[1]; a.rmtem = h.xref[2]; a.rcden = h.xref[3];
    a.rod0 = h.xref[4] - h.xref[5];
    a.ssize = h.xref[6];
    // LDS slots: base (boron folded in) and each present kind after boron
    int slots = 1;
    for (int kd = XK_FTEM; kd < XK_N; ++kd) slots += (on >> kd) & 1;
    const size_t lds = sizeof(double) * slots * K;
    const bool use_lds = c->xs_lds && lds <= 64 * 1024;
    const dim3 blk(use_lds ? 512 : 256);
    const dim3 grid = use_lds ? dim3((unsigned)std::min<long>((N + 511) / 512, 1024)) : dim3((N + 255) / 256);
    HIPCHK(hipEventRecord(c->xev[0], c->s));
#define XS_LAUNCH(GG)                                                                  \
  if (use_lds)                                                                         \
    hipLaunchKernelGGL(k_xs_update_lds<GG>, grid, blk, lds, c->s, a);                  \
  else                                                                                 \
    hipLaunchKernelGGL(k_xs_update<GG>, grid, blk, 0, c->s, a);
    switch (c->G) {
      case 2: XS_LAUNCH(2) break;
      case 4: XS_LAUNCH(4) break;
      case 8: XS_LAUNCH(8) break;
      default: XS_LAUNCH(0) break;
    }
#undef XS_LAUNCH
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->xev[1], c->s));
    }
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, c->xerr, sizeof(int), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->xev[0], c->xev[1]));
    c->timing[0] = ms;
    c->timing[1] = c->timing[2] = c->timing[3] = 0.0;
    if (err) {
      h.xs_set = false;
      return fail(c, ADG_E_XS, "cross section out of range: sigtr < 1e-5, or a negative removal, nu-fission or "
                                "scattering cross section (src/mod_xsec.f90:215,86-150)");
    }
    h.xs_set = true;
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_update_xs(adg_ctx *c, double bcon, const double *ftem, const double *mtem, const double *cden,
                  const double *bpos) {
  return adg_update_xs_core(c, 0, bcon, ftem, mtem, cden, bpos);
}

int adg_get_xs_core(adg_ctx *c, int core, double *D, double *sigr, double *nuf, double *sigf, double *sigs,
                    double *dc) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  if (!c->core[core].xs_set) return fail(c, ADG_E_NOT_SET, "no cross sections set");
  try {
    std::vector<double> st;
    const int G = c->G;
    if (D) gather(c, core, c->D, G, D, st);
    if (sigr) gather(c, core, c->sigr, G, sigr, st);
    if (nuf) gather(c, core, c->nuf, G, nuf, st);
    if (sigf) gather(c, core, c->sigf, G, sigf, st);
    if (sigs) gather(c, core, c->sigs, G * G, sigs, st);
    if (dc) gather(c, core, c->dc, 6 * G, dc, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_get_xs(adg_ctx *c, double *D, double *sigr, double *nuf, double *sigf, double *sigs, double *dc) {
  return adg_get_xs_core(c, 0, D, sigr, nuf, sigf, sigs, dc);
}

int adg_set_xs(adg_ctx *c, const double *D, const double *sigr, const double *sigs, const double *nuf,
               const double *sigf, const double *chi, const int *mat, const double *dc, const double *exsrc) {
  return adg_set_xs_core(c, 0, D, sigr, sigs, nuf, sigf, chi, mat, dc, exsrc);
}

int adg_set_state_core(adg_ctx *c, int core, const double *f0, const double *fs0, double Ke, const double *dn,
                       int first) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  try {
    CoreHost &h = c->core[core];
    std::vector<double> st;
    if (first) {
      h.first = true;
      std::vector<double> z((size_t)6 * c->G * h.N, 0.0);
      scatter(c, core, z.data(), 6 * c->G, c->dn, st);
      return ADG_OK;
    }
    if (!f0 || !fs0 || !dn) return fail(c, ADG_E_ARG, "null state array");
    scatter(c, core, f0, c->G, c->f0, st);
    scatter(c, core, fs0, 1, c->fs0, st);
    scatter(c, core, dn, 6 * c->G, c->dn, st);
    h.Ke_host = Ke;
    h.first = false;
    Scal s;
    scopy(c->s, &s, c->sc + core, sizeof(Scal), hipMemcpyDeviceToHost);
    s.Ke = Ke;
    scopy(c->s, c->sc + core, &s, sizeof(Scal), hipMemcpyHostToDevice);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_set_state(adg_ctx *c, const double *f0, const double *fs0, double Ke, const double *dn, int first) {
  return adg_set_state_core(c, 0, f0, fs0, Ke, dn, first);
}

int adg_get_state_core(adg_ctx *c, int core, double *f0, double *fs0, double *Ke, double *dn, double *df) {
  if (!core_ok(c, core)) return ADG_E_ARG;
  try {
    std::vector<double> st;
    HIPCHK(hipStreamSynchronize(c->s));
    if (f0) gather(c, core, c->f0, c->G, f0, st);
    if (fs0) gather(c, core, c->fs0, 1, fs0, st);
    if (dn) gather(c, core, c->dn, 6 * c->G, dn, st);
    if (df) gather(c, core, c->df, 6 * c->G, df, st);
    if (Ke) {
      Scal s;
      scopy(c->s, &s, c->sc + core, sizeof(Scal), hipMemcpyDeviceToHost);
      *Ke = s.Ke;
    }
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_get_state(adg_ctx *c, double *f0, double *fs0, double *Ke, double *dn, double *df) {
  return adg_get_state_core(c, 0, f0, fs0, Ke, dn, df);
}

int adg_outer(adg_ctx *c, int mode, const adg_iter *it, adg_trace *tr) {
  if (!c || !it) return ADG_E_ARG;
  int rc;
  try {
    rc = run_outer(c, mode, it);
    if (rc != ADG_E_NOT_SET && rc != ADG_E_ARG) copy_trace(c, 0, tr);
  } catch (const HipError &e) {
    rc = hip_fail(c, e);
  }
  return rc;
}

int adg_outer_batch(adg_ctx *c, int mode, const adg_iter *it, int nit, int *status, int *niter, double *keff) {
  if (!c || !it) return ADG_E_ARG;
  int rc;
  try {
    rc = run_outer(c, mode, it, nit);
    if (rc == ADG_E_NOT_SET || rc == ADG_E_ARG) return rc;
    for (int b = 0; b < c->B; ++b) {
      if (status) status[b] = c->core[b].status;
      if (niter) niter[b] = c->core[b].niter;
      if (keff) keff[b] = c->core[b].Ke_host;
    }
  } catch (const HipError &e) {
    rc = hip_fail(c, e);
  }
  return rc;
}

int adg_trace_core(adg_ctx *c, int core, adg_trace *tr) {
  if (!core_ok(c, core) || !tr) return ADG_E_ARG;
  try {
    copy_trace(c, core, tr);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_powdis_core(adg_ctx *c, int core, double *p) {
  if (!core_ok(c, core) || !p) return ADG_E_ARG;
  try {
    const LV lv = c->lv(c->tflat);
    hipLaunchKernelGGL(k_powdis, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->xs(), c->f0, c->pw, c->P(0));
    hipLaunchKernelGGL(k_scale, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->pw, c->P(0));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->hsc, c->sc, c->B * sizeof(Scal), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (!(c->hsc[core].tpow > 0.0)) return fail(c, ADG_E_POWER_NONPOS, "total node power is zero or less");
    std::vector<double> st;
    gather(c, core, c->pw, 1, p, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_powdis(adg_ctx *c, double *p) { return adg_powdis_core(c, 0, p); }

int adg_power_maps_core(adg_ctx *c, int core, int nx, int ny, int nz, const int *xdiv, const int *ydiv,
                        const int *zdiv, double *fasm, int *pmax, double *faxi, int *amax) {
  if (!core_ok(c, core) || !xdiv || !ydiv || !zdiv || !fasm || !faxi) return ADG_E_ARG;
  if (!c->core[core].xs_set) return fail(c, ADG_E_NOT_SET, "cross sections not set");
  const CoreHost &h = c->core[core];
  const int cnt[3] = {nx, ny, nz};
  const int *divs[3] = {xdiv, ydiv, zdiv};
  std::vector<int> ib(2L * std::max(h.nyy, 0) + std::max(nx, 0) + std::max(ny, 0) + std::max(nz, 0) + 3);
  const long nib = map_table(c, h, 3, cnt, divs, ib.data(), (long)ib.size());
  if (nib < 0) return ADG_E_ARG;
  const long ndbl = (long)h.nxx * h.nyy + (long)nx * ny + nz;
  int *di = nullptr;
  double *dd = nullptr;
  try {
    di = dalloc<int>(c->s, nib + 3);
    dd = dalloc<double>(c->s, ndbl);
    const int rc = power_maps_run(c, core, nx, ny, nz, ib.data(), nib, di, dd, fasm, pmax, faxi, amax);
    HIPCHK(hipFree(di));
    di = nullptr;
    HIPCHK(hipFree(dd));
    return rc;
  } catch (const HipError &e) {
    if (di) (void)hipFree(di);
    if (dd) (void)hipFree(dd);
    return hip_fail(c, e);
  }
}

int adg_power_maps_noalloc(adg_ctx *c, int nx, int ny, int nz, const int *xdiv, const int *ydiv, const int *zdiv,
                           double *fasm, int *pmax, double *faxi, int *amax) {
  if (!core_ok(c, 0) || !xdiv || !ydiv || !zdiv || !fasm || !faxi) return ADG_E_ARG;
  if (!c->core[0].xs_set) return fail(c, ADG_E_NOT_SET, "cross sections not set");
  void *base = nullptr;
  hipError_t lookup = hipErrorInvalidDevice;
  if (c->dev >= 0 && c->dev < SCR_MAXDEV) {
    std::lock_guard<std::mutex> lk(scr_lock);
    base = scr_ptr[c->dev];
    lookup = scr_err[c->dev];
  }
  if (!base)
    return fail(c, ADG_E_HIP, std::string("device address of the map scratch unavailable: ") + hipGetErrorString(lookup));
  const CoreHost &h = c->core[0];
  const int cnt[3] = {nx, ny, nz};
  const int *divs[3] = {xdiv, ydiv, zdiv};
  int ib[SCR_INTS];   // host copy of the integer table, on the stack
  const long nib = map_table(c, h, 3, cnt, divs, ib, SCR_INTS - 3);
  if (nib < 0) return ADG_E_ARG;
  const long ndbl = (long)h.nxx * h.nyy + (long)nx * ny + nz;
  if (ndbl > SCR_DBLS)
    return fail(c, ADG_E_ARG, "core too large for the allocation-free maps: nxx*nyy + nx*ny + nz = " +
                                  std::to_string(ndbl) + " exceeds " + std::to_string(SCR_DBLS));
  int *di = (int *)base;
  double *dd = (double *)((char *)base + SCR_INTS * sizeof(int));
  try {
    return power_maps_run(c, 0, nx, ny, nz, ib, nib, di, dd, fasm, pmax, faxi, amax);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
}

int adg_power_maps(adg_ctx *c, int nx, int ny, int nz, const int *xdiv, const int *ydiv, const int *zdiv,
                   double *fasm, int *pmax, double *faxi, int *amax) {
  return adg_power_maps_core(c, 0, nx, ny, nz, xdiv, ydiv, zdiv, fasm, pmax, faxi, amax);
}

int adg_flux_map_core(adg_ctx *c, int core, int nx, int ny, const int *xdiv, const int *ydiv, double norm,
                      double *fasm, int *negf) {
  if (!core_ok(c, core) || !xdiv || !ydiv || !fasm) return ADG_E_ARG;
  const CoreHost &h = c->core[core];
  const int G = c->G;
  if (nx < 1 || ny < 1) return fail(c, ADG_E_ARG, "empty assembly layout");
  std::vector<int> ib;
  std::vector<int> ys0(h.nyy, h.nxx), ys1(h.nyy, -1);
  for (int n = 0; n < h.N; ++n) {
    const long cc = h.cell[n] - h.base;
    const int i = (int)(cc % h.NXP) - 1, j = (int)((cc / h.NXP) % h.NYP) - 1;
    ys0[j] = std::min(ys0[j], i);
    ys1[j] = std::max(ys1[j], i);
  }
  ib.insert(ib.end(), ys0.begin(), ys0.end());
  ib.insert(ib.end(), ys1.begin(), ys1.end());
  const int *divs[2] = {xdiv, ydiv};
  const int cnt[2] = {nx, ny}, tot[2] = {h.nxx, h.nyy};
  for (int d = 0; d < 2; ++d) {
    int s = 0;
    ib.push_back(0);
    for (int q = 0; q < cnt[d]; ++q) {
      if (divs[d][q] < 1) return fail(c, ADG_E_ARG, "assembly division below 1");
      s += divs[d][q];
      ib.push_back(s);
    }
    if (s != tot[d]) return fail(c, ADG_E_ARG, "assembly divisions do not add up to the node counts");
  }
  const long nint = (long)ib.size() + G, ndbl = (long)G * h.nxx * h.nyy + (long)G * nx * ny;
  int *di = nullptr;
  double *dd = nullptr;
  try {
    di = dalloc<int>(c->s, nint);
    dd = dalloc<double>(c->s, ndbl);
    scopy(c->s, di, ib.data(), ib.size() * sizeof(int), hipMemcpyHostToDevice);
    MapArgs a{};
    a.hx = c->d_hx + h.hoff[0];
    a.hy = c->d_hy + h.hoff[1];
    a.hz = c->d_hz + h.hoff[2];
    a.ys0 = di;
    a.ys1 = di + h.nyy;
    a.xa = di + 2 * h.nyy;
    a.ya = a.xa + nx + 1;
    a.negf = di + ib.size();
    a.fnode = dd;
    a.fasm = dd + (long)G * h.nxx * h.nyy;
    a.base = h.base;
    a.sy = h.sy;
    a.sz = h.sz;
    a.nxx = h.nxx; a.nyy = h.nyy; a.nzz = h.nzz;
    a.nx = nx; a.ny = ny; a.nz = 0;
    a.f0 = c->f0;
    a.NC = c->NC;
    a.G = G;
    a.norm = norm;
    hipLaunchKernelGGL(k_flux_fnode, dim3((G * h.nxx * h.nyy + 255) / 256), dim3(256), 0, c->s, a);
    hipLaunchKernelGGL(k_flux_asm, dim3((G * nx * ny + 255) / 256), dim3(256), 0, c->s, a);
    hipLaunchKernelGGL(k_flux_final, dim3(1), dim3(64), 0, c->s, a);
    HIPCHK(hipGetLastError());
    std::vector<int> ng(G, 0);
    HIPCHK(hipMemcpyAsync(fasm, a.fasm, sizeof(double) * G * nx * ny, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipMemcpyAsync(ng.data(), a.negf, sizeof(int) * G, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (negf) *negf = *std::max_element(ng.begin(), ng.end());
    HIPCHK(hipFree(di));
    HIPCHK(hipFree(dd));
  } catch (const HipError &e) {
    if (di) (void)hipFree(di);
    if (dd) (void)hipFree(dd);
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_flux_map(adg_ctx *c, int nx, int ny, const int *xdiv, const int *ydiv, double norm, double *fasm,
                 int *negf) {
  return adg_flux_map_core(c, 0, nx, ny, xdiv, ydiv, norm, fasm, negf);
}

int adg_set_transient_core(adg_ctx *c, int core, const double *wtr) {
  if (!core_ok(c, core) || !wtr) return ADG_E_ARG;
  try {
    std::vector<double> st;
    scatter(c, core, wtr, 1, c->wtr, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_set_transient(adg_ctx *c, const double *wtr) { return adg_set_transient_core(c, 0, wtr); }

int adg_leakage(adg_ctx *c, double *L) {
  if (!c || !L) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "kernel-level entry points take a single-core context");
  try {
    NodArgs a;
    memset(&a, 0, sizeof a);
    a.lv = c->lv(c->tcell);
    a.xs = c->xs();
    a.f0 = c->f0; a.df = c->df; a.dn = c->dn;
    hipLaunchKernelGGL(k_leak, dim3(c->ncell, c->G), dim3(NT), 0, c->s, a, c->S);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    std::vector<double> st;
    gather(c, 0, c->S, 3 * c->G, L, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_nodal_update(adg_ctx *c, int kern, int cmode, double *ndmax, int *loc3) {
  if (!c) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "kernel-level entry points take a single-core context");
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  if (kern != ADG_KERN_PNM && kern != ADG_KERN_SANM) return fail(c, ADG_E_ARG, "nodal kernel must be PNM or SANM");
  try {
    set_scal_start(c, nullptr);
    launch_coupcoef(c);
    HIPCHK(hipGetLastError());
    nodal_update(c, kern, cmode, 0);
    HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (ndmax) *ndmax = c->hsc->ndmax;
    if (loc3) { loc3[0] = c->hsc->im; loc3[1] = c->hsc->jm; loc3[2] = c->hsc->km; }
    if (c->hsc->lu_key != ~0ull) {
      // the first failing system in the reference's sweep order: its node in loc3
      const unsigned long long key = c->hsc->lu_key;
      if (loc3) {
        loc3[0] = (int)((key >> 20) & 1023);
        loc3[1] = (int)((key >> 10) & 1023);
        loc3[2] = (int)(key & 1023);
      }
      char buf[160];
      snprintf(buf, sizeof buf, "LU pivot below 1e-4 at node I = %d, J = %d, K = %d", (int)((key >> 20) & 1023),
               (int)((key >> 10) & 1023), (int)(key & 1023));
      return fail(c, ADG_E_LU_PIVOT, buf);
    }
    if (c->hsc->ndmax > 1.e3) return fail(c, ADG_E_NODAL_UNSTABLE, "nodal update unstable");
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_bicg(adg_ctx *c, int g, int nin, const double *b, double *x) {
  if (!c || !b || !x || g < 0 || g >= c->G || nin < 1) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "kernel-level entry points take a single-core context");
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  try {
    set_scal_start(c, nullptr);
    launch_coupcoef(c);
    launch_matrix(c, 0);
    std::vector<double> st;
    const int N = c->core[0].N;
    // right-hand side goes to c->b, the initial guess to the flux of group g
    scatter(c, 0, b, 1, c->b, st);
    std::vector<double> full((size_t)c->G * N, 0.0);
    gather(c, 0, c->f0, c->G, full.data(), st);
    memcpy(full.data() + (size_t)g * N, x, N * sizeof(double));
    scatter(c, 0, full.data(), c->G, c->f0, st);
    {
      std::unique_lock<std::mutex> plk(persist_mu, std::defer_lock);
      if (c->persist) plk.lock();
      launch_bicg(c, g, nin);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
      HIPCHK(hipStreamSynchronize(c->s));
      if (c->hsc->status == ADG_E_DEVICE_SYNC) {
        reset_psync(c);
        return fail(c, ADG_E_DEVICE_SYNC, "a device-wide wait of the persistent BiCGSTAB kernel timed out");
      }
    }
    gather(c, 0, c->f0, c->G, full.data(), st);
    memcpy(x, full.data() + (size_t)g * N, N * sizeof(double));
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_matvec(adg_ctx *c, int g, const double *x, double *y) {
  if (!c || !x || !y || g < 0 || g >= c->G) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "kernel-level entry points take a single-core context");
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  try {
    launch_coupcoef(c);
    launch_matrix(c, 0);
    std::vector<double> st;
    scatter(c, 0, x, 1, c->b, st);
    HIPCHK(hipMemsetAsync(c->tmp, 0, c->NC * sizeof(double), c->s));
    hipLaunchKernelGGL(k_matvec, dim3(c->nflat), dim3(NTF), 0, c->s, c->lv(c->tflat), c->A64, g, c->b, c->tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    gather(c, 0, c->tmp, 1, y, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_adjoint(adg_ctx *c, const adg_iter *it, double *kadj) {
  if (!c || !it) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "the differentiation entry points take a single-core context");
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  if (c->core[0].first) return fail(c, ADG_E_ARG, "no forward solution: call adg_outer first");
  int rc;
  try {
    diff_alloc(c);
    save_forward(c);
    const LV lv = c->lv(c->tflat);
    // start from the forward flux with its adjoint fission source
    hipLaunchKernelGGL(k_fsrc_init, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->xs(), 1, c->f0, c->fs0, 1);
    c->transpose = true;
    c->variant = 0;
    adg_iter ia = *it;
    ia.kern = ADG_KERN_FDM;   // D-hat held at its converged value
    rc = run_outer(c, ADG_ADJOINT, &ia);
    if (rc == ADG_OK || rc == ADG_E_MAX_OUTER) {
      HIPCHK(hipMemcpyAsync(c->psi, c->f0, (size_t)c->G * c->NC * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      if (kadj) *kadj = c->core[0].Ke_host;
      hipLaunchKernelGGL(k_fdot, dim3(c->nflat), dim3(NTF), 0, c->s, lv, c->xs(), c->phis, c->psi, c->P(0));
      c->psifphi = host_sum(c, c->P(0));
    }
    restore_forward(c);
    c->have_psi = rc == ADG_OK;
  } catch (const HipError &e) {
    rc = hip_fail(c, e);
  }
  return rc;
}

int adg_set_adjoint(adg_ctx *c, const double *psi) {
  if (!c || !psi || c->B != 1) return ADG_E_ARG;
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  try {
    diff_alloc(c);
    std::vector<double> st;
    scatter(c, 0, psi, c->G, c->psi, st);
    hipLaunchKernelGGL(k_fdot, dim3(c->nflat), dim3(NTF), 0, c->s, c->lv(c->tflat), c->xs(), c->f0, c->psi, c->P(0));
    c->psifphi = host_sum(c, c->P(0));
    c->have_psi = true;
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_get_adjoint(adg_ctx *c, double *psi, double *psifphi) {
  if (!c || c->B != 1) return ADG_E_ARG;
  if (!c->have_psi) return fail(c, ADG_E_NOT_SET, "adg_adjoint has not been called");
  try {
    std::vector<double> st;
    if (psi) gather(c, 0, c->psi, c->G, psi, st);
    if (psifphi) *psifphi = c->psifphi;
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_backward(adg_ctx *c, const adg_iter *it, double dJdk, const double *dJdp, double *grad) {
  if (!c || !it || !grad) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "the differentiation entry points take a single-core context");
  if (!c->have_psi) return fail(c, ADG_E_NOT_SET, "adg_adjoint has not been called");
  int rc = ADG_OK;
  try {
    const LV lv = c->lv(c->tflat);
    const XS x = c->xs();
    const int G = c->G;
    std::vector<double> st;
    const double k = c->core[0].Ke_host;
    bool resp = false;
    // node powers at the forward state (PowDis) and dJ/dP of the unnormalised powers
    rc = load_response(c, dJdp, resp);
    if (rc != ADG_OK) return rc;
    if (resp) {
      // generalized adjoint: S* as the external source of an adjoint outer with k fixed
      const size_t n = (size_t)G * c->NC;
      HIPCHK(hipMemcpyAsync(c->exs, c->exsrc, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      hipLaunchKernelGGL(k_sstar, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->dJdP, c->exsrc);
      const std::vector<unsigned> msaved = mask_source_on(c, 0);
      save_forward(c);
      HIPCHK(hipMemsetAsync(c->f0, 0, n * sizeof(double), c->s));
      HIPCHK(hipMemsetAsync(c->fs0, 0, c->NC * sizeof(double), c->s));
      c->transpose = true;
      c->variant = 1;
      adg_iter ig = *it;
      ig.kern = ADG_KERN_FDM;
      rc = run_outer(c, ADG_ADJOINT, &ig);
      HIPCHK(hipMemcpyAsync(c->gam, c->f0, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(c->exsrc, c->exs, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      restore_forward(c);
      restore_mask(c, 0, msaved);
      if (rc != ADG_OK) return rc;
    }
    const double apsi = dJdk * k * k / c->psifphi;
    const int P = nparam(G);
    HIPCHK(hipMemsetAsync(c->grad, 0, (size_t)P * c->NC * sizeof(double), c->s));
    hipLaunchKernelGGL(k_grad, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->psi, apsi,
                       resp ? c->gam : (const double *)nullptr, resp ? c->dJdP : (const double *)nullptr, 1.0 / k,
                       c->grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    gather(c, 0, c->grad, P, grad, st);
  } catch (const HipError &e) {
    rc = hip_fail(c, e);
  }
  return rc;
}

int adg_nodal_vjp(adg_ctx *c, int kern, const double *w, double *phib, double *kb, double *dnb, double *grad) {
  if (!c || !w) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "the differentiation entry points take a single-core context");
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  if (c->core[0].first) return fail(c, ADG_E_ARG, "no state: call adg_set_state or adg_outer first");
  if (kern != ADG_KERN_PNM && kern != ADG_KERN_SANM) return fail(c, ADG_E_ARG, "nodal kernel must be PNM or SANM");
  try {
    vjp_alloc(c);
    std::vector<double> st;
    HIPCHK(hipMemsetAsync(c->vw, 0, (size_t)6 * c->G * c->NC * sizeof(double), c->s));
    scatter(c, 0, w, 6 * c->G, c->vw, st);
    nodal_fwd(c, kern);
    const double k = nodal_vjp(c, kern, c->vw, c->exsrc);
    if (kb) *kb = k;
    if (phib) gather(c, 0, c->va.phib, c->G, phib, st);
    if (dnb) gather(c, 0, c->va.dnb, 6 * c->G, dnb, st);
    if (grad) gather(c, 0, c->va.grad, nparam(c->G), grad, st);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return ADG_OK;
}

int adg_backward_nodal(adg_ctx *c, const adg_iter *it, int kern, double dJdk, const double *dJdp, double tol,
                       int maxit, double *grad, int *nit, double *resid) {
  if (!c || !it || !grad || !(tol > 0.0) || maxit < 1) return ADG_E_ARG;
  if (c->B != 1) return fail(c, ADG_E_ARG, "the differentiation entry points take a single-core context");
  if (!c->have_psi) return fail(c, ADG_E_NOT_SET, "adg_adjoint has not been called");
  if (kern != ADG_KERN_PNM && kern != ADG_KERN_SANM) return fail(c, ADG_E_ARG, "nodal kernel must be PNM or SANM");
  int rc = ADG_OK;
  try {
    vjp_alloc(c);
    const LV lv = c->lv(c->tflat);
    const XS x = c->xs();
    const int G = c->G, P = nparam(c->G);
    const size_t n = (size_t)G * c->NC, nd = 6 * n;
    std::vector<double> st;
    const double k = c->core[0].Ke_host;
    bool resp = false;
    rc = load_response(c, dJdp, resp);
    if (rc != ADG_OK) return rc;
    if (resp) hipLaunchKernelGGL(k_sstar, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->dJdP, c->vsst);
    else HIPCHK(hipMemsetAsync(c->vsst, 0, n * sizeof(double), c->s));
    nodal_fwd(c, kern);
    HIPCHK(hipMemcpyAsync(c->exs, c->exsrc, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipMemsetAsync(c->vw, 0, nd * sizeof(double), c->s));
    const std::vector<unsigned> msaved = mask_source_on(c, 0);
    // w = dJ/d(D-hat) of the CMFD adjoint for the response J + w . N, where
    // N is the nodal update and its transpose adds the flux and k
    // cotangents to the adjoint's sources and the old-D-hat cotangent to w.
    // As the forward outer updates D-hat every few outers, each step runs
    // adj_inner generalized-adjoint outers (warm) between updates of w, so
    // that Gamma and w converge together; the iteration ends when w has
    // settled and the last outers met the adjoint tolerance.
    double apsi = 0.0, res = 1.0;
    bool have_gam = false, gam_on = false, gam_conv = true;
    adg_iter ia = *it;
    if (c->adj_inner > 0) ia.nout = std::min(it->nout, c->adj_inner);
    int iter;
    for (iter = 1; iter <= maxit; ++iter) {
      const bool nod = iter > 1;
      const double kbN = nod ? nodal_vjp(c, kern, c->vw, c->exs) : 0.0;
      gam_on = resp || nod;
      if (gam_on) {
        hipLaunchKernelGGL(k_add2, dim3(c->nflat), dim3(NTF), 0, c->s, lv, (const double *)c->vsst,
                           nod ? (const double *)c->va.phib : (const double *)nullptr, G, c->exsrc);
        rc = gen_adjoint(c, &ia, have_gam);
        gam_conv = rc == ADG_OK;
        if (rc == ADG_E_MAX_OUTER && ia.nout < it->nout) rc = ADG_OK;   // the step's share of outers
        if (rc != ADG_OK) break;
        have_gam = true;
      }
      apsi = (dJdk + kbN) * k * k / c->psifphi;
      HIPCHK(hipMemsetAsync(c->vgdn, 0, nd * sizeof(double), c->s));
      hipLaunchKernelGGL(k_grad_dn, dim3(c->nflat), dim3(NTF), 0, c->s, lv, (const double *)c->f0,
                         (const double *)c->psi, apsi, gam_on ? (const double *)c->gam : (const double *)nullptr,
                         c->vgdn);
      hipLaunchKernelGGL(k_add2, dim3(c->nflat), dim3(NTF), 0, c->s, lv, (const double *)c->vgdn,
                         nod ? (const double *)c->va.dnb : (const double *)nullptr, 6 * G, c->vwn);
      hipLaunchKernelGGL(k_diff_norm, dim3(c->nflat), dim3(NTF), 0, c->s, lv, (const double *)c->vwn,
                         (const double *)c->vw, 6 * G, c->P(0));
      HIPCHK(hipGetLastError());
      double s2[2];
      host_sum2(c, c->P(0), s2);
      res = s2[1] > 0.0 ? sqrt(s2[0] / s2[1]) : 0.0;
      std::swap(c->vw, c->vwn);
      if (nod && res <= tol && gam_conv) break;
    }
    HIPCHK(hipMemcpyAsync(c->exsrc, c->exs, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    restore_mask(c, 0, msaved);
    if (nit) *nit = std::min(iter, maxit);
    if (resid) *resid = res;
    if (rc != ADG_OK) return rc;
    HIPCHK(hipMemsetAsync(c->grad, 0, (size_t)P * c->NC * sizeof(double), c->s));
    hipLaunchKernelGGL(k_grad, dim3(c->nflat), dim3(NTF), 0, c->s, lv, x, c->f0, c->psi, apsi,
                       gam_on ? (const double *)c->gam : (const double *)nullptr,
                       resp ? (const double *)c->dJdP : (const double *)nullptr, 1.0 / k, c->grad);
    if (std::min(iter, maxit) > 1)
      hipLaunchKernelGGL(k_add2, dim3(c->nflat), dim3(NTF), 0, c->s, lv, (const double *)c->grad,
                         (const double *)c->va.grad, P, c->grad);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->s));
    gather(c, 0, c->grad, P, grad, st);
    if (res > tol || !gam_conv) rc = fail(c, ADG_E_MAX_OUTER, "full-model adjoint iteration did not reach its tolerance");
  } catch (const HipError &e) {
    rc = hip_fail(c, e);
  }
  return rc;
}

int adg_autotune(adg_ctx *c, int nin, double *out4) {
  if (!c || nin < 1) return ADG_E_ARG;
  if (!c->all_xs()) return fail(c, ADG_E_NOT_SET, "adg_set_xs has not been called");
  try {
    // candidates: one core, the largest grid; a batch, cells per block
    std::vector<long> cand;
    if (c->B == 1) {
      const CoreHost &h = c->core[0];
      const long full = ((long)h.nzz * h.sz + NTF - 1) / NTF;   // one NTF chunk per block
      for (long g : {256L, 512L, 1024L, 2048L})
        if (g < full) cand.push_back(g);
      cand.push_back(full);
    } else {
      cand = {NTF, 2L * NTF, 4L * NTF};
    }
    const size_t nf = (size_t)c->G * c->NC;
    double *keep = dalloc<double>(c->s, nf);
    HIPCHK(hipMemcpyAsync(keep, c->f0, nf * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    set_scal_start(c, nullptr);
    launch_coupcoef(c);
    launch_matrix(c, 0);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const bool prof = c->prof;
    c->prof = false;
    long best = c->flat_param;
    double tbest = 1e300;
    auto time_solve = [&]() {
      launch_bicg(c, 0, nin);   // warm-up
      float ms = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        HIPCHK(hipEventRecord(e0, c->s));
        launch_bicg(c, 0, nin);
        HIPCHK(hipEventRecord(e1, c->s));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        ms = std::min(ms, t);
      }
      return ms;
    };
    // flat candidates, then (single core with a brick plan) the persistent kernel
    const bool can_persist = c->pkc > 0;
    std::unique_lock<std::mutex> plk(persist_mu, std::defer_lock);
    if (can_persist) plk.lock();
    bool best_persist = false;
    c->persist = false;
    for (long cp : cand) {
      install_flat(c, cp);
      const float ms = time_solve();
      if (ms < tbest) {
        tbest = ms;
        best = cp;
      }
    }
    // the persistent kernel with each checked brick plan
    size_t best_plan = 0;
    if (can_persist) {
      c->persist = true;
      for (size_t q = 0; q < c->pcand.size(); ++q) {
        c->pb = c->pcand[q];
        c->pkc = c->pcand_kc[q];
        const float ms = time_solve();
        HIPCHK(hipMemcpyAsync(c->hsc, c->sc, sizeof(Scal), hipMemcpyDeviceToHost, c->s));
        HIPCHK(hipStreamSynchronize(c->s));
        if (c->hsc->status == ADG_E_DEVICE_SYNC) {
          reset_psync(c);
          set_scal_start(c, nullptr);
        } else if (ms < tbest) {
          tbest = ms;
          best_persist = true;
          best_plan = q;
        }
      }
      c->pb = c->pcand[best_plan];
      c->pkc = c->pcand_kc[best_plan];
    }
    c->prof = prof;
    install_flat(c, best);
    c->persist = best_persist;
    HIPCHK(hipMemcpyAsync(c->f0, keep, nf * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    HIPCHK(hipFree(keep));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (out4) {
      out4[0] = best_persist ? 0.0 : (double)best;
      out4[1] = tbest;
      out4[2] = best_persist ? (double)c->pb.nb : (double)c->nflat;
      out4[3] = (double)(cand.size() + (can_persist ? c->pcand.size() : 0));
    }
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

int adg_persistent_plan(int nxx, int nyy, int nzz, int precision, int ncu, int *plan8) {
  if (!plan8 || nxx < 1 || nyy < 1 || nzz < 1 || nxx > 1022 || nyy > 1022 || nzz > 1022 || ncu < 8 ||
      (precision != ADG_FP64 && precision != ADG_FP32_REFINE))
    return ADG_E_ARG;
  for (int k = 0; k < 8; ++k) plan8[k] = 0;
  const long sy = nxx + 2, sz = sy * (nyy + 2), box = sz * (nzz + 2);
  Brick P{};
  int kc = 0;
  if (!plan_bricks(nxx, nyy, nzz, 0, sy, sz, box, precision == ADG_FP32_REFINE, ncu, P, kc)) return ADG_OK;
  plan8[0] = 1;
  plan8[1] = P.nb;
  plan8[2] = P.BX;
  plan8[3] = P.BY;
  plan8[4] = P.BZ;
  plan8[5] = kc;
  plan8[6] = P.H;
  plan8[7] = verify_plan(P, kc, box).empty() ? 0 : 1;
  return ADG_OK;
}

int adg_solver_info(adg_ctx *c, int *info8) {
  if (!c || !info8) return ADG_E_ARG;
  const Brick &P = c->pb;
  const bool on = c->persist;
  info8[0] = on ? 1 : 0;
  info8[1] = on ? P.nb : c->nflat;
  info8[2] = on ? PNT : NTF;
  info8[3] = on ? P.BX : 0;
  info8[4] = on ? P.BY : 0;
  info8[5] = on ? P.BZ : 0;
  info8[6] = on ? c->pkc : 0;
  info8[7] = on ? P.H : 0;
  return ADG_OK;
}

}  // extern "C"
