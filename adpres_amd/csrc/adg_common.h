// Shared device-side definitions of the gfx950 static k-eigenvalue engine.
//
// Device layout ("padded box"): every node-wise array is stored per group as
// a dense box of (nxx+2) x (nyy+2) x (nzz+2) cells with i fastest, i.e. the
// reference's node list (k slowest, then j, then i; src/mod_io.f90:1310-1323)
// scattered into a box with a one-cell halo.  Cells outside the staggered
// footprint and the halo hold zero cross sections, zero matrix coefficients
// and zero vectors, so the stencil kernels need no bounds tests: every
// neighbour c+-1, c+-sy, c+-sz of an interior-plane cell is inside the box.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace adg {

// Launch shape of the flat stencil kernels (threads per block, largest grid)
// and how a consumer finishes a reduction (per wavefront without barriers,
// or per block); the defaults are the measured best on gfx950.
#ifndef ADG_NT_STORE
#define ADG_NT_STORE 0   // streaming stores for the BiCGSTAB kernels' vector outputs
#endif
#ifndef ADG_KUNROLL
#define ADG_KUNROLL 1   // cells per thread in flight in the BiCGSTAB stencil kernels (1 or 2)
#endif
#ifndef ADG_NTF
#define ADG_NTF 256
#endif
#ifndef ADG_MAXBLK
#define ADG_MAXBLK 1024
#endif
#ifndef ADG_WAVE_REDUCE
#define ADG_WAVE_REDUCE 0
#endif
constexpr int NT = 256;            // threads per block of the per-cell / nodal kernels
constexpr int NTF = ADG_NTF;       // threads per block of the flat stencil kernels
constexpr int MAXG = 8;            // largest group count with a compiled nodal kernel
constexpr int MAXBLK = ADG_MAXBLK; // largest grid of the flat kernels

// per-cell flag bits: bit0 active node; bit (1+s) neighbour present across
// face s, faces s = 0..5 = x+, x-, y+, y-, z+, z- (the reference's 1..6)
constexpr unsigned char F_ACTIVE = 1;
__host__ __device__ inline unsigned char f_nb(int s) { return (unsigned char)(2u << s); }

// Geometry of one core.  A context solves a batch of B >= 1 cores that
// share the group count: core b's padded box occupies cells
// [base, base + (nzz+2) * sz) of every node-wise device array, and a group
// (or face, or group pair) index strides by NC, the batch's total cell count.
struct Geo {
  int nxx, nyy, nzz, G;
  int NXP, NYP;
  long sy, sz, NC;     // strides of this core's box; NC = group stride of the batch arrays
  long base;           // first cell of this core's box
  long c0, M;          // first cell of its interior planes (base + sz), cells in interior planes
  int bc[6];           // per face s: boundary code (east, west, north, south, top, bottom)
  int pfirst, pn;      // partial slots [pfirst, pfirst + pn) of its blocks in the flat launches
  int ffirst, fnb;     // its blocks [ffirst, ffirst + fnb) of the face (nodal pass 2) launches
  int bfirst, bnb;      // its blocks [bfirst, bfirst + bnb) of the boundary-face list launch (0: none)
  const double *hx, *hy, *hz;
  const unsigned char *flag;   // batch-wide, indexed by cell
};

// One block's share of a launch over the batch: flat launches give it the
// cell range [lo, hi), its partial slot, its core's partial slots
// [pfirst, pfirst + pn) and the core's strides; per-node and per-face
// launches give it the local node (or face) indexes [lo, hi) of its core.
struct Blk {
  int core, slot, pfirst, pn;
  long lo, hi, sy, sz;
};

// What every batch kernel receives: per-core geometry and scalars, the
// block table of the launch, and the batch-wide group stride and group
// count.  A flat launch over a single core has no table (blk == nullptr):
// its blocks are derived from the remaining fields, so the drop-in path
// reads no table before its first loads.
struct LV {
  const Geo *geo;
  struct Scal *sc;
  const Blk *blk;
  long NC;
  int G, nb;                     // groups; single core: blocks of its flat launches
  long c0, end, chunk, sy, sz;   // single core: interior cells [c0, end), cells per logical block, strides
};

// Device scalars of one context.  Written by block 0 of the kernel that
// derives them (or by a one-block kernel) and read by later kernels only.
struct Scal {
  // BiCGSTAB (src/mod_cmfd.f90:1213-1241): values of iteration i in slot i & 1,
  // written by block 0 of the kernel that derives them and read by later kernels
  double rho_s[2], alpha_s[2], omega_s[2];
  // outer iteration (src/mod_cmfd.f90:456-496)
  double Ke, f, e1, e2, ser, fer, ndmax, coef;
  double tpow;
  int p, done, status, flags;
  int im, jm, km, init;   // init: the next outer call starts from the first-call state
  unsigned long long lu_key;     // first LU pivot failure (order key << 30 | packed loc)
  // iteration schedule of the running outer call (%ITER): outer p+1 extrapolates
  // when (p+1) % nac == 0 and updates the D-hats when nodal_on && (p+1) % nupd == 0
  int nac, nupd, nodal_on;
  int fixed_k;                   // fixed-source outer (outer_fs): k is not updated
  double serc, ferc;
  double aux;                    // generalized-adjoint outer: psi^T F phi of the projection
};

// true when the outer now running (sc->p + 1) is a nodal-update outer
__device__ __forceinline__ bool nodal_outer(const Scal *sc) {
  return !sc->done && sc->nodal_on && ((sc->p + 1) % sc->nupd) == 0;
}

struct XS {
  const double *D, *sigr, *nuf, *sigf, *chi, *sigs, *dc, *exsrc, *vdel;
  const double *wtr;   // transient fission weight 1 - beta + dfis per cell (outer_tr)
};

struct Range {
  long lo, hi;
};

// The block's entry of the launch table (or, for a single core, the same
// entry computed: the hardware hands block b to XCD b % 8, and with the grid
// a multiple of 8 logical block (b % 8) * (grid / 8) + b / 8 gives every
// XCD one contiguous slab of the box, so the z-neighbour planes of a
// stencil are read through that XCD's own L2) and its core.
struct BlkView {
  Blk b;
  const Geo *q;
  Scal *sc;
};
// TAB: the launch has a table (a batch); the hot kernels are compiled for
// each case so that the single-core view costs no branch before its loads
template <int TAB>   // 1 table, 0 single core, -1 decided at run time
__device__ __forceinline__ BlkView blk_view_t(const LV &lv) {
  BlkView v;
  if (TAB == 1 || (TAB < 0 && lv.blk)) {
    v.b = lv.blk[blockIdx.x];
  } else {
    const int nb = gridDim.x, b = blockIdx.x;
    const int lb = (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
    const long lo = lv.c0 + (long)lb * lv.chunk;
    v.b.core = 0;
    v.b.slot = b;
    v.b.pfirst = 0;
    v.b.pn = nb;
    v.b.lo = lo < lv.end ? lo : lv.end;
    v.b.hi = lo + lv.chunk < lv.end ? lo + lv.chunk : lv.end;
    v.b.sy = lv.sy;
    v.b.sz = lv.sz;
  }
  v.q = lv.geo + v.b.core;
  v.sc = lv.sc + v.b.core;
  return v;
}
__device__ __forceinline__ BlkView blk_view(const LV &lv) { return blk_view_t<-1>(lv); }

// partial slots of core `core` in a per-core (grid = cores) launch
__device__ __forceinline__ void core_slots(const LV &lv, int core, int &first, int &n) {
  if (lv.blk) {
    first = lv.geo[core].pfirst;
    n = lv.geo[core].pn;
  } else {
    first = 0;
    n = lv.nb;
  }
}

// ---------------------------------------------------------------------------
// reductions.  A reducing kernel only writes one partial per block (no
// fences, no atomics) into the block's slot; every block of the kernel that
// consumes the result re-reduces the partials of its core's slots in the
// same fixed order, so each block derives bit-identical scalars and the
// result is deterministic for a given block table.  Kernel boundaries order
// the partial stores before their readers.
// ---------------------------------------------------------------------------

// Sums over the lanes of a wavefront by DPP moves (no LDS round trip), in a
// fixed order, so every wavefront that sums the same values gets the same
// bits, and the result is in every lane.  All 64 lanes must be active.
// dpp_f64 moves a double by a 32-bit DPP control.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// each 8-lane group's sum in every lane of the group: pairs (quad_perm
// 1,0,3,2), quads (quad_perm 2,3,0,1), then the two quads of the group
// (row_half_mirror); each step adds equal operands in both lanes
__device__ __forceinline__ double dpp_sum8(double v) {
  v = v + dpp_f64<0xB1>(v);
  v = v + dpp_f64<0x4E>(v);
  v = v + dpp_f64<0x141>(v);
  return v;
}
// all 64 lanes: the 8-lane sums, the two halves of each 16-lane row
// (row_mirror), then the four row sums read lane by lane
__device__ __forceinline__ double dpp_sum64(double v) {
  v = dpp_sum8(v);
  v = v + dpp_f64<0x140>(v);
  const long long b = __double_as_longlong(v);
  const int lo = (int)(unsigned)(b & 0xffffffffll), hi = (int)(b >> 32);
  double r[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    r[q] = __longlong_as_double((long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, 16 * q) << 32) |
                                            (unsigned)__builtin_amdgcn_readlane(lo, 16 * q)));
  return (r[0] + r[1]) + (r[2] + r[3]);
}

// the flat kernels' lane reduction: the xor butterfly by lane shuffles,
// whose summation order the batching form's FP32+refine outcomes were pinned
// with (the batch tables' launches keep it), or with DPR the DPP sums above
// (single-core launches; sums only)
template <int NV, unsigned MAXMASK, bool DPR = false>
__device__ __forceinline__ void wave_reduce(double (&v)[NV]) {
  static_assert(!DPR || MAXMASK == 0u, "DPP form: sums only");
  if constexpr (DPR) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = dpp_sum64(v[k]);
    return;
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double o = __shfl_xor(v[k], off, 64);
      v[k] = ((MAXMASK >> k) & 1u) ? fmax(v[k], o) : v[k] + o;
    }
  }
}

// block-wide reduction over NTF threads; the result is returned in every thread
template <int NV, unsigned MAXMASK, bool DPR = false>
__device__ __forceinline__ void block_reduce(double (&v)[NV]) {
  __shared__ double sm[(NTF / 64) * NV];
  wave_reduce<NV, MAXMASK, DPR>(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[w * NV + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = sm[k];
    for (int ww = 1; ww < NTF / 64; ++ww)
      a = ((MAXMASK >> k) & 1u) ? fmax(a, sm[ww * NV + k]) : a + sm[ww * NV + k];
    v[k] = a;
  }
}

// this block's contribution -> part[slot * NV + k]
template <int NV, unsigned MAXMASK, bool DPR = false>
__device__ __forceinline__ void block_partial(double (&v)[NV], double *part, int slot) {
  block_reduce<NV, MAXMASK, DPR>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) part[(size_t)slot * NV + k] = v[k];
  }
}

// result over the partial slots [first, first + n) written by an earlier
// kernel.  In the per-wavefront form every wavefront reduces all partials
// on its own (same order in every wavefront, so bit-identical results) with
// no LDS and no barrier.
template <int NV, unsigned MAXMASK, bool DPR = false>
__device__ __forceinline__ void reduce_partials(const double *part, int first, int n, double (&out)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k) out[k] = 0.0;
  const double *pp = part + (size_t)first * NV;
  if (ADG_WAVE_REDUCE) {
    for (int b = (int)(threadIdx.x & 63); b < n; b += 64) {
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const double x = pp[(size_t)b * NV + k];
        out[k] = ((MAXMASK >> k) & 1u) ? fmax(out[k], x) : out[k] + x;
      }
    }
    wave_reduce<NV, MAXMASK, DPR>(out);
    return;
  }
  // four partials per thread are requested before any is added (same
  // summation order as one at a time: slots t, t + NTF, t + 2 NTF, ...)
  for (int b0 = threadIdx.x; b0 < n; b0 += 4 * NTF) {
    double x[4][NV];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * NTF;
#pragma unroll
      for (int k = 0; k < NV; ++k) x[u][k] = b < n ? pp[(size_t)b * NV + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (b0 + u * NTF >= n) break;
#pragma unroll
      for (int k = 0; k < NV; ++k) out[k] = ((MAXMASK >> k) & 1u) ? fmax(out[k], x[u][k]) : out[k] + x[u][k];
    }
  }
  block_reduce<NV, MAXMASK, DPR>(out);
}

}  // namespace adg
