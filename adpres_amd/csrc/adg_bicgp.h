// Persistent BiCGSTAB group solve (cmfd::bicg, src/mod_cmfd.f90:1203-1243)
// for a single core: one launch runs the setup and all nin iterations.
//
// Decomposition.  The node box is cut into bricks of BX x BY x BZ nodes, one
// brick per block.  Every thread keeps up to KC of its brick's nodes -- the
// seven matrix coefficients and x, r, p, v, rs, s, t -- in registers for the
// whole solve; the vector a stencil reads (p, then s) is staged in an LDS
// tile of the brick plus a one-node halo, so the 7-point products read LDS
// only.  Only the brick-surface nodes are exchanged through global memory:
// each block stores s, t, p, v of its surface nodes, and after a device-wide
// synchronisation reads those of the halo nodes and recomputes p (phase A)
// and s (phase B) there from them, so an iteration needs two
// synchronisations, the two reductions of the reference's iteration:
//   A(i): omega_{i-1}, rho_i, beta_i from the previous partials;
//         x += alpha p + omega s, r = s - omega t, p = r + beta (p - omega v)
//         (own nodes; halo nodes' r and p from their stored s, t, p, v);
//         v = A p, partial of (rs, v)                        -> sync
//   B(i): alpha_i = rho_i / (rs, v); s = r - alpha v (halo nodes from the
//         stored v), t = A s; partials (t,t), (t,s), (rs,t), (rs,s) -> sync
// and the final x update follows the last B.  The recurrences and every
// rounding are those of the flat two-kernel form (adg_cmfd.h, KA/KB/KT).
//
// Synchronisation.  A block publishes its partials and then an epoch flag
// in its own 128-byte line; the first block of each XCD (block x leads the
// blocks b = x mod 8) waits for its XCD's flags, sums their partials in a
// fixed order and publishes the XCD sum and flag; every block waits for the
// eight XCD flags and sums the XCD sums in a fixed order.  Every block thus
// derives bit-identical scalars, deterministically.  All exchanged data
// (surface vectors, partials, flags) is written and read with agent-scope
// atomic (coherent) accesses, and a wave publishes a flag only after its
// earlier stores have completed (s_waitcnt), so no cache writeback or
// invalidate is needed.  Every wait is bounded: on timeout the block sets
// an error flag and the core's status and returns, and the other blocks
// return at their next synchronisation.  The blocks must be resident
// together: the launch is sized to the device's occupancy, and the host
// runs one such solve at a time per process.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "adg_common.h"

namespace adg {

constexpr int PNT = 256;              // threads per block of the persistent kernel
constexpr int PKH = 4;                // largest halo nodes per thread
constexpr int PMAXB = 512;            // largest grid (the XCD leader polls nb / 8 <= 64 lines)
constexpr long PSPIN = 1L << 20;      // default polls before a wait gives up
constexpr int PLINE = 16;             // doubles per synchronisation line (128 bytes)
// layout of the synchronisation area (doubles): PMAXB block lines, 8 XCD
// lines, then the control line: [0] epoch base (unsigned), [1] error flag
constexpr long PSYNC_XCD = (long)PLINE * PMAXB;
constexpr long PSYNC_CTL = PSYNC_XCD + 8L * PLINE;
constexpr long PSYNC_DOUBLES = PSYNC_CTL + PLINE;

// the plan of one launch: bricks of the node box and their layout
struct Brick {
  int nxx, nyy, nzz;      // node box
  int nbx, nby, nbz;      // bricks per direction
  int BX, BY, BZ;         // brick extents (the last brick of a direction may be shorter)
  int nbr;                // bricks
  int nb;                 // blocks (nbr rounded up to a multiple of 8)
  int EX, EY, EZ;         // LDS tile: (BX + 2)(BY + 2)(BZ + 2)
  int H;                  // most halo nodes of a brick
  long base, sy, sz, NC;  // cell(i, j, k) = base + (k + 1) sz + (j + 1) sy + i + 1 (0-based i, j, k)
};

struct BrickView {
  int i0, j0, k0, ex, ey, ez;
  bool live;
};

// block b's brick.  XCD x (= b mod 8) takes the contiguous brick range
// [x nb/8, (x + 1) nb/8) in z-major order, so neighbouring bricks share an XCD.
__host__ __device__ inline BrickView brick_of(const Brick &P, int b) {
  BrickView v;
  v.i0 = v.j0 = v.k0 = v.ex = v.ey = v.ez = 0;
  const int per = P.nb / 8;
  const int L = (b % 8) * per + b / 8;
  v.live = L < P.nbr;
  if (!v.live) return v;
  const int bx = L % P.nbx, by = (L / P.nbx) % P.nby, bz = L / (P.nbx * P.nby);
  v.i0 = bx * P.BX;
  v.j0 = by * P.BY;
  v.k0 = bz * P.BZ;
  v.ex = P.nxx - v.i0 < P.BX ? P.nxx - v.i0 : P.BX;
  v.ey = P.nyy - v.j0 < P.BY ? P.nyy - v.j0 : P.BY;
  v.ez = P.nzz - v.k0 < P.BZ ? P.nzz - v.k0 : P.BZ;
  return v;
}

struct PCell {
  long c;      // cell of the padded box
  int e;       // LDS tile index
  bool own;    // a node of the brick (own cells) / inside the node box (halo cells)
  bool surf;   // own cell on the brick's surface
};

__host__ __device__ inline int tile_index(const Brick &P, int lx, int ly, int lz) {
  return (lx + 1) + P.EX * ((ly + 1) + P.EY * (lz + 1));
}

// the q-th node slot of a brick (x fastest)
__host__ __device__ inline PCell own_cell(const Brick &P, const BrickView &v, int q) {
  PCell o;
  o.c = 0;
  o.e = 0;
  o.surf = false;
  const int lx = q % P.BX, ly = (q / P.BX) % P.BY, lz = q / (P.BX * P.BY);
  o.own = v.live && lz < v.ez && ly < v.ey && lx < v.ex;
  if (!o.own) return o;
  o.e = tile_index(P, lx, ly, lz);
  o.c = P.base + (long)(v.k0 + lz + 1) * P.sz + (long)(v.j0 + ly + 1) * P.sy + (v.i0 + lx + 1);
  o.surf = lx == 0 || lx == v.ex - 1 || ly == 0 || ly == v.ey - 1 || lz == 0 || lz == v.ez - 1;
  return o;
}

__host__ __device__ inline int halo_count(const BrickView &v) {
  return v.live ? 2 * (v.ey * v.ez + v.ex * v.ez + v.ex * v.ey) : 0;
}

// the h-th halo node of a brick: the six face layers -x, +x, -y, +y, -z, +z
// (no edges or corners: the 7-point stencil does not read them).  own: the
// node is inside the node box (outside, the padded box holds zeros).
__host__ __device__ inline PCell halo_cell(const Brick &P, const BrickView &v, int h) {
  int lx, ly, lz;
  const int fx = v.ey * v.ez, fy = v.ex * v.ez, fz = v.ex * v.ey;
  if (h < 2 * fx) {
    const int f = h / fx, r = h - f * fx;
    ly = r % v.ey;
    lz = r / v.ey;
    lx = f ? v.ex : -1;
  } else if (h < 2 * fx + 2 * fy) {
    int r = h - 2 * fx;
    const int f = r / fy;
    r -= f * fy;
    lx = r % v.ex;
    lz = r / v.ex;
    ly = f ? v.ey : -1;
  } else {
    int r = h - 2 * fx - 2 * fy;
    const int f = r / fz;
    r -= f * fz;
    lx = r % v.ex;
    ly = r / v.ex;
    lz = f ? v.ez : -1;
  }
  const int i = v.i0 + lx, j = v.j0 + ly, k = v.k0 + lz;
  PCell o;
  o.e = tile_index(P, lx, ly, lz);
  o.own = i >= 0 && i < P.nxx && j >= 0 && j < P.nyy && k >= 0 && k < P.nzz;
  o.c = P.base + (long)(k + 1) * P.sz + (long)(j + 1) * P.sy + (i + 1);
  o.surf = false;
  return o;
}

// the block of the neighbour brick across face d (0..5: -x, +x, -y, +y, -z,
// +z) of block b's brick, or -1
__host__ __device__ inline int neighbour_block(const Brick &P, const BrickView &v, int d) {
  if (!v.live || d > 5) return -1;
  int bx = v.i0 / P.BX, by = v.j0 / P.BY, bz = v.k0 / P.BZ;
  if (d == 0) --bx;
  if (d == 1) ++bx;
  if (d == 2) --by;
  if (d == 3) ++by;
  if (d == 4) --bz;
  if (d == 5) ++bz;
  if (bx < 0 || bx >= P.nbx || by < 0 || by >= P.nby || bz < 0 || bz >= P.nbz) return -1;
  const int L = bx + P.nbx * (by + P.nby * bz), per = P.nb / 8;
  return (L % per) * 8 + L / per;
}

// ---------------------------------------------------------------------------
// host: the plan and its check
// ---------------------------------------------------------------------------

constexpr int PKCMAX = 8;   // most nodes per thread (the kernel instances 1, 2, 4, 8)

inline size_t persist_lds(const Brick &P, bool f32) { return (size_t)P.EX * P.EY * P.EZ * (f32 ? 4 : 8); }

// The brick plans of a node box nxx x nyy x nzz at cell base with strides
// sy, sz: bricks of at most PKCMAX x PNT nodes and PKH x PNT halo nodes whose
// tile fits 64 KB of LDS, at most one block per CU (ncu), ranked by
// (nodes + 2 halo nodes) per brick plus 4 per block (the synchronisation
// grows with the grid), cheapest first (ties: the first in z, y, x extent
// order); at most kmax of them.  kcs: nodes per thread of the kernel
// instance of each plan.  Returns the number of plans.
inline int plan_brick_list(int nxx, int nyy, int nzz, long base, long sy, long sz, long NC, bool f32, int ncu,
                           std::vector<Brick> &plans, std::vector<int> &kcs, int kmax) {
  plans.clear();
  kcs.clear();
  const int maxb = (ncu / 8 * 8 < PMAXB ? ncu / 8 * 8 : PMAXB);
  if (maxb < 8 || NC >= 2147483647L || kmax < 1) return 0;
  struct Cand {
    long cost;
    int BX, BY, BZ;
  };
  std::vector<Cand> cand;
  for (int BZ = 1; BZ <= nzz; ++BZ) {
    const int nbz = (nzz + BZ - 1) / BZ;
    if (BZ > 1 && (nzz + BZ - 2) / (BZ - 1) == nbz) continue;   // the same count with a smaller extent
    for (int BY = 1; BY <= nyy; ++BY) {
      const int nby = (nyy + BY - 1) / BY;
      if (BY > 1 && (nyy + BY - 2) / (BY - 1) == nby) continue;
      for (int BX = 1; BX <= nxx; ++BX) {
        const int nbx = (nxx + BX - 1) / BX;
        if (BX > 1 && (nxx + BX - 2) / (BX - 1) == nbx) continue;
        const long nbr = (long)nbx * nby * nbz;
        const long nb = (nbr + 7) / 8 * 8;
        if (nb > maxb) continue;
        const long cells = (long)BX * BY * BZ;
        const long H = 2L * ((long)BY * BZ + (long)BX * BZ + (long)BX * BY);
        if (cells > (long)PKCMAX * PNT || H > (long)PKH * PNT) continue;
        if ((long)(BX + 2) * (BY + 2) * (BZ + 2) * (f32 ? 4 : 8) > 65536) continue;
        cand.push_back(Cand{cells + 2 * H + 4 * nb, BX, BY, BZ});
      }
    }
  }
  std::stable_sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) { return a.cost < b.cost; });
  for (const Cand &c : cand) {
    if ((int)plans.size() >= kmax) break;
    Brick P{};
    P.BX = c.BX; P.BY = c.BY; P.BZ = c.BZ;
    P.nbx = (nxx + c.BX - 1) / c.BX; P.nby = (nyy + c.BY - 1) / c.BY; P.nbz = (nzz + c.BZ - 1) / c.BZ;
    P.nbr = P.nbx * P.nby * P.nbz;
    P.nb = (P.nbr + 7) / 8 * 8;
    P.H = 2 * (c.BY * c.BZ + c.BX * c.BZ + c.BX * c.BY);
    P.nxx = nxx; P.nyy = nyy; P.nzz = nzz;
    P.EX = P.BX + 2; P.EY = P.BY + 2; P.EZ = P.BZ + 2;
    P.base = base; P.sy = sy; P.sz = sz; P.NC = NC;
    const long cells = (long)P.BX * P.BY * P.BZ;
    int kc = 1;
    while ((long)kc * PNT < cells) kc *= 2;
    if (kc > PKCMAX) continue;
    plans.push_back(P);
    kcs.push_back(kc);
  }
  return (int)plans.size();
}

// the cheapest plan of plan_brick_list; false when no brick shape fits
inline bool plan_bricks(int nxx, int nyy, int nzz, long base, long sy, long sz, long NC, bool f32, int ncu,
                        Brick &best, int &kc) {
  std::vector<Brick> plans;
  std::vector<int> kcs;
  if (plan_brick_list(nxx, nyy, nzz, base, sy, sz, NC, f32, ncu, plans, kcs, 1) < 1) return false;
  best = plans[0];
  kc = kcs[0];
  return true;
}

// Every index the kernel forms for plan P with kc nodes per thread, checked
// on the host before the first launch (box: cells of the padded box): own
// and halo cells inside the padded box, tile indexes inside the tile, every
// node of the box owned by exactly one thread of one block, no tile slot
// written twice in a phase, and every stencil neighbour of an own node
// either written on the tile (own or halo) or a zero slot outside the node
// box.  Returns "" or what failed.
inline std::string verify_plan(const Brick &P, int kc, long box) {
  const long nodes = (long)P.nxx * P.nyy * P.nzz;
  std::vector<unsigned char> owned(nodes, 0);
  const int tile = P.EX * P.EY * P.EZ;
  std::vector<int> mark(tile, -1);
  const long lo = P.base, hi = P.base + box, NYP = P.sz / P.sy;
  if (P.nb % 8 || P.nb > PMAXB || P.nbr > P.nb || kc > PKCMAX) return "grid or instance out of range";
  const int di[6] = {1, -1, 0, 0, 0, 0}, dj[6] = {0, 0, 1, -1, 0, 0}, dk[6] = {0, 0, 0, 0, 1, -1};
  const int de[6] = {1, -1, P.EX, -P.EX, P.EX * P.EY, -P.EX * P.EY};
  for (int b = 0; b < P.nb; ++b) {
    const BrickView v = brick_of(P, b);
    if (v.live && (v.ex < 1 || v.ey < 1 || v.ez < 1)) return "empty brick";
    for (int q = 0; q < kc * PNT; ++q) {
      const PCell o = own_cell(P, v, q);
      if (!o.own) continue;
      if (o.c < lo + P.sz || o.c >= hi - P.sz || o.e < 0 || o.e >= tile) return "own cell out of bounds";
      const long cc = o.c - P.base;
      const int i = (int)(cc % P.sy) - 1, j = (int)((cc / P.sy) % NYP) - 1, k = (int)(cc / P.sz) - 1;
      if (i < 0 || i >= P.nxx || j < 0 || j >= P.nyy || k < 0 || k >= P.nzz) return "own cell outside the node box";
      const long n = ((long)k * P.nyy + j) * P.nxx + i;
      if (owned[n]++) return "node owned twice";
      if (mark[o.e] == b) return "tile slot written twice";
      mark[o.e] = b;
    }
    for (int d = 0; d < 6 && v.live; ++d) {
      // the neighbour brick across face d: its block holds the brick next to this one
      const int nbb = neighbour_block(P, v, d);
      const int di0 = d == 0 ? -P.BX : d == 1 ? P.BX : 0, dj0 = d == 2 ? -P.BY : d == 3 ? P.BY : 0,
                dk0 = d == 4 ? -P.BZ : d == 5 ? P.BZ : 0;
      const int ni0 = v.i0 + di0, nj0 = v.j0 + dj0, nk0 = v.k0 + dk0;
      const bool exists = ni0 >= 0 && ni0 < P.nxx && nj0 >= 0 && nj0 < P.nyy && nk0 >= 0 && nk0 < P.nzz;
      if (exists != (nbb >= 0)) return "neighbour block missing or spurious";
      if (nbb >= 0) {
        if (nbb >= P.nb) return "neighbour block out of range";
        const BrickView w = brick_of(P, nbb);
        if (!w.live || w.i0 != ni0 || w.j0 != nj0 || w.k0 != nk0) return "neighbour block holds another brick";
      }
    }
    const int H = halo_count(v);
    if (H > P.H || H > PKH * PNT) return "halo larger than planned";
    for (int h = 0; h < H; ++h) {
      const PCell o = halo_cell(P, v, h);
      if (o.e < 0 || o.e >= tile) return "halo tile index out of bounds";
      if (o.own && (o.c < lo || o.c >= hi)) return "halo cell out of bounds";
      if (mark[o.e] == b) return "halo slot written twice";
      mark[o.e] = b;
    }
    for (int q = 0; q < kc * PNT; ++q) {
      const PCell o = own_cell(P, v, q);
      if (!o.own) continue;
      const long cc = o.c - P.base;
      const int i = (int)(cc % P.sy) - 1, j = (int)((cc / P.sy) % NYP) - 1, k = (int)(cc / P.sz) - 1;
      for (int d = 0; d < 6; ++d) {
        const int ni = i + di[d], nj = j + dj[d], nk = k + dk[d];
        const bool inside = ni >= 0 && ni < P.nxx && nj >= 0 && nj < P.nyy && nk >= 0 && nk < P.nzz;
        const int ne = o.e + de[d];
        if (ne < 0 || ne >= tile) return "stencil read outside the tile";
        if (inside && mark[ne] != b) return "stencil neighbour not on the tile";
        if (!inside && mark[ne] == b) {
          // a slot outside the node box is written only as a zero halo slot
          bool zero_halo = false;
          for (int h = 0; h < H && !zero_halo; ++h) {
            const PCell ho = halo_cell(P, v, h);
            zero_halo = ho.e == ne && !ho.own;
          }
          if (!zero_halo) return "a zero neighbour slot is overwritten";
        }
      }
    }
  }
  for (long n = 0; n < nodes; ++n)
    if (owned[n] != 1) return "a node of the box is not owned";
  return "";
}

// ---------------------------------------------------------------------------
// coherent accesses and the synchronisation
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ T co_ld(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void co_st(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every memory operation this wave issued has completed
__device__ __forceinline__ void co_drain() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);
  asm volatile("" ::: "memory");
}
__device__ __forceinline__ bool epoch_reached(unsigned v, unsigned e) { return (int)(v - e) >= 0; }
__device__ __forceinline__ unsigned *line_flag(double *line) { return (unsigned *)(line + 8); }

// Device-wide synchronisation with a reduction of four values, in steps so
// that a block can load its halo while the reduction completes:
//   p_publish   every block: its partials, then its epoch flag, in its line
//               (after its earlier stores and loads have completed);
//   p_lead      the first block of each XCD (block x leads the blocks
//               b = x mod 8): waits for its XCD's flags, sums their
//               partials in a fixed order, publishes the XCD sum and flag;
//   p_wait_nb   every block: waits for the flags of its (at most six)
//               neighbour bricks, whose surface values it may then load;
//   p_finish    every block: waits for the eight XCD flags and sums the
//               XCD sums in a fixed order (bit-identical in every block).
// Stores issued by a block before p_publish are seen by coherent loads of
// a neighbour after its p_wait_nb, and of every block after p_finish.  A
// wait that gives up sets the error flag; p_finish then reports failure.

template <int NV>
__device__ __forceinline__ void p_publish(double *sync, unsigned e, const double (&mine)[NV]) {
  static_assert(NV <= 8, "the partials and the flag share a 128-byte line");
  co_drain();
  __syncthreads();
  if (threadIdx.x == 0) {
    double *ln = sync + (long)PLINE * blockIdx.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) co_st(ln + k, mine[k]);
    co_drain();
    co_st(line_flag(ln), e);
  }
}

// the lane sums below are dpp_sum8 / dpp_sum64 of adg_common.h

template <int NV>
__device__ __forceinline__ void p_lead(double *sync, unsigned e, long spin) {
  const int t = threadIdx.x, b = blockIdx.x;
  if (b >= 8 || t >= 64) return;
  int *err = (int *)(sync + PSYNC_CTL) + 1;
  const int per = gridDim.x / 8;
  bool ok = true;
  long n = 0;
  for (;;) {
    const bool got = t >= per || epoch_reached(co_ld(line_flag(sync + (long)PLINE * (t * 8 + b))), e);
    if (__all(got)) break;
    __builtin_amdgcn_s_sleep(1);
    if (++n > spin) {
      ok = false;
      break;
    }
  }
  asm volatile("" ::: "memory");
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  if (t < per) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = co_ld(sync + (long)PLINE * (t * 8 + b) + k);
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = dpp_sum64(v[k]);
  if (t == 0) {
    if (!ok) co_st(err, 1);
    double *xl = sync + PSYNC_XCD + (long)PLINE * b;
#pragma unroll
    for (int k = 0; k < NV; ++k) co_st(xl + k, v[k]);
    co_drain();
    co_st(line_flag(xl), e);
  }
}

// nbr: this thread's neighbour block (lanes 0..5 of wave 0; -1 none)
__device__ __forceinline__ void p_wait_nb(double *sync, unsigned e, int nbr, long spin) {
  __shared__ int snb;
  const int t = threadIdx.x;
  if (t < 64) {
    bool ok = true;
    long n = 0;
    for (;;) {
      const bool got = nbr < 0 || epoch_reached(co_ld(line_flag(sync + (long)PLINE * nbr)), e);
      if (__all(got)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++n > spin) {
        ok = false;
        break;
      }
    }
    if (t == 0) {
      snb = ok ? 0 : 1;
      if (!ok) co_st((int *)(sync + PSYNC_CTL) + 1, 1);
    }
  }
  asm volatile("" ::: "memory");
  __syncthreads();
}

template <int NV>
__device__ __forceinline__ bool p_finish(double *sync, unsigned e, double (&red)[NV], long spin) {
  __shared__ double sred[NV];
  __shared__ int sbad;
  const int t = threadIdx.x;
  int *err = (int *)(sync + PSYNC_CTL) + 1;
  if (t < 64) {
    bool ok = true;
    long n = 0;
    for (;;) {
      const bool got = t >= 8 || epoch_reached(co_ld(line_flag(sync + PSYNC_XCD + (long)PLINE * t)), e);
      if (__all(got)) break;
      __builtin_amdgcn_s_sleep(1);
      if (++n > spin) {
        ok = false;
        break;
      }
    }
    asm volatile("" ::: "memory");
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    if (t < 8) {
#pragma unroll
      for (int k = 0; k < NV; ++k) v[k] = co_ld(sync + PSYNC_XCD + (long)PLINE * t + k);
    }
    const int flagged = t == 0 ? co_ld(err) : 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = dpp_sum8(v[k]);   // lanes 0..7: the eight XCD sums
    if (t == 0) {
      if (!ok) co_st(err, 1);
#pragma unroll
      for (int k = 0; k < NV; ++k) sred[k] = v[k];
      sbad = (!ok || flagged) ? 1 : 0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) red[k] = sred[k];
  return sbad == 0;
}


// block-wide sums of NV values in a fixed order; the result in thread 0
template <int NT, int NV>
__device__ __forceinline__ void p_block_sum(double (&v)[NV]) {
  __shared__ double sm[(NT / 64) * NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = dpp_sum64(v[k]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // no barrier before these writes: between two calls every thread passes
  // p_publish's and p_finish's barriers, which thread 0 reaches only after
  // its reads of sm below
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[w * NV + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = sm[k];
      for (int ww = 1; ww < NT / 64; ++ww) a += sm[ww * NV + k];
      v[k] = a;
    }
  }
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------

template <typename TV>
struct PArgs {
  Brick P;
  const double *A64g;          // FP64 matrix of the group (setup residual)
  const TV *ATg;               // matrix of the iteration (A64g for FP64)
  const double *b;             // source
  double *x64;                 // flux of the group (FP64 iterate, or the base of the FP32 correction)
  TV *S, *T, *P0, *P1, *V0, *V1;   // surface exchange: s, t, and p, v by iteration parity
  double *sync;                // PSYNC_DOUBLES doubles, zeroed at allocation
  Scal *sc;
  int nin;
  int status_timeout;          // status written to the core on a timeout
  long spin;                   // polls before a wait gives up
};

// status and done flag of the core on a timeout (every later kernel skips it)
__device__ __forceinline__ void p_fail(Scal *sc, int status) {
  if (threadIdx.x == 0) {
    co_st(&sc->status, status);
    co_st(&sc->done, 1);
  }
}

// The reduction after v = A p carries one value, (rs, v); the one after
// t = A s four, (t,t), (t,s), (rs,t), (rs,s); the setup's one, (r, r).
template <typename TV, int KC>
__global__ void __launch_bounds__(PNT) __attribute__((amdgpu_waves_per_eu(1, 1))) k_bicg_persist(PArgs<TV> a) {
#pragma clang fp contract(off)
  constexpr int NV = 4;   // values of the reduction after t = A s
  extern __shared__ unsigned char p_lds[];
  TV *W = (TV *)p_lds;
  if (a.sc->done) return;
  const Brick &P = a.P;
  const int t = threadIdx.x;
  const BrickView bv = brick_of(P, blockIdx.x);
  unsigned *ctl = (unsigned *)(a.sync + PSYNC_CTL);
  unsigned e = ctl[0];
  const long NC = P.NC;
  const int sy = (int)P.sy, sz = (int)P.sz;
  const int EX = P.EX, EXY = P.EX * P.EY;
  const int tile = P.EX * P.EY * P.EZ;
  const long spin = a.spin;
  for (int q = t; q < tile; q += PNT) W[q] = (TV)0;
  const int nbr = t < 6 ? neighbour_block(P, bv, t) : -1;
  // this thread's halo nodes, decoded once: cell (-1: none or outside the
  // node box, where the tile keeps zero) and tile slot
  const int H = halo_count(bv);
  int hc[PKH], he[PKH];
#pragma unroll
  for (int k = 0; k < PKH; ++k) {
    const int h = t + k * PNT;
    hc[k] = -1;
    he[k] = -1;
    if (h < H) {
      const PCell o = halo_cell(P, bv, h);
      he[k] = o.e;
      if (o.own) hc[k] = (int)o.c;
    }
  }
  // own nodes: cell (-1: none), tile slot, surface flag
  int oc[KC], oe[KC];
  bool surf[KC];
  TV A[KC][7], x[KC], r[KC], p[KC], v[KC], rs[KC], s[KC], tt[KC];
  double acc[1] = {0.0};
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const PCell o = own_cell(P, bv, t + k * PNT);
    oc[k] = o.own ? (int)o.c : -1;
    oe[k] = o.e;
    surf[k] = o.surf;
    x[k] = r[k] = p[k] = v[k] = rs[k] = s[k] = tt[k] = (TV)0;
#pragma unroll
    for (int q = 0; q < 7; ++q) A[k][q] = (TV)0;
  }
  // setup (bicg lines 1-4): r = b - A x in FP64, rs = r; r goes to S for the
  // neighbours' first p
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    if (oc[k] < 0) continue;
    const long c = oc[k];
    const double *Ag = a.A64g;
    double st = 0.0;
    st = st + Ag[0 * NC + c] * a.x64[c - sz];
    st = st + Ag[1 * NC + c] * a.x64[c - sy];
    st = st + Ag[2 * NC + c] * a.x64[c - 1];
    st = st + Ag[3 * NC + c] * a.x64[c];
    st = st + Ag[4 * NC + c] * a.x64[c + 1];
    st = st + Ag[5 * NC + c] * a.x64[c + sy];
    st = st + Ag[6 * NC + c] * a.x64[c + sz];
    const double rr = a.b[c] - st;
    r[k] = (TV)rr;
    rs[k] = r[k];
    x[k] = sizeof(TV) == 8 ? (TV)a.x64[c] : (TV)0;
#pragma unroll
    for (int q = 0; q < 7; ++q) A[k][q] = a.ATg[q * NC + c];
    acc[0] += (double)r[k] * (double)r[k];
    if (surf[k]) co_st(a.S + c, r[k]);
  }
  // halo values loaded while a synchronisation completes: phase A's s, t, p,
  // v (the first iteration: r in S), phase B's v
  TV hs[PKH], ht[PKH], hp[PKH], hv[PKH], rh[PKH];
#pragma unroll
  for (int k = 0; k < PKH; ++k) hs[k] = ht[k] = hp[k] = hv[k] = (TV)0;
  p_block_sum<PNT, 1>(acc);
  p_publish(a.sync, ++e, acc);
  p_lead<1>(a.sync, e, spin);
  p_wait_nb(a.sync, e, nbr, spin);
#pragma unroll
  for (int k = 0; k < PKH; ++k)
    if (hc[k] >= 0) hs[k] = co_ld(a.S + hc[k]);
  double red0[1];
  if (!p_finish(a.sync, e, red0, spin)) {
    p_fail(a.sc, a.status_timeout);
    return;
  }
  // alpha of the iteration and of the one before; red: the last reduction
  // after t = A s
  double rho = red0[0], alpha = 0.0, alpha_prev = 0.0, omega = 0.0, beta = 0.0;
  double red[NV];
  for (int it = 1; it <= a.nin; ++it) {
    const int cur = it & 1;
    TV *Pw = cur ? a.P1 : a.P0, *Vw = cur ? a.V1 : a.V0;
    if (it > 1) {
      omega = red[1] / red[0];
      const double rn = red[3] - omega * red[2];
      beta = (rn / rho) * (alpha / omega);
      rho = rn;
      alpha_prev = alpha;
    }
    const TV ta = (TV)alpha_prev, tw = (TV)omega, tb = (TV)beta;
    // ---- phase A: p on the tile (halo nodes from their stored s, t, p, v)
#pragma unroll
    for (int k = 0; k < PKH; ++k) {
      rh[k] = (TV)0;
      if (he[k] < 0) continue;
      TV pv = (TV)0;
      if (hc[k] >= 0) {
        if (it == 1) {
          rh[k] = hs[k];
          pv = rh[k];
        } else {
          rh[k] = (TV)(hs[k] - tw * ht[k]);
          pv = (TV)(rh[k] + tb * (hp[k] - tw * hv[k]));
        }
      }
      W[he[k]] = pv;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      if (oc[k] < 0) continue;
      if (it > 1) {
        x[k] = (TV)(x[k] + ta * p[k] + tw * s[k]);
        r[k] = (TV)(s[k] - tw * tt[k]);
        p[k] = (TV)(r[k] + tb * (p[k] - tw * v[k]));
      } else {
        p[k] = r[k];
      }
      W[oe[k]] = p[k];
      if (surf[k]) co_st(Pw + oc[k], p[k]);
    }
    __syncthreads();
    double acca[1] = {0.0};
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      if (oc[k] < 0) continue;
      const int q = oe[k];
      const TV pn[7] = {W[q - EXY], W[q - EX], W[q - 1], W[q], W[q + 1], W[q + EX], W[q + EXY]};
      double sv = 0.0;
#pragma unroll
      for (int m = 0; m < 7; ++m) sv = sv + (double)(A[k][m] * pn[m]);
      v[k] = (TV)sv;
      acca[0] += (double)rs[k] * (double)v[k];
      if (surf[k]) co_st(Vw + oc[k], v[k]);
    }
    p_block_sum<PNT, 1>(acca);
    p_publish(a.sync, ++e, acca);
    p_lead<1>(a.sync, e, spin);
    p_wait_nb(a.sync, e, nbr, spin);
#pragma unroll
    for (int k = 0; k < PKH; ++k)
      if (hc[k] >= 0) hv[k] = co_ld(Vw + hc[k]);
    double ra[1];
    if (!p_finish(a.sync, e, ra, spin)) {
      p_fail(a.sc, a.status_timeout);
      return;
    }
    alpha = rho / ra[0];
    const TV ta2 = (TV)alpha;
    // ---- phase B: s on the tile (halo nodes from their stored v)
#pragma unroll
    for (int k = 0; k < PKH; ++k) {
      if (he[k] < 0) continue;
      W[he[k]] = hc[k] >= 0 ? (TV)(rh[k] - ta2 * hv[k]) : (TV)0;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      if (oc[k] < 0) continue;
      s[k] = (TV)(r[k] - ta2 * v[k]);
      W[oe[k]] = s[k];
      if (surf[k]) co_st(a.S + oc[k], s[k]);
    }
    __syncthreads();
    double accb[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) accb[q] = 0.0;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      if (oc[k] < 0) continue;
      const int q = oe[k];
      const TV sn[7] = {W[q - EXY], W[q - EX], W[q - 1], W[q], W[q + 1], W[q + EX], W[q + EXY]};
      double tv = 0.0;
#pragma unroll
      for (int m = 0; m < 7; ++m) tv = tv + (double)(A[k][m] * sn[m]);
      tt[k] = (TV)tv;
      const double rsd = (double)rs[k];
      accb[0] += (double)tt[k] * (double)tt[k];
      accb[1] += (double)tt[k] * (double)s[k];
      accb[2] += rsd * (double)tt[k];
      accb[3] += rsd * (double)s[k];
      if (surf[k]) co_st(a.T + oc[k], tt[k]);
    }
    p_block_sum<PNT, NV>(accb);
    p_publish(a.sync, ++e, accb);
    p_lead<NV>(a.sync, e, spin);
    if (it < a.nin) {
      // the next phase A's halo: s, t of this phase and p, v of phase A
      p_wait_nb(a.sync, e, nbr, spin);
#pragma unroll
      for (int k = 0; k < PKH; ++k)
        if (hc[k] >= 0) {
          hs[k] = co_ld(a.S + hc[k]);
          ht[k] = co_ld(a.T + hc[k]);
          hp[k] = co_ld(Pw + hc[k]);
          hv[k] = co_ld(Vw + hc[k]);
        }
    }
    if (!p_finish(a.sync, e, red, spin)) {
      p_fail(a.sc, a.status_timeout);
      return;
    }
  }
  // the last iteration's x update (the flat form's KT)
  omega = red[1] / red[0];
  const TV ta = (TV)alpha, tw = (TV)omega;
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    if (oc[k] < 0) continue;
    const TV xn = (TV)(x[k] + ta * p[k] + tw * s[k]);
    if (sizeof(TV) == 8) a.x64[oc[k]] = (double)xn;
    else a.x64[oc[k]] = a.x64[oc[k]] + (double)xn;
  }
  if (blockIdx.x == 0 && t == 0) ctl[0] = e;   // the next launch's epoch base
}

}  // namespace adg
