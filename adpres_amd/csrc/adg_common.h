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

constexpr int NT = 256;        // threads per block of the flat (stencil) kernels
constexpr int MAXG = 8;        // largest group count with a compiled nodal kernel
constexpr int MAXBLK = 2048;   // largest grid of the flat kernels (partials buffer)

// per-cell flag bits: bit0 active node; bit (1+s) neighbour present across
// face s, faces s = 0..5 = x+, x-, y+, y-, z+, z- (the reference's 1..6)
constexpr unsigned char F_ACTIVE = 1;
__host__ __device__ inline unsigned char f_nb(int s) { return (unsigned char)(2u << s); }

struct Geo {
  int nxx, nyy, nzz, G;
  int NXP, NYP;
  long sy, sz, NC;     // strides and total cells of one group box
  long c0, M;          // first cell of interior planes, cells in interior planes
  int bc[6];           // per face s: boundary code (east, west, north, south, top, bottom)
  const double *hx, *hy, *hz;
  const unsigned char *flag;
};

// Device scalars of one context.  Written by the single "last block" of a
// reduction or by the end-of-outer kernel, read by every later kernel.
struct Scal {
  // BiCGSTAB (src/mod_cmfd.f90:1213-1241)
  double rho, rho_prev, alpha, omega, beta;
  // outer iteration (src/mod_cmfd.f90:456-496)
  double Ke, f, e1, e2, ser, fer, ndmax, coef;
  double tpow;
  int p, done, status, flags;
  int im, jm, km, pad;
  unsigned long long lu_key;     // first LU pivot failure (order key << 30 | packed loc)
  // iteration schedule of the running outer call (%ITER): outer p+1 extrapolates
  // when (p+1) % nac == 0 and updates the D-hats when nodal_on && (p+1) % nupd == 0
  int nac, nupd, nodal_on, pad2;
  double serc, ferc;
};

// true when the outer now running (sc->p + 1) is a nodal-update outer
__device__ __forceinline__ bool nodal_outer(const Scal *sc) {
  return !sc->done && sc->nodal_on && ((sc->p + 1) % sc->nupd) == 0;
}

struct XS {
  const double *D, *sigr, *nuf, *sigf, *chi, *sigs, *dc, *exsrc, *vdel;
};

// ---------------------------------------------------------------------------
// reductions: wavefront butterfly (DPP/permute shuffles), 4-wave block
// combine through LDS, per-block partials and a last-block final pass in a
// fixed order, so every reduction is deterministic for a given grid size.
// ---------------------------------------------------------------------------

template <int NV, unsigned MAXMASK>
__device__ __forceinline__ void wave_reduce(double (&v)[NV]) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double o = __shfl_xor(v[k], off, 64);
      v[k] = ((MAXMASK >> k) & 1u) ? fmax(v[k], o) : v[k] + o;
    }
  }
}

__device__ __forceinline__ double load_agent(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Returns true in thread 0 of the last block to finish, with v[] holding the
// grid-wide result.  Must be reached by every thread of every block.
template <int NV, unsigned MAXMASK>
__device__ bool reduce_last(double (&v)[NV], double *part, unsigned *counter) {
  __shared__ double sm[(NT / 64) * NV];
  __shared__ int last;
  wave_reduce<NV, MAXMASK>(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[w * NV + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double a = sm[k];
      for (int ww = 1; ww < NT / 64; ++ww)
        a = ((MAXMASK >> k) & 1u) ? fmax(a, sm[ww * NV + k]) : a + sm[ww * NV + k];
      part[(size_t)blockIdx.x * NV + k] = a;
    }
    __threadfence();
    unsigned t = atomicAdd(counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += NT) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double x = load_agent(part + (size_t)b * NV + k);
      acc[k] = ((MAXMASK >> k) & 1u) ? fmax(acc[k], x) : acc[k] + x;
    }
  }
  wave_reduce<NV, MAXMASK>(acc);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sm[w * NV + k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x != 0) return false;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = sm[k];
    for (int ww = 1; ww < NT / 64; ++ww)
      a = ((MAXMASK >> k) & 1u) ? fmax(a, sm[ww * NV + k]) : a + sm[ww * NV + k];
    v[k] = a;
  }
  *counter = 0u;
  return true;
}

}  // namespace adg
