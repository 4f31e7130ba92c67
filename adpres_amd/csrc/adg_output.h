// Output reductions on the device: the numbers of io::AsmPow
// (src/mod_io.f90:3267-3403, axially averaged assembly powers) and
// io::AxiPow (:3407-3498, radially averaged plane powers) from the node
// powers of PowDis, for the host's printing.  Every sum runs in one thread
// in the reference's loop order (a node column over k, an assembly over its
// node rows and columns, a plane over its node planes and footprint), so
// the numbers are the reference's bit for bit for the same node powers; the
// many independent sums run in parallel.
#pragma once
#include "adg_common.h"

namespace adg {

struct MapArgs {
  const double *pw;            // node powers per cell (zero outside nodes)
  const double *hx, *hy, *hz;  // node sizes
  const double *vdel;          // node volumes per cell (zero outside nodes)
  const int *ys0, *ys1;        // (nyy) footprint of row j: i in [ys0, ys1], 0-based
  const int *xa, *ya, *za;     // (nx+1), (ny+1), (nz+1): first node of each assembly column / row / plane
  double *fnode;               // (nxx nyy) axially averaged node power
  double *fasm;                // (nx ny)
  double *faxi;                // (nz)
  int *imax;                   // i, j of the assembly maximum (1-based), plane of the axial maximum
  long base, sy, sz;           // cell(i, j, k) = base + (k + 1) sz + (j + 1) sy + i + 1
  int nxx, nyy, nzz, nx, ny, nz;
  // the flux map (io::AsmFlux): flux per cell with group stride NC
  const double *f0;
  long NC;
  int G;
  double norm;                 // <= 0: absent (no normalisation)
  int *negf;                   // (G) 1 where an assembly flux of the group is negative
};

__device__ __forceinline__ long map_cell(const MapArgs &a, int i, int j, int k) {
  return a.base + (long)(k + 1) * a.sz + (long)(j + 1) * a.sy + i + 1;
}

// AsmPow's axially averaged node-wise distribution (:3298-3309)
__global__ void k_map_fnode(MapArgs a) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nxx * a.nyy) return;
  const int i = t % a.nxx, j = t / a.nxx;
  double f = 0.0;
  if (i >= a.ys0[j] && i <= a.ys1[j]) {
    double summ = 0.0, vsumm = 0.0;
    for (int k = 0; k < a.nzz; ++k) {
      summ = summ + a.pw[map_cell(a, i, j, k)] * a.hz[k];
      vsumm = vsumm + a.hz[k];
    }
    f = summ / vsumm;
  }
  a.fnode[t] = f;
}

// assembly power (:3312-3336) and plane power (:3436-3456), before normalisation
__global__ void k_map_sums(MapArgs a) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < a.nx * a.ny) {
    const int i = t % a.nx, j = t / a.nx;
    double summ = 0.0, vsumm = 0.0;
    for (int ly = a.ya[j]; ly < a.ya[j + 1]; ++ly)
      for (int lx = a.xa[i]; lx < a.xa[i + 1]; ++lx) {
        summ = summ + a.fnode[lx + (long)a.nxx * ly] * a.hx[lx] * a.hy[ly];
        vsumm = vsumm + a.hx[lx] * a.hy[ly];
      }
    a.fasm[t] = summ / vsumm;
  } else if (t < a.nx * a.ny + a.nz) {
    const int k = t - a.nx * a.ny;
    double summ = 0.0, vsumm = 0.0;
    for (int z = a.za[k]; z < a.za[k + 1]; ++z)
      for (int j = 0; j < a.nyy; ++j)
        for (int i = a.ys0[j]; i <= a.ys1[j]; ++i) {
          const long c = map_cell(a, i, j, z);
          summ = summ + a.pw[c];
          vsumm = vsumm + a.vdel[c];
        }
    a.faxi[k] = summ / vsumm;
  }
}

// the counts, totals, normalisations and maxima (:3330-3353, :3451-3467):
// one thread, in the reference's order
__global__ void k_map_final(MapArgs a) {
#pragma clang fp contract(off)
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int nfuel = 0;
  double totp = 0.0;
  for (int t = 0; t < a.nx * a.ny; ++t)
    if (a.fasm[t] > 0.0) {
      nfuel++;
      totp = totp + a.fasm[t];
    }
  int xm = 1, ym = 1;
  double fmax = 0.0;
  for (int t = 0; t < a.nx * a.ny; ++t) {
    if (totp > 0.0) a.fasm[t] = (double)nfuel / totp * a.fasm[t];
    if (a.fasm[t] > fmax) {
      xm = t % a.nx + 1;
      ym = t / a.nx + 1;
      fmax = a.fasm[t];
    }
  }
  nfuel = 0;
  totp = 0.0;
  for (int k = 0; k < a.nz; ++k)
    if (a.faxi[k] > 0.0) {
      nfuel++;
      totp = totp + a.faxi[k];
    }
  int km = 1;
  fmax = 0.0;
  for (int k = 0; k < a.nz; ++k) {
    a.faxi[k] = (double)nfuel / totp * a.faxi[k];
    if (a.faxi[k] > fmax) {
      km = k + 1;
      fmax = a.faxi[k];
    }
  }
  a.imax[0] = xm;
  a.imax[1] = ym;
  a.imax[2] = km;
}

// io::AsmFlux (src/mod_io.f90:3502-3644): the same reductions per group on
// the flux, with the node volume weight of its axial average
__global__ void k_flux_fnode(MapArgs a) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nxx * a.nyy * a.G) return;
  const int g = t / (a.nxx * a.nyy), r = t - g * a.nxx * a.nyy, i = r % a.nxx, j = r / a.nxx;
  double f = 0.0;
  if (i >= a.ys0[j] && i <= a.ys1[j]) {
    double summ = 0.0, vsumm = 0.0;
    for (int k = 0; k < a.nzz; ++k) {
      summ = summ + a.f0[g * a.NC + map_cell(a, i, j, k)] * a.hx[i] * a.hy[j] * a.hz[k];
      vsumm = vsumm + a.hx[i] * a.hy[j] * a.hz[k];
    }
    f = summ / vsumm;
  }
  a.fnode[t] = f;
}

__global__ void k_flux_asm(MapArgs a) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.nx * a.ny * a.G) return;
  const int g = t / (a.nx * a.ny), r = t - g * a.nx * a.ny, i = r % a.nx, j = r / a.nx;
  const double *fn = a.fnode + (long)g * a.nxx * a.nyy;
  double summ = 0.0, vsumm = 0.0;
  for (int ly = a.ya[j]; ly < a.ya[j + 1]; ++ly)
    for (int lx = a.xa[i]; lx < a.xa[i + 1]; ++lx) {
      summ = summ + fn[lx + (long)a.nxx * ly] * a.hx[lx] * a.hy[ly];
      vsumm = vsumm + a.hx[lx] * a.hy[ly];
    }
  a.fasm[t] = summ / vsumm;
}

// one thread per group: the positive total in the reference's order, the
// negative-flux flag and the normalisation norm / totp * f * norm
__global__ void k_flux_final(MapArgs a) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G) return;
  double *f = a.fasm + (long)g * a.nx * a.ny;
  double totp = 0.0;
  int neg = 0;
  for (int t = 0; t < a.nx * a.ny; ++t) {
    if (f[t] > 0.0) totp = totp + f[t];
    if (f[t] < 0.0) neg = 1;
  }
  a.negf[g] = neg;
  if (a.norm > 0.0)
    for (int t = 0; t < a.nx * a.ny; ++t) f[t] = a.norm / totp * f[t] * a.norm;
}

}  // namespace adg
