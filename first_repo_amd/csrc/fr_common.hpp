// Shared device/host helpers for the gfx950 kernels of libfr_amd.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/fr_amd.h"

#define FR_CHECK(expr)                                                     \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      fprintf(stderr, "fr_amd: %s failed: %s (%s:%d)\n", #expr,            \
              hipGetErrorString(e_), __FILE__, __LINE__);                  \
      return (int)e_;                                                      \
    }                                                                      \
  } while (0)

#define FR_LAUNCH_CHECK()                                                  \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess) {                                                \
      fprintf(stderr, "fr_amd: launch failed: %s (%s:%d)\n",               \
              hipGetErrorString(e_), __FILE__, __LINE__);                  \
      return (int)e_;                                                      \
    }                                                                      \
  } while (0)

// Propagates a nonzero status.
#define FR_RET(expr)       \
  do {                     \
    int rc_ = (expr);      \
    if (rc_) return rc_;   \
  } while (0)

namespace fr {

// Grid constants as the reference caches them (tensor0.cpp:18-26):
// inv_dx = 1./dx, inv_2dx = 1./(2*dx), inv_dx2 = 1./(dx*dx), computed on the
// host in double so device arithmetic sees the same operands.
struct Dev {
  int Nx, Ny, Nz;
  long NxNy;
  long size;
  int box;
  double idx, idy, idz;
  double i2dx, i2dy, i2dz;
  double idx2, idy2, idz2;
};

inline Dev make_dev(const fr_geom* g) {
  Dev d;
  d.Nx = g->Nx;
  d.Ny = g->Ny;
  d.Nz = g->Nz;
  d.NxNy = (long)g->Nx * g->Ny;
  d.size = d.NxNy * g->Nz;
  d.box = g->is_box;
  d.idx = 1. / g->dx;
  d.idy = 1. / g->dy;
  d.idz = 1. / g->dz;
  d.i2dx = 1. / (2 * g->dx);
  d.i2dy = 1. / (2 * g->dy);
  d.i2dz = 1. / (2 * g->dz);
  d.idx2 = 1. / (g->dx * g->dx);
  d.idy2 = 1. / (g->dy * g->dy);
  d.idz2 = 1. / (g->dz * g->dz);
  return d;
}

inline bool geom_ok(const fr_geom* g) {
  return g && g->bs == 1 && g->Nx >= 3 && g->Ny >= 3 && g->Nz >= 3 &&
         (long)g->Nx * g->Ny * g->Nz < (1L << 31);
}

// Periodic image of a ghosted index with one ghost layer: 0 -> N-2, N-1 -> 1.
__device__ __forceinline__ int img(int i, int N) {
  return i == 0 ? N - 2 : (i == N - 1 ? 1 : i);
}

__device__ __forceinline__ long at(const Dev& d, int i, int j, int k) {
  return (long)k * d.NxNy + (long)j * d.Nx + i;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  xcd_block() renumbers the linear workgroup id so that each XCD gets
// one contiguous run of the grid (neighbouring tiles and planes share an L2);
// a tail of fewer than 8 groups keeps its id.  Returns (x, y, z) block
// coordinates of the renumbered id.
// The same renumbering for a one-dimensional grid: the linear id of the
// workgroup's run position (particle kernels: particles stored in cell
// order gather from neighbouring cells, whose rows then share one L2).
__device__ __forceinline__ unsigned xcd_lin() {
  const unsigned nb = gridDim.x;
  unsigned lin = blockIdx.x;
  const unsigned per = nb >> 3;
  if (lin < (per << 3)) lin = (lin & 7u) * per + (lin >> 3);
  return lin;
}

__device__ __forceinline__ uint3 xcd_block() {
  const unsigned nbx = gridDim.x, nby = gridDim.y;
  const unsigned nb = nbx * nby * gridDim.z;
  unsigned lin = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
  const unsigned per = nb >> 3;
  if (lin < (per << 3)) lin = (lin & 7u) * per + (lin >> 3);
  return make_uint3(lin % nbx, (lin / nbx) % nby, lin / (nbx * nby));
}

// Renumbers a linear workgroup id (of nb) so that 2^lg consecutive logical
// ids run on one XCD at nearly the same time: hardware id b runs on XCD
// b % 8, and within each full group of 8 * 2^lg ids, b maps to
// (b % 8) * 2^lg + (b / 8) % 2^lg.  Logical ids a multiple of 8 * 2^lg apart
// keep sharing an XCD, as in plain order; a last partial group keeps its
// ids.  A bijection for every lg >= 0 (lg = 0 is the identity).
__device__ __forceinline__ unsigned xcd_group(unsigned b, unsigned lg, unsigned nb) {
  const unsigned g8 = 8u << lg;
  const unsigned base = b & ~(g8 - 1);
  if (base + g8 > nb) return b;
  return base | ((b & 7u) << lg) | ((b >> 3) & ((1u << lg) - 1));
}

inline hipStream_t S(void* s) { return (hipStream_t)s; }

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace fr
