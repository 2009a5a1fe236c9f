// Preconditioned CG pressure solve with an FP32 geometric-multigrid V-cycle
// as preconditioner (gfx950).  Counterpart of the reference's Hypre branch
// poisson::Solve with Solver = 2 (PCG, poisson.cpp:115-172, 219-248): the same
// 7-point operator div(C grad P) with C = 1/Rho_face, the same two-norm
// relative-residual test at epsilon, at most Iteration iterations; BoomerAMG
// is replaced by the V-cycle, the singular row pin of poisson.cpp:138-145 by
// zero-mean preconditioned residuals.
//
// Precision: the outer iteration (solution, residual, search direction,
// operator, dot products) is FP64; the V-cycle (levels, coefficients,
// red-black Gauss-Seidel, restriction, prolongation) is FP32.  Every
// reduction uses one fixed blocked order, so results do not depend on the
// launch and the CPU oracle reproduces them bit for bit.
//
// Levels hold interior cells only (no ghosts): periodic neighbours by index
// wrap; on the duct the x end faces carry a zero coefficient.
//
// Learned forms (unet.hip; no reference counterpart): mode 1 moves the
// initial iterate along the network's pressure increment d by the step
// (r, d) / (d, S d) that minimises the S-norm error along d (so the error
// never grows); mode 2 also preconditions with z = a N(r) + V(r - a S N(r)),
// a the same optimal step (network, then one V-cycle on what it leaves),
// inside flexible PCG (Polak-Ribiere beta), and falls back to the plain
// V-cycle for the rest of the solve when an iteration reduces the residual
// norm by less than 3x.
// With a pair store attached, every solve records (input, increment) pairs.
#include <cmath>
#include <vector>

#include "fr_internal.hpp"

namespace fr {

constexpr int DNB = 1024, DNT = 256;  // reduction: chunks, threads per chunk

// (i, j, k) of linear index t = (k*ny + j)*nx + i; 32-bit arithmetic (every
// grid here has fewer than 2^31 cells), much cheaper than 64-bit division.
__device__ __forceinline__ void dec3(long t, int nx, int ny, int& i, int& j, int& k) {
  const unsigned u = (unsigned)t, q = u / (unsigned)nx;
  i = (int)(u - q * (unsigned)nx);
  j = (int)(q % (unsigned)ny);
  k = (int)(q / (unsigned)ny);
}

struct MGL {
  int nx, ny, nz;
  float *cx, *cy, *cz, *e, *r, *t;
  __device__ __forceinline__ long at(int i, int j, int k) const {
    return ((long)k * ny + j) * nx + i;
  }
  long n() const { return (long)nx * ny * nz; }
};

__global__ void k_mg_coef0(Dev d, int duct, const double* __restrict__ rfx,
                           const double* __restrict__ rfy, const double* __restrict__ rfz, MGL v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)v.nx * v.ny * v.nz) return;
  int i, j, k;
  dec3(t, v.nx, v.ny, i, j, k);
  const long c = at(d, i + 1, j + 1, k + 1);
  v.cx[t] = (duct && i == 0) ? 0.f : (float)((1. / rfx[c]) * d.idx2);
  v.cy[t] = (float)((1. / rfy[c]) * d.idy2);
  v.cz[t] = (float)((1. / rfz[c]) * d.idz2);
}

// coarse face = sum of the four fine faces it covers / 16
__global__ void k_mg_coarsen(MGL a, MGL c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)c.nx * c.ny * c.nz) return;
  int i, j, k;
  dec3(t, c.nx, c.ny, i, j, k);
  const int fi = 2 * i, fj = 2 * j, fk = 2 * k;
  c.cx[t] = (((a.cx[a.at(fi, fj, fk)] + a.cx[a.at(fi, fj + 1, fk)]) + a.cx[a.at(fi, fj, fk + 1)]) +
             a.cx[a.at(fi, fj + 1, fk + 1)]) * 0.0625f;
  c.cy[t] = (((a.cy[a.at(fi, fj, fk)] + a.cy[a.at(fi + 1, fj, fk)]) + a.cy[a.at(fi, fj, fk + 1)]) +
             a.cy[a.at(fi + 1, fj, fk + 1)]) * 0.0625f;
  c.cz[t] = (((a.cz[a.at(fi, fj, fk)] + a.cz[a.at(fi + 1, fj, fk)]) + a.cz[a.at(fi, fj + 1, fk)]) +
             a.cz[a.at(fi + 1, fj + 1, fk)]) * 0.0625f;
}

// Face coefficients (-x,+x,-y,+y,-z,+z) and neighbour indices of cell (i,j,k).
__device__ __forceinline__ void mg_nbrs(const MGL& v, int duct, int i, int j, int k, float c[6],
                                        long q[6]) {
  const int ip = i + 1 == v.nx ? 0 : i + 1, im = i == 0 ? v.nx - 1 : i - 1;
  const int jp = j + 1 == v.ny ? 0 : j + 1, jm = j == 0 ? v.ny - 1 : j - 1;
  const int kp = k + 1 == v.nz ? 0 : k + 1, km = k == 0 ? v.nz - 1 : k - 1;
  q[0] = v.at(im, j, k);
  q[1] = v.at(ip, j, k);
  q[2] = v.at(i, jm, k);
  q[3] = v.at(i, jp, k);
  q[4] = v.at(i, j, km);
  q[5] = v.at(i, j, kp);
  const long s = v.at(i, j, k);
  c[0] = v.cx[s];
  c[1] = (duct && i + 1 == v.nx) ? 0.f : v.cx[q[1]];
  c[2] = v.cy[s];
  c[3] = v.cy[q[3]];
  c[4] = v.cz[s];
  c[5] = v.cz[q[5]];
}

// red-black half sweep of S e = r: e_c = (r_c + sum C_f e_f) / sum C_f
__global__ void k_mg_smooth(MGL v, int duct, int color) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)v.nx * v.ny * v.nz) return;
  int i, j, k;
  dec3(t, v.nx, v.ny, i, j, k);
  if (((i + j + k) & 1) != color) return;
  float c[6];
  long q[6];
  mg_nbrs(v, duct, i, j, k, c, q);
  float num = v.r[t];
  float den = 0.f;
#pragma unroll
  for (int f = 0; f < 6; f++) {
    num = num + c[f] * v.e[q[f]];
    den = den + c[f];
  }
  v.e[t] = num / den;
}

// t = r - S e
__global__ void k_mg_residual(MGL v, int duct) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)v.nx * v.ny * v.nz) return;
  int i, j, k;
  dec3(t, v.nx, v.ny, i, j, k);
  float c[6];
  long q[6];
  mg_nbrs(v, duct, i, j, k, c, q);
  const float ec = v.e[t];
  float se = 0.f;
#pragma unroll
  for (int f = 0; f < 6; f++) se = se + c[f] * (ec - v.e[q[f]]);
  v.t[t] = v.r[t] - se;
}

// coarse r = mean of the 8 fine residuals (order dk, dj, di); coarse e = 0
__global__ void k_mg_restrict(MGL f, MGL c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)c.nx * c.ny * c.nz) return;
  int i, j, k;
  dec3(t, c.nx, c.ny, i, j, k);
  float a = 0.f;
#pragma unroll
  for (int dk = 0; dk < 2; dk++)
#pragma unroll
    for (int dj = 0; dj < 2; dj++)
#pragma unroll
      for (int di = 0; di < 2; di++) a = a + f.t[f.at(2 * i + di, 2 * j + dj, 2 * k + dk)];
  c.r[t] = a * 0.125f;
  c.e[t] = 0.f;
}

// fine e += coarse e of the parent cell (piecewise constant)
__global__ void k_mg_prolong(MGL f, MGL c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)f.nx * f.ny * f.nz) return;
  int i, j, k;
  dec3(t, f.nx, f.ny, i, j, k);
  f.e[t] = f.e[t] + c.e[c.at(i / 2, j / 2, k / 2)];
}

// The tail of the V-cycle from level l0 down, for small levels, in one
// workgroup: the same per-cell operations as the per-level kernels, with a
// workgroup barrier where those have a launch boundary.
constexpr int MGMAX = 16, SMALL = 1024;
struct MGSet {
  MGL v[MGMAX];
  int nl;
};

__device__ __forceinline__ void sm_smooth(const MGL& v, int duct, int color) {
  const long n = (long)v.nx * v.ny * v.nz;
  for (long t = threadIdx.x; t < n; t += SMALL) {
    int i, j, k;
  dec3(t, v.nx, v.ny, i, j, k);
    if (((i + j + k) & 1) != color) continue;
    float c[6];
    long q[6];
    mg_nbrs(v, duct, i, j, k, c, q);
    float num = v.r[t];
    float den = 0.f;
#pragma unroll
    for (int f = 0; f < 6; f++) {
      num = num + c[f] * v.e[q[f]];
      den = den + c[f];
    }
    v.e[t] = num / den;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(SMALL) k_mg_small(MGSet S, int l0, int duct, int pre, int post,
                                                    int csweeps) {
  const int last = S.nl - 1;
  for (int l = l0; l < last; l++) {
    const MGL& v = S.v[l];
    for (int k = 0; k < pre; k++) {
      sm_smooth(v, duct, 0);
      sm_smooth(v, duct, 1);
    }
    const long n = (long)v.nx * v.ny * v.nz;
    for (long t = threadIdx.x; t < n; t += SMALL) {
      int i, j, k;
  dec3(t, v.nx, v.ny, i, j, k);
      float c[6];
      long q[6];
      mg_nbrs(v, duct, i, j, k, c, q);
      const float ec = v.e[t];
      float se = 0.f;
#pragma unroll
      for (int f = 0; f < 6; f++) se = se + c[f] * (ec - v.e[q[f]]);
      v.t[t] = v.r[t] - se;
    }
    __syncthreads();
    const MGL& c = S.v[l + 1];
    const long nc = (long)c.nx * c.ny * c.nz;
    for (long t = threadIdx.x; t < nc; t += SMALL) {
      int i, j, k;
  dec3(t, c.nx, c.ny, i, j, k);
      float a = 0.f;
      for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
          for (int di = 0; di < 2; di++) a = a + v.t[v.at(2 * i + di, 2 * j + dj, 2 * k + dk)];
      c.r[t] = a * 0.125f;
      c.e[t] = 0.f;
    }
    __syncthreads();
  }
  const MGL& cv = S.v[last];
  for (int k = 0; k < csweeps; k++) {
    sm_smooth(cv, duct, 0);
    sm_smooth(cv, duct, 1);
  }
  for (int k = 0; k < csweeps; k++) {
    sm_smooth(cv, duct, 1);
    sm_smooth(cv, duct, 0);
  }
  for (int l = last - 1; l >= l0; l--) {
    const MGL& v = S.v[l];
    const MGL& c = S.v[l + 1];
    const long n = (long)v.nx * v.ny * v.nz;
    for (long t = threadIdx.x; t < n; t += SMALL) {
      int i, j, k;
  dec3(t, v.nx, v.ny, i, j, k);
      v.e[t] = v.e[t] + c.e[c.at(i / 2, j / 2, k / 2)];
    }
    __syncthreads();
    for (int k = 0; k < post; k++) {
      sm_smooth(v, duct, 1);
      sm_smooth(v, duct, 0);
    }
  }
}

__global__ void k_fill_f(long n, float* __restrict__ a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.f;
}

// ---- fine level, double, ghosted fields --------------------------------------
__device__ __forceinline__ void tijk(long t, int nx, int ny, int& i, int& j, int& k) {
  dec3(t, nx, ny, i, j, k);
  i++;
  j++;
  k++;
}

// S p = -div(C grad p), C = 1./Rho_face, with the operation order of the
// defect-correction residual (poisson.cpp:393-397).
__device__ __forceinline__ double sfine(const Dev& d, const double* __restrict__ p,
                                        const double* __restrict__ rfx,
                                        const double* __restrict__ rfy,
                                        const double* __restrict__ rfz, long c) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  const double Cx0 = 1. / rfx[c], Cx1 = 1. / rfx[c + sx];
  const double Cy0 = 1. / rfy[c], Cy1 = 1. / rfy[c + sy];
  const double Cz0 = 1. / rfz[c], Cz1 = 1. / rfz[c + sz];
  const double pc = p[c];
  const double D = (((p[c + sx] - pc) * d.idx) * Cx1 - ((pc - p[c - sx]) * d.idx) * Cx0) * d.idx +
                   (((p[c + sy] - pc) * d.idy) * Cy1 - ((pc - p[c - sy]) * d.idy) * Cy0) * d.idy +
                   (((p[c + sz] - pc) * d.idz) * Cz1 - ((pc - p[c - sz]) * d.idz) * Cz0) * d.idz;
  return -D;
}

struct Fine {
  Dev d;
  int nx, ny, nz;
  long n;
};

// op 0: r = -RHS - S x; op 1: q = S p
__global__ void k_pcg_apply(Fine F, int op, const double* __restrict__ x,
                            const double* __restrict__ rhs, const double* __restrict__ rfx,
                            const double* __restrict__ rfy, const double* __restrict__ rfz,
                            double* __restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  const long c = at(F.d, i, j, k);
  const double s = sfine(F.d, x, rfx, rfy, rfz, c);
  out[c] = op == 0 ? -rhs[c] - s : s;
}

// Blocked sums over the interior.  mode 0: a.b (ghosted doubles); mode 1:
// sum of the float level-0 array e.
template <int MODE>
__global__ void __launch_bounds__(DNT) k_dsum(Fine F, const double* __restrict__ a,
                                              const double* __restrict__ b,
                                              const float* __restrict__ e,
                                              double* __restrict__ part) {
  __shared__ double sh[DNT];
  const long chunk = (F.n + DNB - 1) / DNB;
  const long lo = (long)blockIdx.x * chunk;
  const long hi = lo + chunk < F.n ? lo + chunk : F.n;
  double acc = 0.0;
  for (long t = lo + threadIdx.x; t < hi; t += DNT) {
    if (MODE == 0) {
      int i, j, k;
      tijk(t, F.nx, F.ny, i, j, k);
      const long c = at(F.d, i, j, k);
      acc += a[c] * b[c];
    } else {
      acc += (double)e[t];
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = DNT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// chunk sums combined by a pairwise tree (DNB/2 threads)
__global__ void __launch_bounds__(DNB / 2) k_dsum_final(const double* __restrict__ part,
                                                        double* __restrict__ out) {
  __shared__ double sh[DNB];
  sh[threadIdx.x] = part[threadIdx.x];
  sh[threadIdx.x + DNB / 2] = part[threadIdx.x + DNB / 2];
  __syncthreads();
  for (int s = DNB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// level-0 r = (float) residual
__global__ void k_pcg_tofloat(Fine F, const double* __restrict__ r, float* __restrict__ f) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  f[t] = (float)r[at(F.d, i, j, k)];
}

// z = (double) e - mean, mean = *msum / n
__global__ void k_pcg_todouble(Fine F, const float* __restrict__ e, const double* __restrict__ msum,
                               double* __restrict__ z) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  const double m = *msum / (double)F.n;
  z[at(F.d, i, j, k)] = (double)e[t] - m;
}

// z (+)= scale * (y - mean), y in interior order, mean = *msum / n
__global__ void k_pcg_netz(Fine F, const float* __restrict__ y, const double* __restrict__ msum,
                           double scale, int add, double* __restrict__ z) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  const long c = at(F.d, i, j, k);
  const double v = scale * ((double)y[t] - *msum / (double)F.n);
  z[c] = add ? z[c] + v : v;
}

// op 0: a = b - s a; op 1: a = a + s b (interior)
__global__ void k_pcg_vec(Fine F, int op, double s, double* __restrict__ a,
                          const double* __restrict__ b) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  const long c = at(F.d, i, j, k);
  a[c] = op == 0 ? b[c] - s * a[c] : a[c] + s * b[c];
}

// op 0: p = z; op 1: x += alpha p, r -= alpha q; op 2: p = z + beta p
__global__ void k_pcg_axpy(Fine F, int op, double s, double* __restrict__ x,
                           double* __restrict__ r, double* __restrict__ p,
                           const double* __restrict__ q) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  const long c = at(F.d, i, j, k);
  if (op == 0) {
    p[c] = q[c];
  } else if (op == 1) {
    x[c] = x[c] + s * p[c];
    r[c] = r[c] - s * q[c];
  } else {
    p[c] = q[c] + s * p[c];
  }
}

// k_pcg_axpy with the coefficient formed on the device from dot products in
// device memory: s = (num - sub) / den (sub optional), or 0 when zero is set
// (the host's alpha = rz/pq, beta = rz2/rz or (rz2 - zro)/rz, in double).
__global__ void k_pcg_axpy_d(Fine F, int op, const double* __restrict__ num,
                             const double* __restrict__ den, const double* __restrict__ sub,
                             int zero, double* __restrict__ x, double* __restrict__ r,
                             double* __restrict__ p, const double* __restrict__ q) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.n) return;
  const double s = zero ? 0.0 : (sub ? (num[0] - sub[0]) : num[0]) / den[0];
  int i, j, k;
  tijk(t, F.nx, F.ny, i, j, k);
  const long c = at(F.d, i, j, k);
  if (op == 1) {
    x[c] = x[c] + s * p[c];
    r[c] = r[c] - s * q[c];
  } else {
    p[c] = q[c] + s * p[c];
  }
}

__global__ void k_neumann_x2(Dev d, double* __restrict__ P) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)d.Ny * d.Nz) return;
  const long base = t * d.Nx;
  P[base] = P[base + 1];
  P[base + d.Nx - 1] = P[base + d.Nx - 2];
}

}  // namespace fr

using namespace fr;

struct fr_pcg {
  fr_geom g;
  Fine F;
  int duct = 0;
  std::vector<MGL> L;
  double *r = nullptr, *z = nullptr, *p = nullptr;  // ghosted fine fields
  double* part = nullptr;
  // scalar results: [0] norms read by the host, [1] the V-cycle's sum, [2]
  // (p, S p), [3] and [4] (r, z) of the current and next iteration, [5]
  // (r_old, z)
  double* dres = nullptr;
  double* hres = nullptr;   // pinned host copy
  int pre = 2, post = 2, coarse_sweeps = 20;
  int l_small = -1;  // first level handled by the one-workgroup kernel
  // learned forms
  fr_unet* net = nullptr;
  int net_mode = 0;
  fr_unet_data* collect = nullptr;
  double *p0 = nullptr, *ta = nullptr, *tb = nullptr, *rold = nullptr;  // ghosted
  int fallbacks = 0, guesses_kept = 0;
  double guess_gain = 0.0;  // sum over kept guesses of log10(rr before / rr after)
};

static int mg_launch_n(long n) { return (int)((n + 255) / 256); }

static int fill_fine_ghosts(fr_pcg* s, double* f, hipStream_t st) {
  int rc = ghost_fill(s->F.d, f, st);
  if (rc) return rc;
  if (s->duct) {
    hipLaunchKernelGGL(k_neumann_x2, dim3(nblocks((long)s->F.d.Ny * s->F.d.Nz, 256)), dim3(256),
                       0, st, s->F.d, f);
    FR_LAUNCH_CHECK();
  }
  return FR_OK;
}

static void smooth_pair(fr_pcg* s, const MGL& v, int first, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_smooth, dim3(mg_launch_n(v.n())), dim3(256), 0, st, v, s->duct, first);
  hipLaunchKernelGGL(k_mg_smooth, dim3(mg_launch_n(v.n())), dim3(256), 0, st, v, s->duct,
                     1 - first);
}

static void vcycle(fr_pcg* s, size_t l, hipStream_t st) {
  const MGL& v = s->L[l];
  if ((int)l == s->l_small) {
    MGSet S;
    S.nl = (int)s->L.size();
    for (int q = 0; q < S.nl; q++) S.v[q] = s->L[q];
    hipLaunchKernelGGL(k_mg_small, dim3(1), dim3(SMALL), 0, st, S, (int)l, s->duct, s->pre,
                       s->post, s->coarse_sweeps);
    return;
  }
  if (l + 1 == s->L.size()) {
    for (int k = 0; k < s->coarse_sweeps; k++) smooth_pair(s, v, 0, st);
    for (int k = 0; k < s->coarse_sweeps; k++) smooth_pair(s, v, 1, st);
    return;
  }
  for (int k = 0; k < s->pre; k++) smooth_pair(s, v, 0, st);
  hipLaunchKernelGGL(k_mg_residual, dim3(mg_launch_n(v.n())), dim3(256), 0, st, v, s->duct);
  const MGL& c = s->L[l + 1];
  hipLaunchKernelGGL(k_mg_restrict, dim3(mg_launch_n(c.n())), dim3(256), 0, st, v, c);
  vcycle(s, l + 1, st);
  hipLaunchKernelGGL(k_mg_prolong, dim3(mg_launch_n(v.n())), dim3(256), 0, st, v, c);
  for (int k = 0; k < s->post; k++) smooth_pair(s, v, 1, st);
}

// blocked sum into s->dres[slot]
static void dsum(fr_pcg* s, int mode, const double* a, const double* b, const float* e, int slot,
                 hipStream_t st) {
  if (mode == 0)
    hipLaunchKernelGGL(k_dsum<0>, dim3(DNB), dim3(DNT), 0, st, s->F, a, b, e, s->part);
  else
    hipLaunchKernelGGL(k_dsum<1>, dim3(DNB), dim3(DNT), 0, st, s->F, a, b, e, s->part);
  hipLaunchKernelGGL(k_dsum_final, dim3(1), dim3(DNB / 2), 0, st, s->part, s->dres + slot);
}

static int read_scalar(fr_pcg* s, int slot, double* v, hipStream_t st) {
  FR_LAUNCH_CHECK();
  FR_CHECK(hipMemcpyAsync(s->hres, s->dres + slot, sizeof(double), hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  *v = s->hres[0];
  return FR_OK;
}

// z = M r: r to float, one V-cycle, zero-mean z in double
static int precond(fr_pcg* s, hipStream_t st) {
  const MGL& f = s->L[0];
  const unsigned nb = mg_launch_n(s->F.n);
  hipLaunchKernelGGL(k_pcg_tofloat, dim3(nb), dim3(256), 0, st, s->F, (const double*)s->r, f.r);
  hipLaunchKernelGGL(k_fill_f, dim3(nb), dim3(256), 0, st, f.n(), f.e);
  vcycle(s, 0, st);
  dsum(s, 1, nullptr, nullptr, f.e, 1, st);
  hipLaunchKernelGGL(k_pcg_todouble, dim3(nb), dim3(256), 0, st, s->F, (const float*)f.e,
                     (const double*)(s->dres + 1), s->z);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

namespace fr {

static int ensure_buf(fr_pcg* s, double** p) {
  if (*p) return FR_OK;
  const size_t gs = (size_t)s->g.Nx * s->g.Ny * s->g.Nz;
  FR_CHECK(hipMalloc(p, sizeof(double) * gs));
  FR_CHECK(hipMemset(*p, 0, sizeof(double) * gs));
  return FR_OK;
}

// Optimal step along d with q = S d: (r, d) / (d, q), 0 if not a descent.
static int opt_step(fr_pcg* s, const double* d, const double* q, double* a, hipStream_t st) {
  double rd, dq;
  dsum(s, 0, s->r, d, nullptr, 0, st);
  FR_RET(read_scalar(s, 0, &rd, st));
  dsum(s, 0, d, q, nullptr, 0, st);
  FR_RET(read_scalar(s, 0, &dq, st));
  *a = (dq > 0 && rd > 0) ? rd / dq : 0.0;
  return FR_OK;
}

// z = a N(r) + V(r - a S N(r)), N(r) = scale * (net(-r) - mean).
static int precond_net(fr_pcg* s, const double* rfx, const double* rfy, const double* rfz,
                       double rrms, hipStream_t st) {
  const Fine& F = s->F;
  const unsigned nb = mg_launch_n(F.n);
  const float* y;
  double sc, a;
  FR_RET(unet_infer(s->net, F.d, s->r, -1.0, rrms, rfx, rfy, rfz, nullptr, &y, &sc, st));
  dsum(s, 1, nullptr, nullptr, y, 1, st);
  hipLaunchKernelGGL(k_pcg_netz, dim3(nb), dim3(256), 0, st, F, y, (const double*)(s->dres + 1),
                     sc, 0, s->tb);
  FR_RET(fill_fine_ghosts(s, s->tb, st));
  hipLaunchKernelGGL(k_pcg_apply, dim3(nb), dim3(256), 0, st, F, 1, (const double*)s->tb,
                     (const double*)nullptr, rfx, rfy, rfz, s->ta);
  FR_RET(opt_step(s, s->tb, s->ta, &a, st));
  hipLaunchKernelGGL(k_pcg_vec, dim3(nb), dim3(256), 0, st, F, 0, a, s->ta, (const double*)s->r);
  std::swap(s->r, s->ta);
  const int rc = precond(s, st);
  std::swap(s->r, s->ta);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pcg_vec, dim3(nb), dim3(256), 0, st, F, 1, a, s->z, (const double*)s->tb);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int pcg_solve(fr_pcg* s, const double* rfx, const double* rfy, const double* rfz, double* P,
              const double* RHS, int maxit, double tol, int* iters, double* relres,
              hipStream_t st) {
  if (!s || !rfx || !rfy || !rfz || !P || !RHS) return FR_EINVAL;
  const Fine& F = s->F;
  const unsigned nb = mg_launch_n(F.n);
  // level coefficients
  hipLaunchKernelGGL(k_mg_coef0, dim3(nb), dim3(256), 0, st, F.d, s->duct, rfx, rfy, rfz, s->L[0]);
  for (size_t l = 1; l < s->L.size(); l++)
    hipLaunchKernelGGL(k_mg_coarsen, dim3(mg_launch_n(s->L[l].n())), dim3(256), 0, st, s->L[l - 1],
                       s->L[l]);
  int rc = fill_fine_ghosts(s, P, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pcg_apply, dim3(nb), dim3(256), 0, st, F, 0, (const double*)P, RHS, rfx, rfy,
                     rfz, s->r);
  double bn2, rr;
  dsum(s, 0, RHS, RHS, nullptr, 0, st);
  if ((rc = read_scalar(s, 0, &bn2, st))) return rc;
  const double bn = std::sqrt(bn2);
  dsum(s, 0, s->r, s->r, nullptr, 0, st);
  if ((rc = read_scalar(s, 0, &rr, st))) return rc;
  int it = 0;
  const size_t gbytes = sizeof(double) * (size_t)s->g.Nx * s->g.Ny * s->g.Nz;
  const bool todo = std::sqrt(rr) > tol * bn;
  // pair collection: the input at entry, the increment at exit
  double cscale = 0;
  const bool collecting = todo && s->collect && unet_data_room(s->collect);
  if (collecting) {
    if ((rc = ensure_buf(s, &s->p0))) return rc;
    FR_CHECK(hipMemcpyAsync(s->p0, P, gbytes, hipMemcpyDeviceToDevice, st));
    if ((rc = unet_data_begin(s->collect, F.d, s->r, -1.0, std::sqrt(rr / F.n), rfx, rfy, rfz, P,
                              &cscale, st)))
      return rc;
  }
  bool flex = false;
  if (todo && s->net && s->net_mode >= 1) {
    // learned initial iterate: P += a d along the network's increment d
    const float* y;
    double sc, a;
    if ((rc = unet_infer(s->net, F.d, s->r, -1.0, std::sqrt(rr / F.n), rfx, rfy, rfz, P, &y, &sc,
                         st)))
      return rc;
    dsum(s, 1, nullptr, nullptr, y, 1, st);
    hipLaunchKernelGGL(k_pcg_netz, dim3(nb), dim3(256), 0, st, F, y, (const double*)(s->dres + 1),
                       sc, 0, s->p);
    if ((rc = fill_fine_ghosts(s, s->p, st))) return rc;
    hipLaunchKernelGGL(k_pcg_apply, dim3(nb), dim3(256), 0, st, F, 1, (const double*)s->p, RHS,
                       rfx, rfy, rfz, s->z);
    if ((rc = opt_step(s, s->p, s->z, &a, st))) return rc;
    if (a > 0) {
      hipLaunchKernelGGL(k_pcg_axpy, dim3(nb), dim3(256), 0, st, F, 1, a, P, s->r, s->p,
                         (const double*)s->z);
      const double rr0 = rr;
      dsum(s, 0, s->r, s->r, nullptr, 0, st);
      if ((rc = read_scalar(s, 0, &rr, st))) return rc;
      s->guesses_kept++;
      if (rr > 0) s->guess_gain += std::log10(rr0 / rr);
    }
    if (s->net_mode >= 2) {
      for (double** b : {&s->ta, &s->tb, &s->rold})
        if ((rc = ensure_buf(s, b))) return rc;
      flex = true;
    }
  }
  if (std::sqrt(rr) > tol * bn) {
    if ((rc = flex ? precond_net(s, rfx, rfy, rfz, std::sqrt(rr / F.n), st) : precond(s, st)))
      return rc;
    hipLaunchKernelGGL(k_pcg_axpy, dim3(nb), dim3(256), 0, st, F, 0, 0.0, P, s->r, s->p,
                       (const double*)s->z);
    // (r, z) stays on the device (slots 3 and 4 alternate); alpha and beta
    // are formed there, so the host reads only (r, r) for the stopping test
    int rz_slot = 3;
    dsum(s, 0, s->r, s->z, nullptr, rz_slot, st);
    while (it < maxit) {
      if ((rc = fill_fine_ghosts(s, s->p, st))) return rc;
      // q = S p in z's storage
      hipLaunchKernelGGL(k_pcg_apply, dim3(nb), dim3(256), 0, st, F, 1, (const double*)s->p, RHS,
                         rfx, rfy, rfz, s->z);
      dsum(s, 0, s->p, s->z, nullptr, 2, st);  // pq
      if (flex) FR_CHECK(hipMemcpyAsync(s->rold, s->r, gbytes, hipMemcpyDeviceToDevice, st));
      // alpha = rz / pq
      hipLaunchKernelGGL(k_pcg_axpy_d, dim3(nb), dim3(256), 0, st, F, 1,
                         (const double*)(s->dres + rz_slot), (const double*)(s->dres + 2),
                         (const double*)nullptr, 0, P, s->r, s->p, (const double*)s->z);
      it++;
      const double rr_prev = rr;
      dsum(s, 0, s->r, s->r, nullptr, 0, st);
      if ((rc = read_scalar(s, 0, &rr, st))) return rc;
      if (std::sqrt(rr) <= tol * bn) break;
      bool restart = false;
      if (flex && rr > rr_prev / 9.0) {  // the network stopped paying: plain V-cycles
        flex = false;
        restart = true;
        s->fallbacks++;
      }
      if ((rc = flex ? precond_net(s, rfx, rfy, rfz, std::sqrt(rr / F.n), st) : precond(s, st)))
        return rc;
      const int rz2_slot = 7 - rz_slot;
      dsum(s, 0, s->r, s->z, nullptr, rz2_slot, st);
      // beta = rz2 / rz; Polak-Ribiere in the flexible form:
      // (z, r_new - r_old) / rz = (rz2 - (r_old, z)) / rz; 0 at a restart
      if (flex) dsum(s, 0, s->rold, s->z, nullptr, 5, st);
      hipLaunchKernelGGL(k_pcg_axpy_d, dim3(nb), dim3(256), 0, st, F, 2,
                         (const double*)(s->dres + rz2_slot), (const double*)(s->dres + rz_slot),
                         flex ? (const double*)(s->dres + 5) : (const double*)nullptr,
                         restart ? 1 : 0, P, s->r, s->p, (const double*)s->z);
      rz_slot = rz2_slot;
    }
  }
  if (collecting) {
    if ((rc = fill_fine_ghosts(s, P, st))) return rc;
    if ((rc = unet_data_end(s->collect, F.d, s->p0, P, cscale, st))) return rc;
  }
  if ((rc = fill_fine_ghosts(s, P, st))) return rc;
  FR_LAUNCH_CHECK();
  if (iters) *iters = it;
  if (relres) *relres = bn > 0 ? std::sqrt(rr) / bn : 0.0;
  return FR_OK;
}

}  // namespace fr

extern "C" {

int fr_pcg_create(const fr_geom* g, fr_pcg** out) {
  if (!geom_ok(g) || !out) return FR_EINVAL;
  fr_pcg* s = new fr_pcg();
  s->g = *g;
  s->duct = g->is_box ? 0 : 1;
  s->F.d = make_dev(g);
  s->F.nx = g->Nx - 2;
  s->F.ny = g->Ny - 2;
  s->F.nz = g->Nz - 2;
  s->F.n = (long)s->F.nx * s->F.ny * s->F.nz;
  int nx = s->F.nx, ny = s->F.ny, nz = s->F.nz;
  while (true) {
    MGL v;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    const size_t n = (size_t)nx * ny * nz;
    float** a[6] = {&v.cx, &v.cy, &v.cz, &v.e, &v.r, &v.t};
    for (auto pp : a) FR_CHECK(hipMalloc(pp, sizeof(float) * n));
    s->L.push_back(v);
    if (nx % 2 || ny % 2 || nz % 2 || nx < 8 || ny < 8 || nz < 8) break;
    nx /= 2;
    ny /= 2;
    nz /= 2;
  }
  if ((int)s->L.size() > MGMAX) return FR_EINVAL;
  for (size_t l = 0; l < s->L.size(); l++)
    if (s->L[l].n() <= 4096) {
      s->l_small = (int)l;
      break;
    }
  const size_t gs = (size_t)g->Nx * g->Ny * g->Nz;
  for (double** pp : {&s->r, &s->z, &s->p}) {
    FR_CHECK(hipMalloc(pp, sizeof(double) * gs));
    FR_CHECK(hipMemset(*pp, 0, sizeof(double) * gs));
  }
  FR_CHECK(hipMalloc(&s->part, sizeof(double) * DNB));
  FR_CHECK(hipMalloc(&s->dres, sizeof(double) * 8));
  FR_CHECK(hipHostMalloc(&s->hres, sizeof(double) * 2, hipHostMallocDefault));
  *out = s;
  return FR_OK;
}

int fr_pcg_destroy(fr_pcg* s) {
  if (!s) return FR_OK;
  for (MGL& v : s->L) {
    float* a[6] = {v.cx, v.cy, v.cz, v.e, v.r, v.t};
    for (float* p : a) (void)hipFree(p);
  }
  for (double* p : {s->r, s->z, s->p, s->p0, s->ta, s->tb, s->rold}) (void)hipFree(p);
  (void)hipFree(s->part);
  (void)hipFree(s->dres);
  (void)hipHostFree(s->hres);
  delete s;
  return FR_OK;
}

int fr_pcg_set_learned(fr_pcg* s, fr_unet* net, int mode) {
  if (!s || mode < 0 || mode > 2 || (mode > 0 && !net)) return FR_EINVAL;
  s->net = mode ? net : nullptr;
  s->net_mode = mode;
  return FR_OK;
}

int fr_pcg_set_collect(fr_pcg* s, fr_unet_data* d) {
  if (!s) return FR_EINVAL;
  s->collect = d;
  return FR_OK;
}

int fr_pcg_learned_stats(fr_pcg* s, int* guesses_kept, int* fallbacks, double* log10_gain) {
  if (!s) return FR_EINVAL;
  if (guesses_kept) *guesses_kept = s->guesses_kept;
  if (fallbacks) *fallbacks = s->fallbacks;
  if (log10_gain) *log10_gain = s->guess_gain;
  return FR_OK;
}

int fr_pcg_solve(fr_pcg* s, const double* rfx, const double* rfy, const double* rfz, double* P,
                 const double* RHS, int iterations, double epsilon, int* num_iter,
                 double* relres, void* stream) {
  return fr::pcg_solve(s, rfx, rfy, rfz, P, RHS, iterations, epsilon, num_iter, relres,
                       S(stream));
}

}  // extern "C"
