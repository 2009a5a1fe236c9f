// tensor0/tensor1 operators on gfx950: stencils with the reference write
// ranges and single-rank halo semantics, element-wise arithmetic, cumulative
// halos and deterministic reductions.
//
// Halo semantics folded into each stencil (communicator.cpp:177-396): after
// an operator with x write range [xlo,xhi) and its Update_Ghosts, the value
// at ghosted cell (i,j,k) equals the operator evaluated at
//   box : (img(i), img(j), img(k))            (periodic in x, y, z)
//   duct: (i, img(j), img(k)) when xlo <= i < xhi, else the prior content of
//         (i, img(j), img(k))                  (y, z wrap; x ghosts kept)
// so every kernel writes the whole ghosted array in one pass and reads only
// cells the reference reads.
#include <mutex>

#include "fr_internal.hpp"

namespace fr {

enum StencilOp {
  ST_DIV = 0,   // tensor0::Equal_Div_F2C  tensor0.cpp:303-310
  ST_DEL2,      // tensor0::Equal_Del2     tensor0.cpp:312-319
  ST_IX_C2F,    // tensor0.cpp:353-360
  ST_IY_C2F,    // tensor0.cpp:362-369
  ST_IZ_C2F,    // tensor0.cpp:371-378
  ST_IX_F2C,    // tensor0.cpp:381-388
  ST_IY_F2C,    // tensor0.cpp:390-397
  ST_IZ_F2C,    // tensor0.cpp:399-406
  ST_I_F2C,     // tensor0::Equal_I_F2C    tensor0.cpp:408-416
  ST_GX,        // tensor1::Equal_Grad_C2F x  tensor1.cpp:244-247
  ST_GY,        //                        y  tensor1.cpp:249-252
  ST_GZ,        //                        z  tensor1.cpp:254-257
  ST_COPY,      // Update_Ghosts alone
};

template <int OP, typename T>
__device__ __forceinline__ T stencil_at(const Dev& d, const T* __restrict__ a,
                                        const T* __restrict__ b, const T* __restrict__ c, int i,
                                        int j, int k) {
  const T idx = (T)d.idx, idy = (T)d.idy, idz = (T)d.idz;
  const T idx2 = (T)d.idx2, idy2 = (T)d.idy2, idz2 = (T)d.idz2;
  const T h = (T)0.5;
  const long n = at(d, i, j, k);
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  if (OP == ST_DIV)
    return (a[n + sx] - a[n]) * idx + (b[n + sy] - b[n]) * idy + (c[n + sz] - c[n]) * idz;
  if (OP == ST_DEL2)
    return (a[n + sx] - 2 * a[n] + a[n - sx]) * idx2 + (a[n + sy] - 2 * a[n] + a[n - sy]) * idy2 +
           (a[n + sz] - 2 * a[n] + a[n - sz]) * idz2;
  if (OP == ST_IX_C2F) return (a[n] + a[n - sx]) * h;
  if (OP == ST_IY_C2F) return (a[n] + a[n - sy]) * h;
  if (OP == ST_IZ_C2F) return (a[n] + a[n - sz]) * h;
  if (OP == ST_IX_F2C) return (a[n] + a[n + sx]) * h;
  if (OP == ST_IY_F2C) return (a[n] + a[n + sy]) * h;
  if (OP == ST_IZ_F2C) return (a[n] + a[n + sz]) * h;
  if (OP == ST_I_F2C) {
    const T OneThird = (T)(1. / 3.);
    return OneThird * ((a[n] + a[n + sx]) * h + (b[n] + b[n + sy]) * h + (c[n] + c[n + sz]) * h);
  }
  if (OP == ST_GX) return (a[n] - a[n - sx]) * idx;
  if (OP == ST_GY) return (a[n] - a[n - sy]) * idy;
  if (OP == ST_GZ) return (a[n] - a[n - sz]) * idz;
  return (T)0;
}

template <int OP, typename T>
__global__ void __launch_bounds__(256) k_stencil(Dev d, T* __restrict__ out,
                                                 const T* __restrict__ a,
                                                 const T* __restrict__ b,
                                                 const T* __restrict__ c, int xlo, int xhi) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.Nx || j >= d.Ny) return;
  const int jj = img(j, d.Ny), kk = img(k, d.Nz);
  const int ii = d.box ? img(i, d.Nx) : i;
  T v;
  if (OP != ST_COPY && ii >= xlo && ii < xhi) {
    v = stencil_at<OP, T>(d, a, b, c, ii, jj, kk);
  } else {
    if (ii == i && jj == j && kk == k) return;
    v = out[at(d, ii, jj, kk)];
  }
  out[at(d, i, j, k)] = v;
}

static dim3 grid3(const Dev& d) { return dim3(nblocks(d.Nx, 64), nblocks(d.Ny, 4), d.Nz); }

template <typename T>
struct nodeduce {
  using type = T;
};

template <int OP, typename T>
static int launch_stencil(const fr_geom* g, T* out, const typename nodeduce<T>::type* a,
                          const typename nodeduce<T>::type* b,
                          const typename nodeduce<T>::type* c, int xlo, int xhi, void* s) {
  if (!geom_ok(g) || !out) return FR_EINVAL;
  Dev d = make_dev(g);
  hipLaunchKernelGGL((k_stencil<OP, T>), grid3(d), dim3(64, 4), 0, S(s), d, out, a, b, c, xlo,
                     xhi);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

// Ghost shell only: k faces (all i,j), j faces (interior k, all i), and for
// the box the i faces (interior j,k).  Element t of the shell of f.
template <typename T>
__device__ __forceinline__ void shell_one(const Dev& d, T* __restrict__ f, long t) {
  const long nk = 2 * d.NxNy;
  const long nj = 2L * d.Nx * (d.Nz - 2);
  const long ni = d.box ? 2L * (d.Ny - 2) * (d.Nz - 2) : 0;
  int i, j, k;
  if (t < nk) {
    k = (t < d.NxNy) ? 0 : d.Nz - 1;
    long r = t % d.NxNy;
    j = (int)(r / d.Nx);
    i = (int)(r % d.Nx);
  } else if ((t -= nk) < nj) {
    long half = (long)d.Nx * (d.Nz - 2);
    j = (t < half) ? 0 : d.Ny - 1;
    long r = t % half;
    k = 1 + (int)(r / d.Nx);
    i = (int)(r % d.Nx);
  } else if ((t -= nj) < ni) {
    long half = (long)(d.Ny - 2) * (d.Nz - 2);
    i = (t < half) ? 0 : d.Nx - 1;
    long r = t % half;
    k = 1 + (int)(r / (d.Ny - 2));
    j = 1 + (int)(r % (d.Ny - 2));
  } else {
    return;
  }
  const int ii = d.box ? img(i, d.Nx) : i;
  f[at(d, i, j, k)] = f[at(d, ii, img(j, d.Ny), img(k, d.Nz))];
}

// Up to FMAX fields per launch (blockIdx.y selects the field).
constexpr int FMAX = 6;
template <typename T>
struct FieldSet {
  T* f[FMAX];
};

template <typename T>
__global__ void k_shell(Dev d, FieldSet<T> fs) {
  shell_one(d, fs.f[blockIdx.y], (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename T>
int ghost_fill_n(const Dev& d, T* const* f, int nf, hipStream_t s) {
  if (nf < 1 || nf > FMAX) return FR_EINVAL;
  FieldSet<T> fs{};
  for (int q = 0; q < nf; q++) fs.f[q] = f[q];
  long n = 2 * d.NxNy + 2L * d.Nx * (d.Nz - 2) + (d.box ? 2L * (d.Ny - 2) * (d.Nz - 2) : 0);
  hipLaunchKernelGGL(k_shell<T>, dim3(nblocks(n, 256), nf), dim3(256), 0, s, d, fs);
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int ghost_fill_n<double>(const Dev&, double* const*, int, hipStream_t);
template int ghost_fill_n<float>(const Dev&, float* const*, int, hipStream_t);

template <typename T>
int ghost_fill(const Dev& d, T* f, hipStream_t s) {
  return ghost_fill_n(d, &f, 1, s);
}
template int ghost_fill<double>(const Dev&, double*, hipStream_t);
template int ghost_fill<float>(const Dev&, float*, hipStream_t);

// Cumulative halo on one rank (communicator.cpp:422-511 box, 520-626 duct):
// z phase (all i,j), y phase (all i, interior k), x phase (box; interior
// j,k).  Each thread reads both ghosts of its line before adding, as the
// reference packs both send buffers before unpacking.
template <typename T>
__global__ void k_cum_z(Dev d, FieldSet<T> fs) {
  T* __restrict__ f = fs.f[blockIdx.y];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.NxNy) return;
  const long top = (long)(d.Nz - 1) * d.NxNy + t, bot = t;
  const T s0 = f[top], s1 = f[bot];
  f[(long)(d.Nz - 2) * d.NxNy + t] += s1;
  f[d.NxNy + t] += s0;
}
template <typename T>
__global__ void k_cum_y(Dev d, FieldSet<T> fs) {
  T* __restrict__ f = fs.f[blockIdx.y];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)d.Nx * (d.Nz - 2);
  if (t >= n) return;
  const int i = (int)(t % d.Nx), k = 1 + (int)(t / d.Nx);
  const T s0 = f[at(d, i, d.Ny - 1, k)], s1 = f[at(d, i, 0, k)];
  f[at(d, i, d.Ny - 2, k)] += s1;
  f[at(d, i, 1, k)] += s0;
}
template <typename T>
__global__ void k_cum_x(Dev d, FieldSet<T> fs) {
  T* __restrict__ f = fs.f[blockIdx.y];
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)(d.Ny - 2) * (d.Nz - 2);
  if (t >= n) return;
  const int j = 1 + (int)(t % (d.Ny - 2)), k = 1 + (int)(t / (d.Ny - 2));
  const T s0 = f[at(d, d.Nx - 1, j, k)], s1 = f[at(d, 0, j, k)];
  f[at(d, d.Nx - 2, j, k)] += s1;
  f[at(d, 1, j, k)] += s0;
}

template <typename T>
int cum_fold_n(const Dev& d, T* const* f, int nf, hipStream_t s) {
  if (nf < 1 || nf > FMAX) return FR_EINVAL;
  FieldSet<T> fs{};
  for (int q = 0; q < nf; q++) fs.f[q] = f[q];
  hipLaunchKernelGGL(k_cum_z<T>, dim3(nblocks(d.NxNy, 256), nf), dim3(256), 0, s, d, fs);
  hipLaunchKernelGGL(k_cum_y<T>, dim3(nblocks((long)d.Nx * (d.Nz - 2), 256), nf), dim3(256), 0,
                     s, d, fs);
  if (d.box)
    hipLaunchKernelGGL(k_cum_x<T>, dim3(nblocks((long)(d.Ny - 2) * (d.Nz - 2), 256), nf),
                       dim3(256), 0, s, d, fs);
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int cum_fold_n<double>(const Dev&, double* const*, int, hipStream_t);
template int cum_fold_n<float>(const Dev&, float* const*, int, hipStream_t);

template <typename T>
int cum_fold(const Dev& d, T* f, hipStream_t s) {
  return cum_fold_n(d, &f, 1, s);
}
template int cum_fold<double>(const Dev&, double*, hipStream_t);
template int cum_fold<float>(const Dev&, float*, hipStream_t);

// Element-wise forms of tensor0.cpp:93-241 (see fr_t0_elementwise).
template <typename T>
__global__ void k_elem(long n, int op, T* __restrict__ out, double a_, const T* __restrict__ A,
                       double b_, const T* __restrict__ B, double c_, const T* __restrict__ C) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const T a = (T)a_, b = (T)b_, c = (T)c_;
  T v;
  switch (op) {
    case 0:
      v = a * A[t];
      if (B) v = v + b * B[t];
      if (C) v = v + c * C[t];
      out[t] = v;
      break;
    case 1: out[t] += a * A[t]; break;
    case 2: out[t] = A[t] * B[t] + b * C[t]; break;
    case 3: out[t] = a / A[t]; break;
    case 4: out[t] = A[t] / B[t]; break;
    case 5: out[t] = a; break;
    case 6: out[t] += a; break;
    case 7: out[t] *= a; break;
    case 8: out[t] = A[t]; break;
    case 9: out[t] *= A[t]; break;
    case 10: out[t] -= A[t]; break;
    case 11: out[t] += A[t]; break;
  }
}

// ---- deterministic reductions ------------------------------------------------
constexpr int RB = 256;   // threads per reduction block
constexpr int RG = 1024;  // fixed number of partial blocks

template <int OP>
__device__ __forceinline__ double rcomb(double x, double y) {
  if (OP == 0 || OP == 5) return x + y;
  if (OP == 2 || OP == 4) return y > x ? y : x;
  return y < x ? y : x;  // min
}

template <int OP>
__device__ __forceinline__ double rmap(double v) {
  if (OP == 4) return v >= 0 ? v : -v;
  if (OP == 5) return v * v;
  return v;
}

template <int OP>
__device__ double block_reduce(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = rcomb<OP>(sh[t], sh[t + s]);
    __syncthreads();
  }
  return sh[0];
}

template <int OP, typename T>
__global__ void __launch_bounds__(RB) k_reduce_part(Dev d, const T* __restrict__ f, double init,
                                                    double* __restrict__ part) {
  __shared__ double sh[RB];
  const int ni = d.Nx - 2, nj = d.Ny - 2, nk = d.Nz - 2;
  const int rows = nj * nk;
  double acc = init;
  // block b takes interior rows b, b + RG, ...; its threads stride along x.
  // Eight rows are loaded before they are combined (in row order), so eight
  // loads per thread are in flight instead of one.
  constexpr int UB = 8;
  for (int i = 1 + threadIdx.x; i <= ni; i += RB) {
    for (int r0 = blockIdx.x; r0 < rows; r0 += UB * RG) {
      double v[UB];
#pragma unroll
      for (int u = 0; u < UB; u++) {
        const int row = r0 + u * RG;
        v[u] = 0.0;
        if (row < rows) {
          const int j = 1 + row % nj, k = 1 + row / nj;
          v[u] = (double)f[at(d, i, j, k)];
        }
      }
#pragma unroll
      for (int u = 0; u < UB; u++)
        if (r0 + u * RG < rows) acc = rcomb<OP>(acc, rmap<OP>(v[u]));
    }
  }
  double r = block_reduce<OP>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <int OP>
__global__ void __launch_bounds__(RB) k_reduce_final(const double* __restrict__ part,
                                                     double init, double* __restrict__ out) {
  __shared__ double sh[RB];
  double acc = init;
  for (int t = threadIdx.x; t < RG; t += RB) acc = rcomb<OP>(acc, part[t]);
  double r = block_reduce<OP>(acc, sh);
  if (threadIdx.x == 0) *out = r;
}

// Reduction workspace (partials + result), one per process.
static double* g_red_ws = nullptr;
static double* g_red_host = nullptr;

static int red_ws() {
  if (!g_red_ws) {
    FR_CHECK(hipMalloc(&g_red_ws, sizeof(double) * (RG + 8)));
    FR_CHECK(hipHostMalloc(&g_red_host, sizeof(double) * 8, hipHostMallocDefault));
  }
  return FR_OK;
}

// Launch the reduction and leave the result in device memory at *dres.  Sums
// and extrema are formed in double whatever the field type.
template <typename T>
int reduce_async_ws(const Dev& d, const T* f, int op, double* dres, double* g_red_ws,
                    hipStream_t s);

template <typename T>
int reduce_async(const Dev& d, const T* f, int op, double* dres, hipStream_t s) {
  int rc = red_ws();
  if (rc) return rc;
  return reduce_async_ws(d, f, op, dres, g_red_ws, s);
}
template int reduce_async<double>(const Dev&, const double*, int, double*, hipStream_t);
template int reduce_async<float>(const Dev&, const float*, int, double*, hipStream_t);

// The same with a caller-owned partial-sum workspace of RG doubles (for a
// reduction on a second stream beside the default workspace's users).
template <typename T>
int reduce_async_ws(const Dev& d, const T* f, int op, double* dres, double* g_red_ws,
                    hipStream_t s) {
  const double first = 0.0;
  switch (op) {
    case 0:
    case 1:
      hipLaunchKernelGGL((k_reduce_part<0, T>), dim3(RG), dim3(RB), 0, s, d, f, first, g_red_ws);
      hipLaunchKernelGGL(k_reduce_final<0>, dim3(1), dim3(RB), 0, s, g_red_ws, first, dres);
      break;
    case 5:
      hipLaunchKernelGGL((k_reduce_part<5, T>), dim3(RG), dim3(RB), 0, s, d, f, first, g_red_ws);
      hipLaunchKernelGGL(k_reduce_final<5>, dim3(1), dim3(RB), 0, s, g_red_ws, first, dres);
      break;
    case 2:
      hipLaunchKernelGGL((k_reduce_part<2, T>), dim3(RG), dim3(RB), 0, s, d, f, -HUGE_VAL,
                         g_red_ws);
      hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), dim3(RB), 0, s, g_red_ws, -HUGE_VAL, dres);
      break;
    case 3:
      hipLaunchKernelGGL((k_reduce_part<3, T>), dim3(RG), dim3(RB), 0, s, d, f, HUGE_VAL,
                         g_red_ws);
      hipLaunchKernelGGL(k_reduce_final<3>, dim3(1), dim3(RB), 0, s, g_red_ws, HUGE_VAL, dres);
      break;
    case 4:
      hipLaunchKernelGGL((k_reduce_part<4, T>), dim3(RG), dim3(RB), 0, s, d, f, 0.0, g_red_ws);
      hipLaunchKernelGGL(k_reduce_final<4>, dim3(1), dim3(RB), 0, s, g_red_ws, 0.0, dres);
      break;
    default:
      return FR_EINVAL;
  }
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int reduce_async_ws<double>(const Dev&, const double*, int, double*, double*,
                                     hipStream_t);
template int reduce_async_ws<float>(const Dev&, const float*, int, double*, double*, hipStream_t);

// (*this) -= mean (tensor0.cpp:437-441) with the mean kept on the device:
// every element gets + (T)(-(sum/denom)), the value tensor0::add(-mean)
// adds, with no host round trip.
template <typename T>
__global__ void k_add_neg_mean(long n, T* __restrict__ f, const double* __restrict__ dsum,
                               double denom) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const T a = (T)(-(dsum[0] / denom));
  f[t] += a;
}

template <typename T>
int make_mean_zero_async(const Dev& d, T* f, long n, double* ws, hipStream_t s) {
  FR_RET(reduce_async_ws(d, (const T*)f, 0, ws + RED_PARTS, ws, s));
  const double denom = (double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2);
  hipLaunchKernelGGL(k_add_neg_mean<T>, dim3(nblocks(n, 256)), dim3(256), 0, s, n, f,
                     (const double*)(ws + RED_PARTS), denom);
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int make_mean_zero_async<double>(const Dev&, double*, long, double*, hipStream_t);
template int make_mean_zero_async<float>(const Dev&, float*, long, double*, hipStream_t);

constexpr int ST_SUM = 1024;
// out[0] = the sum; with denom > 0 also out[1] = -(sum/denom), the value
// make_mean_zero adds (k_add_neg_mean forms the same expression).  Workgroup
// g sums part[g*n .. g*n+n) into out[2g], out[2g+1].
__global__ void __launch_bounds__(ST_SUM) k_sum_parts(const double* __restrict__ part, long n,
                                                      double* __restrict__ out, double denom) {
  __shared__ double ws[ST_SUM / 64];
  part += blockIdx.x * n;
  out += 2 * blockIdx.x;
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long t = threadIdx.x;
  for (; t + 7 * ST_SUM < n; t += 8 * ST_SUM) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = part[t + q * ST_SUM];
#pragma unroll
    for (int q = 0; q < 8; q++) a[q] += v[q];
  }
  for (; t < n; t += ST_SUM) a[0] += part[t];
  double r = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x >= 64) return;
  r = threadIdx.x < ST_SUM / 64 ? ws[threadIdx.x] : 0.0;
  for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o);
  if (threadIdx.x == 0) {
    out[0] = r;
    if (denom > 0) out[1] = -(r / denom);
  }
}

int sum_parts(const double* part, long n, double* out, hipStream_t s, double denom, int sets) {
  hipLaunchKernelGGL(k_sum_parts, dim3(sets), dim3(ST_SUM), 0, s, part, n, out, denom);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

template <typename T>
int add_neg_mean_async(const Dev& d, T* f, long n, const double* dsum, hipStream_t s) {
  const double denom = (double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2);
  hipLaunchKernelGGL(k_add_neg_mean<T>, dim3(nblocks(n, 256)), dim3(256), 0, s, n, f, dsum,
                     denom);
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int add_neg_mean_async<double>(const Dev&, double*, long, const double*, hipStream_t);
template int add_neg_mean_async<float>(const Dev&, float*, long, const double*, hipStream_t);

template <typename T>
__global__ void k_add_neg_mean_set(long n, T* __restrict__ f, T* __restrict__ g,
                                   const double* __restrict__ dsum, double denom, int has_c,
                                   double c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const T a = (T)(-(dsum[0] / denom));
  T v = f[t] + a;
  if (has_c) v = v + (T)c;
  f[t] = v;
  g[t] = v;
}

template <typename T>
int make_mean_zero_set_async(const Dev& d, T* f, T* g, long n, double* ws, bool has_c, double c,
                             hipStream_t s) {
  FR_RET(reduce_async_ws(d, (const T*)f, 0, ws + RED_PARTS, ws, s));
  const double denom = (double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2);
  hipLaunchKernelGGL(k_add_neg_mean_set<T>, dim3(nblocks(n, 256)), dim3(256), 0, s, n, f, g,
                     (const double*)(ws + RED_PARTS), denom, has_c ? 1 : 0, c);
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int make_mean_zero_set_async<double>(const Dev&, double*, double*, long, double*, bool,
                                              double, hipStream_t);
template int make_mean_zero_set_async<float>(const Dev&, float*, float*, long, double*, bool,
                                             double, hipStream_t);

template <typename T>
int reduce_sync(const Dev& d, const T* f, int op, double* result, hipStream_t s) {
  int rc = red_ws();
  if (rc) return rc;
  rc = reduce_async(d, f, op, g_red_ws + RG, s);
  if (rc) return rc;
  FR_CHECK(hipMemcpyAsync(g_red_host, g_red_ws + RG, sizeof(double), hipMemcpyDeviceToHost, s));
  FR_CHECK(hipStreamSynchronize(s));
  double v = g_red_host[0];
  if (op == 1) v = v / ((double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2));
  *result = v;
  return FR_OK;
}
template int reduce_sync<double>(const Dev&, const double*, int, double*, hipStream_t);
template int reduce_sync<float>(const Dev&, const float*, int, double*, hipStream_t);

// ---- scratch of the allocation-free reduction (fr_t0_reduce_noalloc) ---------
// The RED_PARTS partial results and the result of one reduction live in a
// file-scope device array, `scratch`, which every GPU the library runs on
// holds (8.3 KB) until the process ends.  Its name in host code is a host-side
// stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for it, looked up once per device with that
// device current and kept per device (each GPU holds its own copy at its own
// address).  Calls that overlap on one GPU share this array and race.
constexpr int NOALLOC_DOUBLES = RED_PARTS + 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

int noalloc_scratch(double** out) {
  int dev = 0;
  FR_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    fprintf(stderr, "fr_amd: device %d beyond the %d the scratch table holds\n", dev,
            NOALLOC_MAX_DEVICES);
    return FR_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FR_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return FR_OK;
}

// communicator.cpp:635-646: one block per x index i.
__global__ void __launch_bounds__(RB) k_reducex(Dev d, const double* __restrict__ f,
                                                double denom, double* __restrict__ out) {
  __shared__ double sh[RB];
  const int i = blockIdx.x;
  const int nj = d.Ny - 2, nk = d.Nz - 2;
  double acc = 0;
  for (long t = threadIdx.x; t < (long)nj * nk; t += RB) {
    const int j = 1 + (int)(t % nj), k = 1 + (int)(t / nj);
    acc += f[at(d, i, j, k)];
  }
  double r = block_reduce<0>(acc, sh);
  if (threadIdx.x == 0) out[i] = r / denom;
}

int reducex(const Dev& d, const double* f, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reducex, dim3(d.Nx), dim3(RB), 0, s, d, f,
                     (double)((d.Ny - 2) * (d.Nz - 2)), out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

template <typename T>
int elementwise(long n, int op, T* out, double a, const T* A, double b, const T* B, double c,
                const T* C, hipStream_t s) {
  hipLaunchKernelGGL(k_elem<T>, dim3(nblocks(n, 256)), dim3(256), 0, s, n, op, out, a, A, b, B,
                     c, C);
  FR_LAUNCH_CHECK();
  return FR_OK;
}
template int elementwise<double>(long, int, double*, double, const double*, double,
                                 const double*, double, const double*, hipStream_t);
template int elementwise<float>(long, int, float*, double, const float*, double, const float*,
                                double, const float*, hipStream_t);

// The reference's element-wise members, one expression each with its own
// operand order (tensor0.cpp:93-241; IEEE products commute, sums of two
// terms commute, three terms add left to right).  Operands may alias out:
// every lane reads its own element before writing it.
__global__ void k_arith(long n, int op, double* out, double a, const double* A, const double* B,
                        double b, const double* C, double c, const double* D) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double o = out[t];
  double v;
  switch (op) {
    case AR_SET_S: v = a; break;
    case AR_ADD_S: v = o + a; break;
    case AR_SUB_S: v = o - a; break;
    case AR_MUL_S: v = o * a; break;
    case AR_DIV_S: v = o / a; break;
    case AR_SET_T: v = A[t]; break;
    case AR_ADD_T: v = o + A[t]; break;
    case AR_SUB_T: v = o - A[t]; break;
    case AR_MUL_T: v = o * A[t]; break;
    case AR_DIV_T: v = o / A[t]; break;
    case AR_PEM_TT: v = o + A[t] * B[t]; break;
    case AR_PEM_ST: v = o + a * A[t]; break;
    case AR_DIV_TT: v = A[t] / B[t]; break;
    case AR_DIV_TS: v = A[t] / a; break;
    case AR_DIV_ST: v = a / A[t]; break;
    case AR_MUL_TT: v = A[t] * B[t]; break;
    case AR_MUL_ST: v = a * A[t]; break;
    case AR_LC_STST: v = a * A[t] + b * C[t]; break;
    case AR_LC_TTST: v = A[t] * B[t] + b * C[t]; break;
    case AR_LC_TTTT: v = A[t] * B[t] + C[t] * D[t]; break;
    case AR_LC_STTT: v = a * A[t] + C[t] * D[t]; break;
    case AR_LC_STSTST: v = a * A[t] + b * C[t] + c * D[t]; break;
    default: v = o;
  }
  out[t] = v;
}

int t0_arith(long n, int op, double* out, double a, const double* A, const double* B, double b,
             const double* C, double c, const double* D, hipStream_t s) {
  if (op < 0 || op >= AR_NOPS || !out) return FR_EINVAL;
  hipLaunchKernelGGL(k_arith, dim3(nblocks(n, 256)), dim3(256), 0, s, n, op, out, a, A, B, b, C,
                     c, D);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

__global__ void k_f2d(long n, const float* __restrict__ a, double* __restrict__ b) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (double)a[t];
}

int to_double(const float* src, double* dst, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(nblocks(n, 256)), dim3(256), 0, s, n, src, dst);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

// Equal_I{x,y,z}_C2F (tensor0.cpp:353-378) into `out`, any field type.
template <typename T>
int interp_c2f(const fr_geom* g, T* out, const T* in, int axis, hipStream_t s) {
  const int Nx = g->Nx;
  if (axis == 0) return launch_stencil<ST_IX_C2F, T>(g, out, in, nullptr, nullptr, 1, Nx, s);
  if (axis == 1) return launch_stencil<ST_IY_C2F, T>(g, out, in, nullptr, nullptr, 0, Nx, s);
  return launch_stencil<ST_IZ_C2F, T>(g, out, in, nullptr, nullptr, 0, Nx, s);
}
template int interp_c2f<double>(const fr_geom*, double*, const double*, int, hipStream_t);
template int interp_c2f<float>(const fr_geom*, float*, const float*, int, hipStream_t);

}  // namespace fr

using namespace fr;

extern "C" {

const char* fr_version(void) { return "fr_amd 0.1 gfx950"; }

int fr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int fr_t0_div_f2c(const fr_geom* g, double* out, const double* tx, const double* ty,
                  const double* tz, void* s) {
  if (!tx || !ty || !tz) return FR_EINVAL;
  return launch_stencil<ST_DIV>(g, out, tx, ty, tz, 0, g->Nx - 1, s);
}

int fr_t0_del2(const fr_geom* g, double* out, const double* in, void* s) {
  if (!in) return FR_EINVAL;
  return launch_stencil<ST_DEL2>(g, out, in, nullptr, nullptr, 1, g->Nx - 1, s);
}

int fr_t1_grad_c2f(const fr_geom* g, double* ox, double* oy, double* oz,
                   const double* in, void* s) {
  if (!in || !ox || !oy || !oz) return FR_EINVAL;
  int rc = launch_stencil<ST_GX>(g, ox, in, nullptr, nullptr, 1, g->Nx, s);
  if (!rc) rc = launch_stencil<ST_GY>(g, oy, in, nullptr, nullptr, 0, g->Nx, s);
  if (!rc) rc = launch_stencil<ST_GZ>(g, oz, in, nullptr, nullptr, 0, g->Nx, s);
  return rc;
}

int fr_t0_interp(const fr_geom* g, double* out, const double* in, int axis, int dir,
                 void* s) {
  if (!in || axis < 0 || axis > 2 || dir < 0 || dir > 1) return FR_EINVAL;
  const int Nx = g->Nx;
  if (dir == 0) {
    if (axis == 0) return launch_stencil<ST_IX_C2F>(g, out, in, nullptr, nullptr, 1, Nx, s);
    if (axis == 1) return launch_stencil<ST_IY_C2F>(g, out, in, nullptr, nullptr, 0, Nx, s);
    return launch_stencil<ST_IZ_C2F>(g, out, in, nullptr, nullptr, 0, Nx, s);
  }
  if (axis == 0) return launch_stencil<ST_IX_F2C>(g, out, in, nullptr, nullptr, 0, Nx - 1, s);
  if (axis == 1) return launch_stencil<ST_IY_F2C>(g, out, in, nullptr, nullptr, 0, Nx, s);
  return launch_stencil<ST_IZ_F2C>(g, out, in, nullptr, nullptr, 0, Nx, s);
}

int fr_t0_i_f2c(const fr_geom* g, double* out, const double* tx, const double* ty,
                const double* tz, void* s) {
  if (!tx || !ty || !tz) return FR_EINVAL;
  return launch_stencil<ST_I_F2C>(g, out, tx, ty, tz, 0, g->Nx - 1, s);
}

int fr_t0_update_ghosts(const fr_geom* g, double* f, int cumulative, void* s) {
  if (!geom_ok(g) || !f) return FR_EINVAL;
  Dev d = make_dev(g);
  if (cumulative) return cum_fold(d, f, S(s));
  return ghost_fill(d, f, S(s));
}

int fr_t0_elementwise(const fr_geom* g, int op, double* out, double a, const double* A,
                      double b, const double* B, double c, const double* Cf, void* s) {
  if (!geom_ok(g) || !out || op < 0 || op > 11) return FR_EINVAL;
  if ((op == 0 || op == 1 || op == 3 || op == 8 || op == 9 || op == 10 || op == 11) && !A)
    return FR_EINVAL;
  if ((op == 2 || op == 4) && (!A || !B)) return FR_EINVAL;
  if (op == 2 && !Cf) return FR_EINVAL;
  Dev d = make_dev(g);
  return elementwise(d.size, op, out, a, A, b, B, c, Cf, S(s));
}

int fr_t0_reduce(const fr_geom* g, const double* f, int op, double* result, void* s) {
  if (!geom_ok(g) || !f || !result || op < 0 || op > 5) return FR_EINVAL;
  Dev d = make_dev(g);
  return reduce_sync(d, f, op, result, S(s));
}

int fr_t0_reduce_noalloc(const fr_geom* g, const double* f, int op, double* result, void* s) {
  if (!geom_ok(g) || !f || !result || op < 0 || op > 5) {
    fprintf(stderr,
            "fr_amd: fr_t0_reduce_noalloc: bad argument (op %d; needs bs = 1, ghosted sizes "
            ">= 3 and fewer than 2^31 ghosted cells)\n",
            op);
    return FR_EINVAL;
  }
  double* ws = nullptr;
  FR_RET(fr::noalloc_scratch(&ws));
  Dev d = make_dev(g);
  // partial results in ws[0, RED_PARTS), the result at ws[RED_PARTS]; every
  // partial is written by the call before it is read, so nothing is cleared
  FR_RET(reduce_async_ws(d, f, op, ws + RED_PARTS, ws, S(s)));
  FR_CHECK(hipMemcpyAsync(result, ws + RED_PARTS, sizeof(double), hipMemcpyDeviceToHost, S(s)));
  FR_CHECK(hipStreamSynchronize(S(s)));
  if (op == 1) *result = *result / ((double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2));
  return FR_OK;
}

int fr_t0_reducex(const fr_geom* g, const double* f, double* out, void* s) {
  if (!geom_ok(g) || !f || !out) return FR_EINVAL;
  Dev d = make_dev(g);
  return reducex(d, f, out, S(s));
}

}  // extern "C"
