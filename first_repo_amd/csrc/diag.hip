// Step diagnostics of grid::Statistics (grid.cpp:714-811) needed to compare
// runs: TKE (tensor1::Equal_TKE, tensor1.cpp:415-426), the dissipation
// profile (tensor0::Equal_SijSij/Equal_dissipation, tensor0.cpp:584-625),
// plus the mean particle temperature.
#include "fr_internal.hpp"

namespace fr {

// u = RU_m/Rho (whole array, grid.cpp:738) was stored in `u`; red = its
// plane means.  out = 0.5*Rho*(u - red[i])*(u - red[i]).
__global__ void k_tke(Dev d, const double* __restrict__ u, const double* __restrict__ rho,
                      const double* __restrict__ red, double* __restrict__ out) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.size) return;
  const int i = (int)(n % d.Nx);
  const double a = u[n] - red[i];
  out[n] = 0.5 * rho[n] * a * a;
}

// 2*nu*SijSij at interior cells, nu = Mu0/Rho; SijSij of the face velocity
// U (tensor0.cpp:584-595), squares taken as products.
__global__ void k_diss(Dev d, const double* __restrict__ ux, const double* __restrict__ uy,
                       const double* __restrict__ uz, const double* __restrict__ rho,
                       double mu, double* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i < 1 || j < 1 || k < 1 || i >= d.Nx - 1 || j >= d.Ny - 1 || k >= d.Nz - 1) return;
  const long n = at(d, i, j, k), sx = 1, sy = d.Nx, sz = d.NxNy;
  auto d1 = [&](const double* f, long q, long s, double id) { return (f[q + s] - f[q]) * id; };
  auto c2 = [&](const double* f, long q, long s, double i2) { return (f[q + s] - f[q - s]) * i2; };
  const double a = d1(ux, n, sx, d.idx), b = d1(uy, n, sy, d.idy), e = d1(uz, n, sz, d.idz);
  double r = a * a + b * b + e * e;
  double t = c2(ux, n, sy, d.i2dy) + c2(ux, n + sx, sy, d.i2dy) + c2(uy, n, sx, d.i2dx) +
             c2(uy, n + sy, sx, d.i2dx);
  r += 0.125 * (t * t);
  t = c2(ux, n, sz, d.i2dz) + c2(ux, n + sx, sz, d.i2dz) + c2(uz, n, sx, d.i2dx) +
      c2(uz, n + sz, sx, d.i2dx);
  r += 0.125 * (t * t);
  t = c2(uy, n, sz, d.i2dz) + c2(uy, n + sy, sz, d.i2dz) + c2(uz, n, sy, d.i2dy) +
      c2(uz, n + sz, sy, d.i2dy);
  r += 0.125 * (t * t);
  const double nu = mu / rho[n];
  out[n] = 2 * nu * r;
}

__global__ void k_psum(const double* __restrict__ v, int n, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += 256L * gridDim.x) acc += v[t];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

int diag_tke(const Dev& d, const double* u, const double* rho, const double* red,
             double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_tke, dim3(nblocks(d.size, 256)), dim3(256), 0, s, d, u, rho, red, out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int diag_diss(const Dev& d, const double* const u[3], const double* rho, double mu,
              double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_diss, dim3(nblocks(d.Nx, 64), nblocks(d.Ny, 4), d.Nz), dim3(64, 4), 0, s,
                     d, u[0], u[1], u[2], rho, mu, out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

// Deterministic sum of n doubles: 256 partial blocks then one.
int vsum(const double* v, int n, double* part257, double* host, hipStream_t s) {
  hipLaunchKernelGGL(k_psum, dim3(256), dim3(256), 0, s, v, n, part257);
  hipLaunchKernelGGL(k_psum, dim3(1), dim3(256), 0, s, part257, 256, part257 + 256);
  FR_LAUNCH_CHECK();
  FR_CHECK(hipMemcpyAsync(host, part257 + 256, sizeof(double), hipMemcpyDeviceToHost, s));
  FR_CHECK(hipStreamSynchronize(s));
  return FR_OK;
}

}  // namespace fr
