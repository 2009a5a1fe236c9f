// Field operators of grid::Statistics (grid.cpp:714-866) on gfx950: the
// post-process forms of tensor0.cpp:575-663 and tensor1.cpp:375-439, each a
// whole-array (or, for the strain and rotation sums, interior) pass whose
// operands are fields and per-x plane means; the plane means themselves
// come from reducex (communicator.cpp:635-646).  Every expression keeps the
// reference's operand order.  Statistics are outside the timed step, so
// these favour plainness over fusion.
#include "fr_internal.hpp"

namespace fr {

// OP codes of k_stat_point (see stat_point in fr_internal.hpp).
template <int OP>
__global__ void k_stat_point(Dev d, double* __restrict__ o0, double* __restrict__ o1,
                             double* __restrict__ o2, const double* __restrict__ a0,
                             const double* __restrict__ a1, const double* __restrict__ a2,
                             const double* __restrict__ b0, const double* __restrict__ b1,
                             const double* __restrict__ b2, const double* __restrict__ r,
                             const double* __restrict__ ra, const double* __restrict__ rr,
                             double s) {
  const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= d.size) return;
  const int i = (int)(n % d.Nx);
  const double* A[3] = {a0, a1, a2};
  const double* B[3] = {b0, b1, b2};
  double* O[3] = {o0, o1, o2};
  if (OP == ST_RSD) {  // tensor0.cpp:652-663 (T = r, reduced)
    o0[n] = (r[n] - rr[i]) * (r[n] - rr[i]);
  } else if (OP == ST_ST) {  // tensor0.cpp:642-650, s = 2*Tp*Tp
    o0[n] = sqrt(s * a0[n]);
  } else if (OP == ST_NU) {  // tensor0.cpp:627-631: Equal_Divide(mu, Rho)
    o0[n] = s / r[n];
  } else if (OP == ST_DISS) {  // tensor0.cpp:618-625
    o0[n] = 2 * a0[n] * b0[n];
  } else if (OP == ST_ETA) {  // tensor0.cpp:633-640
    o0[n] = pow(a0[n] * a0[n] * a0[n] / b0[n], 0.25);
  } else {
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const double* red_m = ra + (long)m * d.Nx;
      if (OP == ST_TKE2) {  // tensor1.cpp:375-386 (B: RU reduced in ra, R in rr)
        O[m][n] = 0.5 * r[n] * (A[m][n] - red_m[i] / rr[i]) * (A[m][n] - red_m[i] / rr[i]);
      } else if (OP == ST_UIUI2) {  // tensor1.cpp:388-399
        O[m][n] = (A[m][n] - red_m[i] / rr[i]) * (A[m][n] - red_m[i] / rr[i]);
      } else if (OP == ST_CORRU) {  // tensor1.cpp:401-412
        O[m][n] = (A[m][n] - red_m[i]) * (r[n] - rr[i]);
      } else if (OP == ST_TKE) {  // tensor1.cpp:415-426
        O[m][n] = 0.5 * r[n] * (A[m][n] - red_m[i]) * (A[m][n] - red_m[i]);
      } else if (OP == ST_UIUI) {  // tensor1.cpp:428-439
        O[m][n] = (A[m][n] - red_m[i]) * (A[m][n] - red_m[i]);
      } else if (OP == ST_DIV0) {  // tensor1.cpp:174-179: B_m / R
        O[m][n] = B[m][n] / r[n];
      } else if (OP == ST_DIV1) {  // tensor1.cpp:153-158: A_m / B_m
        O[m][n] = A[m][n] / B[m][n];
      }
    }
  }
}

// tensor0::Equal_SijSij (tensor0.cpp:584-595, WIJ = 0) or Equal_WijWij
// (597-607, WIJ = 1, added to out's prior content) of the face velocity U
// at interior cells; pow(t, 2) is t*t exactly.
template <int WIJ>
__global__ void k_strain(Dev d, const double* __restrict__ ux, const double* __restrict__ uy,
                         const double* __restrict__ uz, double* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i < 1 || j < 1 || k < 1 || i >= d.Nx - 1 || j >= d.Ny - 1 || k >= d.Nz - 1) return;
  const long n = at(d, i, j, k), sx = 1, sy = d.Nx, sz = d.NxNy;
  auto d1 = [&](const double* f, long q, long s, double id) { return (f[q + s] - f[q]) * id; };
  auto c2 = [&](const double* f, long q, long s, double i2) { return (f[q + s] - f[q - s]) * i2; };
  double r, t;
  if (WIJ == 0) {
    const double a = d1(ux, n, sx, d.idx), b = d1(uy, n, sy, d.idy), e = d1(uz, n, sz, d.idz);
    r = a * a + b * b + e * e;
    t = c2(ux, n, sy, d.i2dy) + c2(ux, n + sx, sy, d.i2dy) + c2(uy, n, sx, d.i2dx) +
        c2(uy, n + sy, sx, d.i2dx);
    r += 0.125 * (t * t);
    t = c2(ux, n, sz, d.i2dz) + c2(ux, n + sx, sz, d.i2dz) + c2(uz, n, sx, d.i2dx) +
        c2(uz, n + sz, sx, d.i2dx);
    r += 0.125 * (t * t);
    t = c2(uy, n, sz, d.i2dz) + c2(uy, n + sy, sz, d.i2dz) + c2(uz, n, sy, d.i2dy) +
        c2(uz, n + sz, sy, d.i2dy);
    r += 0.125 * (t * t);
  } else {
    r = out[n];
    t = c2(ux, n, sy, d.i2dy) + c2(ux, n + sx, sy, d.i2dy) - c2(uy, n, sx, d.i2dx) -
        c2(uy, n + sy, sx, d.i2dx);
    r += 0.125 * (t * t);
    t = c2(ux, n, sz, d.i2dz) + c2(ux, n + sx, sz, d.i2dz) - c2(uz, n, sx, d.i2dx) -
        c2(uz, n + sz, sx, d.i2dx);
    r += 0.125 * (t * t);
    t = c2(uy, n, sz, d.i2dz) + c2(uy, n + sy, sz, d.i2dz) - c2(uz, n, sy, d.i2dy) -
        c2(uz, n + sz, sy, d.i2dy);
    r += 0.125 * (t * t);
  }
  out[n] = r;
}

int stat_point(const Dev& d, int op, double* const o[3], const double* const a[3],
               const double* const b[3], const double* r, const double* ra, const double* rr,
               double s, hipStream_t st) {
  const dim3 g(nblocks(d.size, 256)), bl(256);
#define FR_STAT_CASE(OPC)                                                                    \
  case OPC:                                                                                  \
    hipLaunchKernelGGL(k_stat_point<OPC>, g, bl, 0, st, d, o[0], o[1], o[2], a[0], a[1], a[2], \
                       b[0], b[1], b[2], r, ra, rr, s);                                     \
    break;
  switch (op) {
    FR_STAT_CASE(ST_RSD)
    FR_STAT_CASE(ST_TKE2)
    FR_STAT_CASE(ST_UIUI2)
    FR_STAT_CASE(ST_CORRU)
    FR_STAT_CASE(ST_TKE)
    FR_STAT_CASE(ST_UIUI)
    FR_STAT_CASE(ST_ST)
    FR_STAT_CASE(ST_NU)
    FR_STAT_CASE(ST_DISS)
    FR_STAT_CASE(ST_ETA)
    FR_STAT_CASE(ST_DIV0)
    FR_STAT_CASE(ST_DIV1)
    default:
      return FR_EINVAL;
  }
#undef FR_STAT_CASE
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int stat_strain(const Dev& d, bool rotation, const double* const u[3], double* out,
                hipStream_t st) {
  const dim3 g(nblocks(d.Nx, 64), nblocks(d.Ny, 4), d.Nz), bl(64, 4);
  if (rotation)
    hipLaunchKernelGGL(k_strain<1>, g, bl, 0, st, d, u[0], u[1], u[2], out);
  else
    hipLaunchKernelGGL(k_strain<0>, g, bl, 0, st, d, u[0], u[1], u[2], out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

}  // namespace fr
