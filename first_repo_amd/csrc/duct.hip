// RK4-substep kernels of the inflow/outflow duct (grid.cpp:335-385, 427-537,
// 556-657 with IsBox = 0), for one device that holds the whole duct (the
// inlet and the outlet rank of the reference at once).
//
// Duct fields wrap periodically in y and z only; the x ghost planes are the
// inlet (i = 0) and outlet (i = Nx-1) slices.  Each kernel writes every
// ghosted cell: y,z ghosts hold the value at their periodic image, the inlet
// slice takes the copies from the inflow planes, the outlet slice the
// convective outflow update.  Where the reference leaves a stale value in an
// x ghost cell that nothing reads before it is overwritten (RU.x and the
// _np1 stages on the inlet planes, the RHS of x-ghost cells), the kernels
// store a freshly computed one instead.
#include "fr_internal.hpp"

namespace fr {

#define FR_DUCT_PROLOGUE                                    \
  const int i = blockIdx.x * 64 + threadIdx.x;              \
  const int j = blockIdx.y * 4 + threadIdx.y;               \
  const int k = blockIdx.z;                                 \
  if (i >= d.Nx || j >= d.Ny) return;                       \
  const long n = at(d, i, j, k);                            \
  const long c = at(d, i, img(j, d.Ny), img(k, d.Nz));      \
  const long pl = (long)k * d.Ny + j; /* inflow plane index */ \
  const long sx = 1, sy = d.Nx, sz = d.NxNy;                \
  (void)sx; (void)sy; (void)sz; (void)pl;

static dim3 grid3d(const Dev& d) { return dim3(nblocks(d.Nx, 64), nblocks(d.Ny, 4), d.Nz); }

// grid::Update_Rho(grid& g) (grid.cpp:335-385) fused with U = RU_int/Rho_face
// (grid.cpp:684).  RHS_Rho is zero on the outlet slice.
__global__ void __launch_bounds__(256) k_duct_update_rho(
    Dev d, BoxCoef cf, double cbc, const double* __restrict__ rux,
    const double* __restrict__ ruy, const double* __restrict__ ruz,
    const double* __restrict__ rho, const double* __restrict__ rho_int,
    double* __restrict__ rho_np1, double* __restrict__ rho_new, const double* __restrict__ rfx,
    const double* __restrict__ rfy, const double* __restrict__ rfz, double* __restrict__ ux,
    double* __restrict__ uy, double* __restrict__ uz, const double* __restrict__ in_new,
    const double* __restrict__ in_np1) {
  FR_DUCT_PROLOGUE
  double rhs = 0.0;
  if (i < d.Nx - 1)
    rhs = (rux[c + sx] - rux[c]) * d.idx + (ruy[c + sy] - ruy[c]) * d.idy +
          (ruz[c + sz] - ruz[c]) * d.idz;
  double np1 = rho_np1[n] + cf.cpost_neg * rhs;
  double nw = cf.last ? np1 : 1 * rho[n] + cf.cpre_neg * rhs;
  if (i == 0) {  // inflow copies (grid.cpp:352-364)
    nw = in_new[pl];
    if (cf.last) np1 = in_np1[pl];
  } else if (i == d.Nx - 1) {  // convective outflow (grid.cpp:367-384)
    const double diff = rho_int[n] - rho_int[n - 1];
    np1 = np1 - cf.cpost * cbc * diff * d.idx;
    nw = cf.last ? np1 : rho[n] - cf.cpre * cbc * diff * d.idx;
  }
  rho_np1[n] = np1;
  rho_new[n] = nw;
  ux[n] = rux[n] / rfx[n];
  uy[n] = ruy[n] / rfy[n];
  uz[n] = ruz[n] / rfz[n];
}

__device__ __forceinline__ double ddivu(const Dev& d, const double* ux, const double* uy,
                                        const double* uz, long q) {
  return (ux[q + 1] - ux[q]) * d.idx + (uy[q + d.Nx] - uy[q]) * d.idy +
         (uz[q + d.NxNy] - uz[q]) * d.idz;
}

template <int M>
__device__ __forceinline__ double dconv(const Dev& d, const double* __restrict__ um,
                                        const double* __restrict__ rux,
                                        const double* __restrict__ ruy,
                                        const double* __restrict__ ruz, long c) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  const long sm = M == 0 ? sx : (M == 1 ? sy : sz);
  auto Fx = [&](long q) { return ((um[q] + um[q - sx]) * 0.5) * ((rux[q] + rux[q - sm]) * 0.5); };
  auto Fy = [&](long q) { return ((um[q] + um[q - sy]) * 0.5) * ((ruy[q] + ruy[q - sm]) * 0.5); };
  auto Fz = [&](long q) { return ((um[q] + um[q - sz]) * 0.5) * ((ruz[q] + ruz[q - sm]) * 0.5); };
  return (Fx(c + sx) - Fx(c)) * d.idx + (Fy(c + sy) - Fy(c)) * d.idy +
         (Fz(c + sz) - Fz(c)) * d.idz;
}

// grid::Update_RU_WOP(grid& g) (grid.cpp:427-537): viscous, grad-div and
// convective terms without forcing or gravity; RHS zero on the outlet slice;
// RU_new.x(1), RU_new.y(0), RU_new.z(0) (and RU_np1 at the last stage) from
// the inflow planes; convective outflow on the outlet slice.
__global__ void __launch_bounds__(256) k_duct_update_ru_wop(
    Dev d, BoxCoef cf, WopCoef w, double cbc, const double* __restrict__ ux,
    const double* __restrict__ uy, const double* __restrict__ uz,
    const double* __restrict__ rux, const double* __restrict__ ruy,
    const double* __restrict__ ruz, const double* __restrict__ sxp,
    const double* __restrict__ syp, const double* __restrict__ szp,
    const double* __restrict__ ru0x, const double* __restrict__ ru0y,
    const double* __restrict__ ru0z, double* __restrict__ np1x, double* __restrict__ np1y,
    double* __restrict__ np1z, double* __restrict__ newx, double* __restrict__ newy,
    double* __restrict__ newz, const double* __restrict__ in_new,
    const double* __restrict__ in_np1, long plane) {
  FR_DUCT_PROLOGUE
  const bool out = i == d.Nx - 1;
  double g[3] = {0.0, 0.0, 0.0};
  if (!out) {
    const double dv = ddivu(d, ux, uy, uz, c);
    g[0] = (dv - ddivu(d, ux, uy, uz, c - sx)) * d.idx;
    g[1] = (dv - ddivu(d, ux, uy, uz, c - sy)) * d.idy;
    g[2] = (dv - ddivu(d, ux, uy, uz, c - sz)) * d.idz;
  }
  const double* U[3] = {ux, uy, uz};
  const double* Sd[3] = {sxp, syp, szp};
  const double* R0[3] = {ru0x, ru0y, ru0z};
  const double* RI[3] = {rux, ruy, ruz};
  double* N1[3] = {np1x, np1y, np1z};
  double* NW[3] = {newx, newy, newz};
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double r = 0.0;
    if (!out) {
      const double* u = U[m];
      const double lap = (u[c + sx] - 2 * u[c] + u[c - sx]) * d.idx2 +
                         (u[c + sy] - 2 * u[c] + u[c - sy]) * d.idy2 +
                         (u[c + sz] - 2 * u[c] + u[c - sz]) * d.idz2;
      r = w.a_s * Sd[m][c] + w.b_g * g[m] + w.c_l * lap;
      const double cv = m == 0 ? dconv<0>(d, u, rux, ruy, ruz, c)
                       : m == 1 ? dconv<1>(d, u, rux, ruy, ruz, c)
                                : dconv<2>(d, u, rux, ruy, ruz, c);
      r = r - cv;
    }
    double np1 = N1[m][n] + cf.cpost * r;
    double nw = cf.last ? np1 : 1 * R0[m][n] + cf.cpre * r;
    if (i == (m == 0 ? 1 : 0)) {  // inflow copies (grid.cpp:482-504)
      nw = in_new[m * plane + pl];
      if (cf.last) np1 = in_np1[m * plane + pl];
    }
    if (out) {  // convective outflow (grid.cpp:507-536)
      const double diff = RI[m][n] - RI[m][n - 1];
      np1 = np1 - cf.cpost * cbc * diff * d.idx;
      nw = cf.last ? np1 : R0[m][n] - cf.cpre * cbc * diff * d.idx;
    }
    N1[m][n] = np1;
    NW[m][n] = nw;
  }
}

// grid::Compute_Div_U_new (grid.cpp:556-563) at interior cells into `dvg`,
// and Rho_face = I_C2F(Rho_new) (grid.cpp:582): x faces i >= 1, y and z
// faces everywhere.
__global__ void __launch_bounds__(256) k_duct_div(Dev d, DivCoef q, const double* __restrict__ rn,
                                                  const double* __restrict__ Q,
                                                  double* __restrict__ rfx,
                                                  double* __restrict__ rfy,
                                                  double* __restrict__ rfz,
                                                  double* __restrict__ dvg) {
  FR_DUCT_PROLOGUE
  if (i >= 1) rfx[n] = (rn[c] + rn[c - sx]) * 0.5;
  rfy[n] = (rn[c] + rn[c - sy]) * 0.5;
  rfz[n] = (rn[c] + rn[c - sz]) * 0.5;
  if (i >= 1 && i < d.Nx - 1 && j >= 1 && j < d.Ny - 1 && k >= 1 && k < d.Nz - 1) {
    auto f = [&](long p) { return 1. / rn[p]; };
    const double fc = f(c);
    const double del2 = (f(c + sx) - 2 * fc + f(c - sx)) * d.idx2 +
                        (f(c + sy) - 2 * fc + f(c - sy)) * d.idy2 +
                        (f(c + sz) - 2 * fc + f(c - sz)) * d.idz2;
    const double qq = q.cooling ? Q[c] - q.qmean : Q[c];
    dvg[n] = q.kc * del2 + q.bq * qq;
  }
}

// The two x-face sums of grid::Correct_OutFlow (grid.cpp:585-600): U.x =
// RU_new.x/Rho_face.x summed over the interior (j,k) of the outlet slice and
// of the first inlet face (i = 1); one workgroup, fixed order.
__global__ void __launch_bounds__(256) k_duct_planes(Dev d, const double* __restrict__ rnx,
                                                     const double* __restrict__ rfx,
                                                     double* __restrict__ out) {
  __shared__ double so[256], si[256];
  const int nj = d.Ny - 2, nk = d.Nz - 2;
  double a = 0.0, b = 0.0;
  for (long t = threadIdx.x; t < (long)nj * nk; t += 256) {
    const int j = 1 + (int)(t % nj), k = 1 + (int)(t / nj);
    const long o = at(d, d.Nx - 1, j, k), p = at(d, 1, j, k);
    a += rnx[o] / rfx[o];
    b += rnx[p] / rfx[p];
  }
  so[threadIdx.x] = a;
  si[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      so[threadIdx.x] += so[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = so[0];
    out[1] = si[0];
  }
}

// RU_new.x (and RU_np1.x at the last stage) on the outlet slice times ratio
// (grid.cpp:611-624).
__global__ void k_duct_scale(Dev d, double ratio, int last, double* __restrict__ rnx,
                             double* __restrict__ r1x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)d.Ny * d.Nz) return;
  const long o = t * d.Nx + d.Nx - 1;
  rnx[o] = rnx[o] * ratio;
  if (last) r1x[o] = r1x[o] * ratio;
}

// grid::Compute_RHS_Pois without the Rho_face update (grid.cpp:629-637):
// RHS_Pois = (Div_F2C(RU_new/Rho_face) - divergence) * (1/(dt*pre)) at the
// interior (the pressure solve reads nothing else); zero mean is taken after.
__global__ void __launch_bounds__(256) k_duct_rhs(Dev d, double s, const double* __restrict__ rnx,
                                                  const double* __restrict__ rny,
                                                  const double* __restrict__ rnz,
                                                  const double* __restrict__ rfx,
                                                  const double* __restrict__ rfy,
                                                  const double* __restrict__ rfz,
                                                  const double* __restrict__ dvg,
                                                  double* __restrict__ rhs) {
  FR_DUCT_PROLOGUE
  if (i < 1 || i >= d.Nx - 1 || j < 1 || j >= d.Ny - 1 || k < 1 || k >= d.Nz - 1) return;
  const double D = (rnx[c + sx] / rfx[c + sx] - rnx[c] / rfx[c]) * d.idx +
                   (rny[c + sy] / rfy[c + sy] - rny[c] / rfy[c]) * d.idy +
                   (rnz[c + sz] / rfz[c + sz] - rnz[c] / rfz[c]) * d.idz;
  double r = D - dvg[n];
  r = r * s;
  rhs[n] = r;
}

// grid::Update_RU_WP (grid.cpp:651-657) on the duct: Grad_C2F(P) with its x
// component from i >= 1 (the inlet x face is not a pressure face).
__global__ void __launch_bounds__(256) k_duct_update_ru_wp(
    Dev d, BoxCoef cf, const double* __restrict__ P, double* __restrict__ np1x,
    double* __restrict__ np1y, double* __restrict__ np1z, double* __restrict__ newx,
    double* __restrict__ newy, double* __restrict__ newz) {
  FR_DUCT_PROLOGUE
  const double gx = i >= 1 ? (P[c] - P[c - sx]) * d.idx : 0.0;
  const double gy = (P[c] - P[c - sy]) * d.idy;
  const double gz = (P[c] - P[c - sz]) * d.idz;
  const double ax = np1x[n] + cf.cpost_neg * gx;
  const double ay = np1y[n] + cf.cpost_neg * gy;
  const double az = np1z[n] + cf.cpost_neg * gz;
  np1x[n] = ax;
  np1y[n] = ay;
  np1z[n] = az;
  if (cf.last) {
    newx[n] = ax;
    newy[n] = ay;
    newz[n] = az;
  } else {
    newx[n] = newx[n] + cf.cpre_neg * gx;
    newy[n] = newy[n] + cf.cpre_neg * gy;
    newz[n] = newz[n] + cf.cpre_neg * gz;
  }
}

int duct_update_rho(const Dev& d, const BoxCoef& cf, double cbc, const double* const ru_int[3],
                    const double* rho, const double* rho_int, double* rho_np1, double* rho_new,
                    const double* const rf[3], double* const u[3], const double* in_new,
                    const double* in_np1, hipStream_t s) {
  hipLaunchKernelGGL(k_duct_update_rho, grid3d(d), dim3(64, 4), 0, s, d, cf, cbc, ru_int[0],
                     ru_int[1], ru_int[2], rho, rho_int, rho_np1, rho_new, rf[0], rf[1], rf[2],
                     u[0], u[1], u[2], in_new, in_np1);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int duct_update_ru_wop(const Dev& d, const BoxCoef& cf, const WopCoef& w, double cbc,
                       const double* const u[3], const double* const ru_int[3],
                       const double* const S[3], const double* const ru[3],
                       double* const ru_np1[3], double* const ru_new[3], const double* in_new,
                       const double* in_np1, hipStream_t s) {
  const long plane = (long)d.Ny * d.Nz;
  hipLaunchKernelGGL(k_duct_update_ru_wop, grid3d(d), dim3(64, 4), 0, s, d, cf, w, cbc, u[0], u[1],
                     u[2], ru_int[0], ru_int[1], ru_int[2], S[0], S[1], S[2], ru[0], ru[1], ru[2],
                     ru_np1[0], ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2], in_new,
                     in_np1, plane);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int duct_div(const Dev& d, const DivCoef& q, const double* rho_new, const double* Q,
             double* const rf[3], double* dvg, hipStream_t s) {
  hipLaunchKernelGGL(k_duct_div, grid3d(d), dim3(64, 4), 0, s, d, q, rho_new, Q, rf[0], rf[1],
                     rf[2], dvg);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int duct_planes(const Dev& d, const double* ru_new_x, const double* rfx, double* out2,
                hipStream_t s) {
  hipLaunchKernelGGL(k_duct_planes, dim3(1), dim3(256), 0, s, d, ru_new_x, rfx, out2);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int duct_scale(const Dev& d, double ratio, int last, double* ru_new_x, double* ru_np1_x,
               hipStream_t s) {
  hipLaunchKernelGGL(k_duct_scale, dim3(nblocks((long)d.Ny * d.Nz, 256)), dim3(256), 0, s, d,
                     ratio, last, ru_new_x, ru_np1_x);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int duct_rhs(const Dev& d, double sc, const double* const ru_new[3], const double* const rf[3],
             const double* dvg, double* rhs, hipStream_t s) {
  hipLaunchKernelGGL(k_duct_rhs, grid3d(d), dim3(64, 4), 0, s, d, sc, ru_new[0], ru_new[1],
                     ru_new[2], rf[0], rf[1], rf[2], dvg, rhs);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

// The duct inlet of a coupled run reads the box's first interior slice
// (grid.cpp:352-363, 485-503, bs = 1): plane q of `out` ([8][Nz][Ny], the
// order rho_new, rho_np1, RU_new x,y,z, RU_np1 x,y,z) is src[q] of the box at
// x index xi[q] over the whole ghosted (j, k) plane.  One lane per (j, k).
struct PlaneSrc {
  const double* p[8];
  int xi[8];
};

__global__ void __launch_bounds__(256) k_pair_planes(Dev b, PlaneSrc s, double* __restrict__ out) {
  const long m = (long)b.Ny * b.Nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int j = (int)(t % b.Ny), k = (int)(t / b.Ny);
#pragma unroll
  for (int q = 0; q < 8; q++) out[q * m + t] = s.p[q][at(b, s.xi[q], j, k)];
}

int pair_planes(const Dev& b, const double* const src[8], const int xi[8], double* out,
                hipStream_t s) {
  PlaneSrc ps;
  for (int q = 0; q < 8; q++) {
    if (xi[q] < 0 || xi[q] >= b.Nx) return FR_EINVAL;
    ps.p[q] = src[q];
    ps.xi[q] = xi[q];
  }
  hipLaunchKernelGGL(k_pair_planes, dim3(nblocks((long)b.Ny * b.Nz, 256)), dim3(256), 0, s, b,
                     ps, out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int duct_update_ru_wp(const Dev& d, const BoxCoef& cf, const double* P, double* const ru_np1[3],
                      double* const ru_new[3], hipStream_t s) {
  hipLaunchKernelGGL(k_duct_update_ru_wp, grid3d(d), dim3(64, 4), 0, s, d, cf, P, ru_np1[0],
                     ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

}  // namespace fr
