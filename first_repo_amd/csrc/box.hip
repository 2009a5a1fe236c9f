// Fused RK4-substep kernels of the triply periodic box (grid.cpp:327-657).
//
// Each kernel evaluates, per ghosted cell, exactly the floating-point
// expression tree the reference builds from its sequence of tensor0/tensor1
// operators (every intermediate is recomputed inline from the same operands,
// so no value changes), and writes every cell including the ghosts (the
// ghost of a box field is the value at its periodic image).  Inputs must
// have valid periodic ghosts, which every kernel here maintains.
#include "fr_internal.hpp"

namespace fr {

#define FR_CELL_PROLOGUE                                          \
  const uint3 bk = xcd_block();                                   \
  const int i = bk.x * 64 + threadIdx.x;                          \
  const int j = bk.y * 4 + threadIdx.y;                           \
  const int k = bk.z;                                             \
  if (i >= d.Nx || j >= d.Ny) return;                            \
  const long n = at(d, i, j, k);                                  \
  const long c = at(d, img(i, d.Nx), img(j, d.Ny), img(k, d.Nz)); \
  const long sx = 1, sy = d.Nx, sz = d.NxNy;                      \
  (void)sx; (void)sy; (void)sz;

static dim3 grid3(const Dev& d) { return dim3(nblocks(d.Nx, 64), nblocks(d.Ny, 4), d.Nz); }

// grid::Update_Rho (grid.cpp:327-333) fused with U = RU_int / Rho_face
// (grid.cpp:662, the first line of grid::Update_Particle).
__global__ void __launch_bounds__(256) k_update_rho(
    Dev d, BoxCoef cf, const double* __restrict__ rux, const double* __restrict__ ruy,
    const double* __restrict__ ruz, const double* __restrict__ rho,
    double* __restrict__ rho_np1, double* __restrict__ rho_new,
    const double* __restrict__ rfx, const double* __restrict__ rfy,
    const double* __restrict__ rfz, double* __restrict__ ux, double* __restrict__ uy,
    double* __restrict__ uz) {
  FR_CELL_PROLOGUE
  // RHS_Rho = Div_F2C(RU_int)
  const double rhs = (rux[c + sx] - rux[c]) * d.idx + (ruy[c + sy] - ruy[c]) * d.idy +
                     (ruz[c + sz] - ruz[c]) * d.idz;
  const double np1 = rho_np1[n] + cf.cpost_neg * rhs;
  rho_np1[n] = np1;
  rho_new[n] = cf.last ? np1 : 1 * rho[n] + cf.cpre_neg * rhs;
  ux[n] = rux[n] / rfx[n];
  uy[n] = ruy[n] / rfy[n];
  uz[n] = ruz[n] / rfz[n];
}

// Div_F2C(U) at cell q (any q with q+e in range), grid.cpp:390.
__device__ __forceinline__ double divu(const Dev& d, const double* ux, const double* uy,
                                       const double* uz, long q) {
  return (ux[q + 1] - ux[q]) * d.idx + (uy[q + d.Nx] - uy[q]) * d.idy +
         (uz[q + d.NxNy] - uz[q]) * d.idz;
}

// Convection of momentum component m (grid.cpp:394-411): the flux
// F_d = I_d^{C2F}(U_m) * I_m^{C2F}(RU_int_d) at faces, then Div_F2C(F).
template <int M>
__device__ __forceinline__ double conv(const Dev& d, const double* __restrict__ um,
                                       const double* __restrict__ rux,
                                       const double* __restrict__ ruy,
                                       const double* __restrict__ ruz, long c) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  const long sm = M == 0 ? sx : (M == 1 ? sy : sz);
  // F_x at q: ((U_m(q)+U_m(q-sx))*0.5) * ((RU_x(q)+RU_x(q-sm))*0.5)
  auto Fx = [&](long q) { return ((um[q] + um[q - sx]) * 0.5) * ((rux[q] + rux[q - sm]) * 0.5); };
  auto Fy = [&](long q) { return ((um[q] + um[q - sy]) * 0.5) * ((ruy[q] + ruy[q - sm]) * 0.5); };
  auto Fz = [&](long q) { return ((um[q] + um[q - sz]) * 0.5) * ((ruz[q] + ruz[q - sm]) * 0.5); };
  return (Fx(c + sx) - Fx(c)) * d.idx + (Fy(c + sy) - Fy(c)) * d.idy +
         (Fz(c + sz) - Fz(c)) * d.idz;
}

// grid::Update_RU_WOP (grid.cpp:387-425), all three components.
template <int GRAV>
__global__ void __launch_bounds__(256) k_update_ru_wop(
    Dev d, BoxCoef cf, WopCoef w, const double* __restrict__ ux,
    const double* __restrict__ uy, const double* __restrict__ uz,
    const double* __restrict__ rux, const double* __restrict__ ruy,
    const double* __restrict__ ruz, const double* __restrict__ sxp,
    const double* __restrict__ syp, const double* __restrict__ szp,
    const double* __restrict__ ru0x, const double* __restrict__ ru0y,
    const double* __restrict__ ru0z, double* __restrict__ np1x, double* __restrict__ np1y,
    double* __restrict__ np1z, double* __restrict__ newx, double* __restrict__ newy,
    double* __restrict__ newz, const double* __restrict__ rfx,
    const double* __restrict__ rfy, const double* __restrict__ rfz) {
  FR_CELL_PROLOGUE
  const double dv = divu(d, ux, uy, uz, c);
  const double g[3] = {(dv - divu(d, ux, uy, uz, c - sx)) * d.idx,
                       (dv - divu(d, ux, uy, uz, c - sy)) * d.idy,
                       (dv - divu(d, ux, uy, uz, c - sz)) * d.idz};
  const double* U[3] = {ux, uy, uz};
  const double* Sd[3] = {sxp, syp, szp};
  const double* R0[3] = {ru0x, ru0y, ru0z};
  const double* RI[3] = {rux, ruy, ruz};
  double* N1[3] = {np1x, np1y, np1z};
  double* NW[3] = {newx, newy, newz};
  const double* RF[3] = {rfx, rfy, rfz};
  const double gv[3] = {w.gx, w.gy, w.gz};
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const double* u = U[m];
    const double lap = (u[c + sx] - 2 * u[c] + u[c - sx]) * d.idx2 +
                       (u[c + sy] - 2 * u[c] + u[c - sy]) * d.idy2 +
                       (u[c + sz] - 2 * u[c] + u[c - sz]) * d.idz2;
    double r = w.a_s * Sd[m][c] + w.b_g * g[m] + w.c_l * lap;
    const double cv = m == 0 ? conv<0>(d, u, rux, ruy, ruz, c)
                     : m == 1 ? conv<1>(d, u, rux, ruy, ruz, c)
                              : conv<2>(d, u, rux, ruy, ruz, c);
    r = r - cv;
    r = r + w.A * RI[m][c];
    if (m == 0) r = r - w.sx_forc;
    if (GRAV) r = r + (RF[m][c] - w.rho0) * gv[m];
    const double np1 = N1[m][n] + cf.cpost * r;
    N1[m][n] = np1;
    NW[m][n] = cf.last ? np1 : 1 * R0[m][n] + cf.cpre * r;
  }
}

// grid::Compute_Div_U_new (grid.cpp:556-563) + grid::Compute_RHS_Pois
// (grid.cpp:629-637) up to the zero-mean shift: writes Rho_face = I_C2F(
// Rho_new) and RHS_Pois = (Div_F2C(RU_new/Rho_face) - divergence)/(dt*pre).
__global__ void __launch_bounds__(256) k_div_rhs(
    Dev d, DivCoef q, const double* __restrict__ rn, const double* __restrict__ Q,
    const double* __restrict__ rnx, const double* __restrict__ rny,
    const double* __restrict__ rnz, double* __restrict__ rfx, double* __restrict__ rfy,
    double* __restrict__ rfz, double* __restrict__ rhs) {
  FR_CELL_PROLOGUE
  auto f = [&](long p) { return 1. / rn[p]; };
  const double fc = f(c);
  const double del2 = (f(c + sx) - 2 * fc + f(c - sx)) * d.idx2 +
                      (f(c + sy) - 2 * fc + f(c - sy)) * d.idy2 +
                      (f(c + sz) - 2 * fc + f(c - sz)) * d.idz2;
  const double qq = q.cooling ? Q[c] - q.qmean : Q[c];
  const double dvg = q.kc * del2 + q.bq * qq;
  // Rho_face = I_C2F(Rho_new) at the three lower faces and the upper ones
  const double fx0 = (rn[c] + rn[c - sx]) * 0.5, fx1 = (rn[c + sx] + rn[c]) * 0.5;
  const double fy0 = (rn[c] + rn[c - sy]) * 0.5, fy1 = (rn[c + sy] + rn[c]) * 0.5;
  const double fz0 = (rn[c] + rn[c - sz]) * 0.5, fz1 = (rn[c + sz] + rn[c]) * 0.5;
  rfx[n] = fx0;
  rfy[n] = fy0;
  rfz[n] = fz0;
  const double D = (rnx[c + sx] / fx1 - rnx[c] / fx0) * d.idx +
                   (rny[c + sy] / fy1 - rny[c] / fy0) * d.idy +
                   (rnz[c + sz] / fz1 - rnz[c] / fz0) * d.idz;
  double r = D - dvg;
  r = r * q.s;
  rhs[n] = r;
}

// grid::Update_RU_WP (grid.cpp:651-657)
__global__ void __launch_bounds__(256) k_update_ru_wp(
    Dev d, BoxCoef cf, const double* __restrict__ P, double* __restrict__ np1x,
    double* __restrict__ np1y, double* __restrict__ np1z, double* __restrict__ newx,
    double* __restrict__ newy, double* __restrict__ newz) {
  FR_CELL_PROLOGUE
  const double gx = (P[c] - P[c - sx]) * d.idx;
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

int box_update_rho(const Dev& d, const BoxCoef& cf, const double* const ru_int[3],
                   const double* rho, double* rho_np1, double* rho_new,
                   const double* const rf[3], double* const u[3], hipStream_t s) {
  hipLaunchKernelGGL(k_update_rho, grid3(d), dim3(64, 4), 0, s, d, cf, ru_int[0], ru_int[1],
                     ru_int[2], rho, rho_np1, rho_new, rf[0], rf[1], rf[2], u[0], u[1], u[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int box_update_ru_wop(const Dev& d, const BoxCoef& cf, const WopCoef& w,
                      const double* const u[3], const double* const ru_int[3],
                      const double* const S[3], const double* const ru[3],
                      double* const ru_np1[3], double* const ru_new[3],
                      const double* const rf[3], hipStream_t s) {
  if (w.gravity)
    hipLaunchKernelGGL(k_update_ru_wop<1>, grid3(d), dim3(64, 4), 0, s, d, cf, w, u[0], u[1],
                       u[2], ru_int[0], ru_int[1], ru_int[2], S[0], S[1], S[2], ru[0], ru[1],
                       ru[2], ru_np1[0], ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2],
                       rf[0], rf[1], rf[2]);
  else
    hipLaunchKernelGGL(k_update_ru_wop<0>, grid3(d), dim3(64, 4), 0, s, d, cf, w, u[0], u[1],
                       u[2], ru_int[0], ru_int[1], ru_int[2], S[0], S[1], S[2], ru[0], ru[1],
                       ru[2], ru_np1[0], ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2],
                       rf[0], rf[1], rf[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int box_div_rhs(const Dev& d, const DivCoef& q, const double* rho_new, const double* Q,
                const double* const ru_new[3], double* const rf[3], double* rhs,
                hipStream_t s) {
  hipLaunchKernelGGL(k_div_rhs, grid3(d), dim3(64, 4), 0, s, d, q, rho_new, Q, ru_new[0],
                     ru_new[1], ru_new[2], rf[0], rf[1], rf[2], rhs);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int box_update_ru_wp(const Dev& d, const BoxCoef& cf, const double* P,
                     double* const ru_np1[3], double* const ru_new[3], hipStream_t s) {
  hipLaunchKernelGGL(k_update_ru_wp, grid3(d), dim3(64, 4), 0, s, d, cf, P, ru_np1[0],
                     ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

}  // namespace fr
