// Fused RK4-substep kernels of the triply periodic box (grid.cpp:327-657).
//
// Each kernel evaluates, per ghosted cell, exactly the floating-point
// expression tree the reference builds from its sequence of tensor0/tensor1
// operators (every intermediate is recomputed inline from the same operands,
// so no value changes), and writes every cell including the ghosts (the
// ghost of a box field is the value at its periodic image).  Inputs must
// have valid periodic ghosts, which every kernel here maintains.
#include <type_traits>

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
// (grid.cpp:662, the first line of grid::Update_Particle).  SET (last stage
// of a full step only): the end-of-step copy Rho = Rho_np1 of
// grid::TimeAdvance (grid.cpp:317) is made here, into rho_set; the last
// stage does not read Rho.  V (the BoxCoef flags, as template arguments so
// that each form compiles to straight-line code): RHO_PLAIN; RHO_BASE (first
// stage, Rho stands in for Rho_np1); RHO_SET (SET above); RHO_SET_LEAN (SET,
// and Rho_np1 itself is not written: it equals Rho from there on).  FI: rfx
// holds Rho_int and Rho_face = I_C2F(Rho_int) (the divergence kernel of the
// previous stage formed it from the same values, which the stage exchange
// has moved into Rho_int): the face values are formed here as that kernel
// forms them (at the image cell, in double, rounded to T); rfy, rfz unused.
enum { RHO_PLAIN, RHO_BASE, RHO_SET, RHO_SET_LEAN };
template <typename T, int V, bool FI>
__global__ void __launch_bounds__(256) k_update_rho(
    Dev d, BoxCoef cf, const T* __restrict__ rux, const T* __restrict__ ruy,
    const T* __restrict__ ruz, const T* __restrict__ rho, T* __restrict__ rho_np1,
    T* __restrict__ rho_new, const T* __restrict__ rfx, const T* __restrict__ rfy,
    const T* __restrict__ rfz, T* __restrict__ ux, T* __restrict__ uy, T* __restrict__ uz,
    T* __restrict__ rho_set) {
  FR_CELL_PROLOGUE
  // RHS_Rho = Div_F2C(RU_int)
  const T rhs = (rux[c + sx] - rux[c]) * (T)d.idx + (ruy[c + sy] - ruy[c]) * (T)d.idy +
                (ruz[c + sz] - ruz[c]) * (T)d.idz;
  constexpr bool SET = V >= RHO_SET;
  T base;
  if constexpr (V == RHO_BASE)
    base = rho[n];
  else
    base = rho_np1[n];
  const T np1 = base + (T)cf.cpost_neg * rhs;
  if constexpr (V != RHO_SET_LEAN) rho_np1[n] = np1;
  if constexpr (SET) {
    rho_new[n] = np1;
    rho_set[n] = np1;
  } else {
    rho_new[n] = cf.last ? np1 : 1 * rho[n] + (T)cf.cpre_neg * rhs;
  }
  if constexpr (FI) {
    const double r0 = (double)rfx[c];
    ux[n] = rux[n] / (T)((r0 + (double)rfx[c - sx]) * 0.5);
    uy[n] = ruy[n] / (T)((r0 + (double)rfx[c - sy]) * 0.5);
    uz[n] = ruz[n] / (T)((r0 + (double)rfx[c - sz]) * 0.5);
  } else {
    ux[n] = rux[n] / rfx[n];
    uy[n] = ruy[n] / rfy[n];
    uz[n] = ruz[n] / rfz[n];
  }
}

// Rho_face = I_C2F(rn) at every ghosted cell, the values k_div_rhs stores
// (at the image cell, in double, rounded to T): the faces of a step that
// left them unstored, for a reader outside the kernels that rebuild them.
template <typename T>
__global__ void __launch_bounds__(256) k_faces_of(Dev d, const T* __restrict__ rn,
                                                  T* __restrict__ rfx, T* __restrict__ rfy,
                                                  T* __restrict__ rfz) {
  FR_CELL_PROLOGUE
  const double r0 = (double)rn[c];
  rfx[n] = (T)((r0 + (double)rn[c - sx]) * 0.5);
  rfy[n] = (T)((r0 + (double)rn[c - sy]) * 0.5);
  rfz[n] = (T)((r0 + (double)rn[c - sz]) * 0.5);
}

// Div_F2C(U) at cell q (any q with q+e in range), grid.cpp:390.
template <typename T>
__device__ __forceinline__ T divu(const Dev& d, const T* ux, const T* uy, const T* uz, long q) {
  return (ux[q + 1] - ux[q]) * (T)d.idx + (uy[q + d.Nx] - uy[q]) * (T)d.idy +
         (uz[q + d.NxNy] - uz[q]) * (T)d.idz;
}

// Convection of momentum component m (grid.cpp:394-411): the flux
// F_d = I_d^{C2F}(U_m) * I_m^{C2F}(RU_int_d) at faces, then Div_F2C(F).
template <int M, typename T>
__device__ __forceinline__ T conv(const Dev& d, const T* __restrict__ um, const T* __restrict__ rux,
                                  const T* __restrict__ ruy, const T* __restrict__ ruz, long c) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  const long sm = M == 0 ? sx : (M == 1 ? sy : sz);
  const T h = (T)0.5;
  // F_x at q: ((U_m(q)+U_m(q-sx))*0.5) * ((RU_x(q)+RU_x(q-sm))*0.5)
  auto Fx = [&](long q) { return ((um[q] + um[q - sx]) * h) * ((rux[q] + rux[q - sm]) * h); };
  auto Fy = [&](long q) { return ((um[q] + um[q - sy]) * h) * ((ruy[q] + ruy[q - sm]) * h); };
  auto Fz = [&](long q) { return ((um[q] + um[q - sz]) * h) * ((ruz[q] + ruz[q - sm]) * h); };
  return (Fx(c + sx) - Fx(c)) * (T)d.idx + (Fy(c + sy) - Fy(c)) * (T)d.idy +
         (Fz(c + sz) - Fz(c)) * (T)d.idz;
}

// grid::Update_RU_WOP (grid.cpp:387-425), all three components.
template <int GRAV, typename T>
__global__ void __launch_bounds__(256) k_update_ru_wop(
    Dev d, BoxCoef cf, WopCoef w, const T* __restrict__ ux, const T* __restrict__ uy,
    const T* __restrict__ uz, const T* __restrict__ rux, const T* __restrict__ ruy,
    const T* __restrict__ ruz, const T* __restrict__ sxp, const T* __restrict__ syp,
    const T* __restrict__ szp, const T* __restrict__ ru0x, const T* __restrict__ ru0y,
    const T* __restrict__ ru0z, T* __restrict__ np1x, T* __restrict__ np1y, T* __restrict__ np1z,
    T* __restrict__ newx, T* __restrict__ newy, T* __restrict__ newz, const T* __restrict__ rfx,
    const T* __restrict__ rfy, const T* __restrict__ rfz) {
  FR_CELL_PROLOGUE
  const T dv = divu(d, ux, uy, uz, c);
  const T g[3] = {(dv - divu(d, ux, uy, uz, c - sx)) * (T)d.idx,
                  (dv - divu(d, ux, uy, uz, c - sy)) * (T)d.idy,
                  (dv - divu(d, ux, uy, uz, c - sz)) * (T)d.idz};
  const T* U[3] = {ux, uy, uz};
  const T* Sd[3] = {sxp, syp, szp};
  const T* R0[3] = {ru0x, ru0y, ru0z};
  const T* RI[3] = {rux, ruy, ruz};
  T* N1[3] = {np1x, np1y, np1z};
  T* NW[3] = {newx, newy, newz};
  const T* RF[3] = {rfx, rfy, rfz};
  const T gv[3] = {(T)w.gx, (T)w.gy, (T)w.gz};
#pragma unroll
  for (int m = 0; m < 3; m++) {
    const T* u = U[m];
    const T lap = (u[c + sx] - 2 * u[c] + u[c - sx]) * (T)d.idx2 +
                  (u[c + sy] - 2 * u[c] + u[c - sy]) * (T)d.idy2 +
                  (u[c + sz] - 2 * u[c] + u[c - sz]) * (T)d.idz2;
    T r = (T)w.a_s * Sd[m][c] + (T)w.b_g * g[m] + (T)w.c_l * lap;
    const T cv = m == 0 ? conv<0>(d, u, rux, ruy, ruz, c)
                : m == 1 ? conv<1>(d, u, rux, ruy, ruz, c)
                         : conv<2>(d, u, rux, ruy, ruz, c);
    r = r - cv;
    r = r + (T)w.A * RI[m][c];
    if (m == 0) r = r - (T)w.sx_forc;
    if (GRAV) r = r + (RF[m][c] - (T)w.rho0) * gv[m];
    const T np1 = N1[m][n] + (T)cf.cpost * r;
    N1[m][n] = np1;
    NW[m][n] = cf.last ? np1 : 1 * R0[m][n] + (T)cf.cpre * r;
  }
}

// The same update, z-marching: a workgroup owns a 64 x 8 column of interior
// cells over KZ planes and keeps planes k-1, k, k+1 of the six stencil
// arrays (U and RU_int, with a one-cell x/y halo) in an LDS ring; plane k+2
// is fetched into registers while plane k is computed, so every U/RU_int
// value is read from HBM about once.  The expressions are those of
// k_update_ru_wop with each u[c + offset] read from the ring.  Only interior
// cells are written; the caller restores the periodic ghosts of the six
// outputs afterwards (a ghost's value is its image's, as k_update_ru_wop
// computes it).
constexpr int WX = 64, WY = 8, WKZ = 16, WHX = WX + 2, WHY = WY + 2;
constexpr int WPL = 6 * WHY * WHX;               // values per plane of the ring
constexpr int WPT = (WPL + WX * WY - 1) / (WX * WY);  // per thread

// MODE (the BoxCoef flags, as template arguments): WOP_PLAIN; WOP_BASE
// (first stage: RU stands in for RU_np1, which is not read); WOP_LEAN (last
// stage of a full step: RU_np1 is not written, RU_new takes its value, and
// RU, which the last stage does not use, is not read).
enum { WOP_PLAIN, WOP_BASE, WOP_LEAN };
template <int GRAV, int MODE, typename T>
__global__ void __launch_bounds__(WX * WY) k_update_ru_wop_z(
    Dev d, BoxCoef cf, WopCoef w, const T* __restrict__ ux, const T* __restrict__ uy,
    const T* __restrict__ uz, const T* __restrict__ rux, const T* __restrict__ ruy,
    const T* __restrict__ ruz, const T* __restrict__ sxp, const T* __restrict__ syp,
    const T* __restrict__ szp, const T* __restrict__ ru0x, const T* __restrict__ ru0y,
    const T* __restrict__ ru0z, T* __restrict__ np1x, T* __restrict__ np1y, T* __restrict__ np1z,
    T* __restrict__ newx, T* __restrict__ newy, T* __restrict__ newz, const T* __restrict__ rfx,
    const T* __restrict__ rfy, const T* __restrict__ rfz) {
  __shared__ T ring[3][6][WHY][WHX];
  const T* src[6] = {ux, uy, uz, rux, ruy, ruz};
  const uint3 bk = xcd_block();
  const int i0 = 1 + bk.x * WX, j0 = 1 + bk.y * WY, k0 = 1 + bk.z * WKZ;
  const int k1 = min(d.Nz - 1, k0 + WKZ);
  const int tid = threadIdx.y * WX + threadIdx.x;
  // this thread's share of a plane: value t = tid + WX*WY q -> (array, y, x)
  auto fetch = [&](int kk, T* reg) {
#pragma unroll
    for (int q = 0; q < WPT; q++) {
      const int t = tid + q * WX * WY;
      T v = (T)0;
      if (t < WPL) {
        const int a = t / (WHY * WHX), r = t % (WHY * WHX);
        const int yy = j0 - 1 + r / WHX, xx = i0 - 1 + r % WHX;
        if (xx < d.Nx && yy < d.Ny) v = src[a][at(d, xx, yy, kk)];
      }
      reg[q] = v;
    }
  };
  auto store = [&](int slot, const T* reg) {
#pragma unroll
    for (int q = 0; q < WPT; q++) {
      const int t = tid + q * WX * WY;
      if (t < WPL) (&ring[slot][0][0][0])[t] = reg[q];
    }
  };
  T reg[WPT];
  fetch(k0 - 1, reg);
  store((k0 - 1) % 3, reg);
  fetch(k0, reg);
  store(k0 % 3, reg);
  if (k0 < k1) fetch(k0 + 1, reg);
  const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
  const bool live = i < d.Nx - 1 && j < d.Ny - 1;
  const T* Sd[3] = {sxp, syp, szp};
  const T* R0[3] = {ru0x, ru0y, ru0z};
  const T* RI[3] = {rux, ruy, ruz};
  T* N1[3] = {np1x, np1y, np1z};
  T* NW[3] = {newx, newy, newz};
  const T* RF[3] = {rfx, rfy, rfz};
  const T gv[3] = {(T)w.gx, (T)w.gy, (T)w.gz};
  const int tx = threadIdx.x + 1, ty = threadIdx.y + 1;
  // per-cell operands of the next plane, fetched one iteration ahead:
  // S, RU_int, RU, RU_np1 (x, y, z each) and Rho_face under gravity
  constexpr int NOP = GRAV ? 15 : 12;
  T op[NOP], opn[NOP];
  auto fetch_ops = [&](int kk, T* o) {
    const long c = at(d, i, j, kk);
#pragma unroll
    for (int m = 0; m < 3; m++) {
      o[m] = Sd[m][c];
      o[3 + m] = RI[m][c];
      if constexpr (MODE != WOP_LEAN) o[6 + m] = R0[m][c];  // the last stage does not use RU
      if constexpr (MODE != WOP_BASE) o[9 + m] = N1[m][c];
      if constexpr (GRAV != 0) o[12 + m] = RF[m][c];
    }
  };
  if (live) fetch_ops(k0, op);
  for (int k = k0; k < k1; k++) {
    __syncthreads();  // plane k-2's slot is free
    store((k + 1) % 3, reg);
    __syncthreads();
    if (k + 1 < k1) {
      fetch(k + 2, reg);
      if (live) fetch_ops(k + 1, opn);
    }
    if (!live) continue;
    const int sl[3] = {(k + 2) % 3, k % 3, (k + 1) % 3};  // slots of k-1, k, k+1
    auto A = [&](int a, int ox, int oy, int oz) -> T {
      return ring[sl[oz + 1]][a][ty + oy][tx + ox];
    };
    auto dv_at = [&](int ox, int oy, int oz) -> T {
      return (A(0, ox + 1, oy, oz) - A(0, ox, oy, oz)) * (T)d.idx +
             (A(1, ox, oy + 1, oz) - A(1, ox, oy, oz)) * (T)d.idy +
             (A(2, ox, oy, oz + 1) - A(2, ox, oy, oz)) * (T)d.idz;
    };
    const long c = at(d, i, j, k);
    const T dv = dv_at(0, 0, 0);
    const T g[3] = {(dv - dv_at(-1, 0, 0)) * (T)d.idx, (dv - dv_at(0, -1, 0)) * (T)d.idy,
                    (dv - dv_at(0, 0, -1)) * (T)d.idz};
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const T lap = (A(m, 1, 0, 0) - 2 * A(m, 0, 0, 0) + A(m, -1, 0, 0)) * (T)d.idx2 +
                    (A(m, 0, 1, 0) - 2 * A(m, 0, 0, 0) + A(m, 0, -1, 0)) * (T)d.idy2 +
                    (A(m, 0, 0, 1) - 2 * A(m, 0, 0, 0) + A(m, 0, 0, -1)) * (T)d.idz2;
      T r = (T)w.a_s * op[m] + (T)w.b_g * g[m] + (T)w.c_l * lap;
      // convection: F_e(o) = ((U_m(o) + U_m(o - e)) * 0.5) * ((RU_e(o) + RU_e(o - e_m)) * 0.5)
      const int mx = m == 0, my = m == 1, mz = m == 2;
      const T h = (T)0.5;
      auto F = [&](int e, int ox, int oy, int oz) -> T {
        const int ex = e == 0, ey = e == 1, ez = e == 2;
        return ((A(m, ox, oy, oz) + A(m, ox - ex, oy - ey, oz - ez)) * h) *
               ((A(3 + e, ox, oy, oz) + A(3 + e, ox - mx, oy - my, oz - mz)) * h);
      };
      const T cv = (F(0, 1, 0, 0) - F(0, 0, 0, 0)) * (T)d.idx +
                   (F(1, 0, 1, 0) - F(1, 0, 0, 0)) * (T)d.idy +
                   (F(2, 0, 0, 1) - F(2, 0, 0, 0)) * (T)d.idz;
      r = r - cv;
      r = r + (T)w.A * op[3 + m];
      if (m == 0) r = r - (T)w.sx_forc;
      if constexpr (GRAV != 0) r = r + (op[12 + m] - (T)w.rho0) * gv[m];
      const T np1 = (MODE == WOP_BASE ? op[6 + m] : op[9 + m]) + (T)cf.cpost * r;
      if constexpr (MODE != WOP_LEAN) N1[m][c] = np1;
      if constexpr (MODE == WOP_LEAN)
        NW[m][c] = np1;
      else
        NW[m][c] = cf.last ? np1 : 1 * op[6 + m] + (T)cf.cpre * r;
    }
#pragma unroll
    for (int q = 0; q < NOP; q++) op[q] = opn[q];
  }
}

// grid::Compute_Div_U_new (grid.cpp:556-563) + grid::Compute_RHS_Pois
// (grid.cpp:629-637) up to the zero-mean shift: writes Rho_face = I_C2F(
// Rho_new) and RHS_Pois = (Div_F2C(RU_new/Rho_face) - divergence)/(dt*pre)
// at every ghosted cell (a ghost takes its image's values).  The right-hand
// side is formed in double from fields of either precision: it is the small
// difference of large face fluxes, and the pressure solve iterates to 1e-10
// on it.
//
// A workgroup is a 64x4 column tile marching through DKZ planes: Rho_new at
// the planes below and above, its reciprocals, the upper z face value and
// the z flux RU_new_z/Rho_face_z of the upper face are the next plane's
// lower ones, so each is loaded or formed once (the same expressions, so
// the same values).  The march reloads where the image plane sequence jumps
// (the periodic ghost planes).  With part: per plane, the tile's sum of
// RHS_Pois over its interior cells into part[(k*gy + ty)*gx + tx], summed
// later in a fixed order.
constexpr int DKZ = 8;

template <typename T, bool SF>
__global__ void __launch_bounds__(256) k_div_rhs(Dev d, DivCoef q, const T* __restrict__ rn,
                                                 const T* __restrict__ Q,
                                                 const T* __restrict__ rnx,
                                                 const T* __restrict__ rny,
                                                 const T* __restrict__ rnz, T* __restrict__ rfx,
                                                 T* __restrict__ rfy, T* __restrict__ rfz,
                                                 double* __restrict__ rhs,
                                                 double* __restrict__ part) {
  __shared__ double wsum[2][4];
  const uint3 bk = xcd_block();
  const int i = bk.x * 64 + threadIdx.x, j = bk.y * 4 + threadIdx.y;
  const bool live = i < d.Nx && j < d.Ny;
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  const long cxy = live ? (long)img(j, d.Ny) * d.Nx + img(i, d.Nx) : 0;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  auto v = [](const T* a, long p) { return (double)a[p]; };
  const double h = 0.5;
  double rm = 0, rc = 0, rp = 0;  // Rho_new at c - sz, c, c + sz
  double im = 0, ic = 0, ip = 0;  // their reciprocals
  double fz1 = 0, qz1 = 0;        // upper z face value and RU_new_z / face there
  int prev = -2;
  const int k0 = bk.z * DKZ, k1 = min(k0 + DKZ, d.Nz);
  for (int k = k0; k < k1; k++) {
    const int ck = img(k, d.Nz);
    const long c = (long)ck * sz + cxy;
    double fz0 = 0, qz0 = 0;
    if (ck == prev + 1) {
      rm = rc;
      rc = rp;
      im = ic;
      ic = ip;
      fz0 = fz1;
      qz0 = qz1;
      if (live) {
        rp = v(rn, c + sz);
        ip = 1. / rp;
        fz1 = (rp + rc) * h;
        qz1 = v(rnz, c + sz) / fz1;
      }
    } else if (live) {
      rm = v(rn, c - sz);
      rc = v(rn, c);
      rp = v(rn, c + sz);
      im = 1. / rm;
      ic = 1. / rc;
      ip = 1. / rp;
      fz0 = (rc + rm) * h;
      qz0 = v(rnz, c) / fz0;
      fz1 = (rp + rc) * h;
      qz1 = v(rnz, c + sz) / fz1;
    }
    prev = ck;
    double r = 0.0;
    if (live) {
      const long n = at(d, i, j, k);
      const double rxp = v(rn, c + sx), rxm = v(rn, c - sx);
      const double ryp = v(rn, c + sy), rym = v(rn, c - sy);
      const double fc = ic;
      const double del2 = (1. / rxp - 2 * fc + 1. / rxm) * d.idx2 +
                          (1. / ryp - 2 * fc + 1. / rym) * d.idy2 +
                          (ip - 2 * fc + im) * d.idz2;
      const double qq = q.cooling ? v(Q, c) - q.qmean : v(Q, c);
      const double dvg = q.kc * del2 + q.bq * qq;
      // Rho_face = I_C2F(Rho_new) at the three lower faces and the upper ones
      const double fx0 = (rc + rxm) * h, fx1 = (rxp + rc) * h;
      const double fy0 = (rc + rym) * h, fy1 = (ryp + rc) * h;
      if constexpr (SF) {
        rfx[n] = (T)fx0;
        rfy[n] = (T)fy0;
        rfz[n] = (T)fz0;
      }
      const double D = (v(rnx, c + sx) / fx1 - v(rnx, c) / fx0) * d.idx +
                       (v(rny, c + sy) / fy1 - v(rny, c) / fy0) * d.idy + (qz1 - qz0) * d.idz;
      r = D - dvg;
      r = r * q.s;
      rhs[n] = r;
    }
    if (part) {
      const bool in = i >= 1 && i <= d.Nx - 2 && j >= 1 && j <= d.Ny - 2 && k >= 1 &&
                      k <= d.Nz - 2;
      double w = in ? r : 0.0;
      for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
      if ((tid & 63) == 0) wsum[k & 1][tid >> 6] = w;
      __syncthreads();
      if (tid == 0)
        part[((long)k * gridDim.y + bk.y) * gridDim.x + bk.x] =
            (wsum[k & 1][0] + wsum[k & 1][1]) + (wsum[k & 1][2] + wsum[k & 1][3]);
    }
  }
}

// grid::Update_RU_WP (grid.cpp:651-657)
// The pressure gradient is formed from the FP64 pressure in double.
template <typename T>
__global__ void __launch_bounds__(256) k_update_ru_wp(
    Dev d, BoxCoef cf, const double* __restrict__ P, T* __restrict__ np1x,
    T* __restrict__ np1y, T* __restrict__ np1z, T* __restrict__ newx, T* __restrict__ newy,
    T* __restrict__ newz) {
  FR_CELL_PROLOGUE
  const T gx = (T)((P[c] - P[c - sx]) * d.idx);
  const T gy = (T)((P[c] - P[c - sy]) * d.idy);
  const T gz = (T)((P[c] - P[c - sz]) * d.idz);
  const T cpn = (T)cf.cpost_neg, cqn = (T)cf.cpre_neg;
  const T ax = np1x[n] + cpn * gx;
  const T ay = np1y[n] + cpn * gy;
  const T az = np1z[n] + cpn * gz;
  np1x[n] = ax;
  np1y[n] = ay;
  np1z[n] = az;
  if (cf.last) {
    newx[n] = ax;
    newy[n] = ay;
    newz[n] = az;
  } else {
    newx[n] = newx[n] + cqn * gx;
    newy[n] = newy[n] + cqn * gy;
    newz[n] = newz[n] + cqn * gz;
  }
}

// The last stage of a full box step (BoxCoef::lean): RU_new holds the stage's
// RU_np1 (the momentum update wrote it there only) and takes the update of
// k_update_ru_wp in place, the same expressions.  A workgroup is a 64x4
// column tile through PKZ planes; it also sums each component over the
// interior cells it covers (each thread in plane order, then the workgroup
// in a fixed tree) into part[m * gridsize + block], which sum_parts reduces
// to the means that grid.cpp:565-577 removes.
constexpr int PKZ = 8;

template <typename T>
__global__ void __launch_bounds__(256) k_update_ru_wp_lean(Dev d, BoxCoef cf,
                                                           const double* __restrict__ P,
                                                           T* __restrict__ nx, T* __restrict__ ny,
                                                           T* __restrict__ nz,
                                                           double* __restrict__ part) {
  __shared__ double ws[3][4];
  const uint3 bk = xcd_block();
  const int i = bk.x * 64 + threadIdx.x, j = bk.y * 4 + threadIdx.y;
  const bool live = i < d.Nx && j < d.Ny;
  const bool inxy = i >= 1 && i <= d.Nx - 2 && j >= 1 && j <= d.Ny - 2;
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  const long cxy = live ? (long)img(j, d.Ny) * d.Nx + img(i, d.Nx) : 0;
  const T cpn = (T)cf.cpost_neg;
  double s[3] = {0.0, 0.0, 0.0};
  const int k0 = bk.z * PKZ, k1 = min(k0 + PKZ, d.Nz);
  if (live) {
    for (int k = k0; k < k1; k++) {
      const long n = at(d, i, j, k);
      const long c = (long)img(k, d.Nz) * sz + cxy;
      const T gx = (T)((P[c] - P[c - sx]) * d.idx);
      const T gy = (T)((P[c] - P[c - sy]) * d.idy);
      const T gz = (T)((P[c] - P[c - sz]) * d.idz);
      const T ax = nx[n] + cpn * gx;
      const T ay = ny[n] + cpn * gy;
      const T az = nz[n] + cpn * gz;
      nx[n] = ax;
      ny[n] = ay;
      nz[n] = az;
      if (inxy && k >= 1 && k <= d.Nz - 2) {
        s[0] += (double)ax;
        s[1] += (double)ay;
        s[2] += (double)az;
      }
    }
  }
  const int tid = threadIdx.y * 64 + threadIdx.x;
#pragma unroll
  for (int m = 0; m < 3; m++) {
    double w = s[m];
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
    if ((tid & 63) == 0) ws[m][tid >> 6] = w;
  }
  __syncthreads();
  if (tid < 3) {
    const long nb = (long)gridDim.x * gridDim.y * gridDim.z;
    const long b = ((long)bk.z * gridDim.y + bk.y) * gridDim.x + bk.x;
    part[tid * nb + b] = (ws[tid][0] + ws[tid][1]) + (ws[tid][2] + ws[tid][3]);
  }
}

// RU = RU_np1 with its mean removed, plus c on the x component
// (grid.cpp:565-577 and the copy of grid.cpp:318), from src (the lean last
// stage's RU_np1, which RU_int holds after the stage exchange) and the sums
// sum_parts left in sums[2m], sums[2m+1] = -(sum/denom); the expressions of
// k_add_neg_mean_set.
template <typename T>
__global__ void k_set_mean_free3(long n, T* __restrict__ d0, T* __restrict__ d1,
                                 T* __restrict__ d2, const T* __restrict__ s0,
                                 const T* __restrict__ s1, const T* __restrict__ s2,
                                 const double* __restrict__ sums, double c) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int m = blockIdx.y;
  const T* src = m == 0 ? s0 : (m == 1 ? s1 : s2);
  T* dst = m == 0 ? d0 : (m == 1 ? d1 : d2);
  T v = src[t] + (T)sums[2 * m + 1];
  if (m == 0) v = v + (T)c;
  dst[t] = v;
}

long box_wp_parts(const Dev& d) {
  return 3L * nblocks(d.Nx, 64) * nblocks(d.Ny, 4) * nblocks(d.Nz, PKZ);
}

template <typename T>
int box_set_mean_free(const Dev& d, T* const ru[3], const T* const src[3], const double* sums,
                      double c, hipStream_t s) {
  const long n = (long)d.Nx * d.Ny * d.Nz;
  hipLaunchKernelGGL(k_set_mean_free3<T>, dim3(nblocks(n, 256), 3), dim3(256), 0, s, n, ru[0],
                     ru[1], ru[2], src[0], src[1], src[2], sums, c);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

template <typename T>
int box_update_rho(const Dev& d, const BoxCoef& cf, const T* const ru_int[3], const T* rho,
                   T* rho_np1, T* rho_new, const T* const rf[3], T* const u[3], hipStream_t s,
                   T* rho_set, const T* rho_int) {
  if (rho_set && !cf.last) return FR_EINVAL;
  if ((cf.lean && !rho_set) || (cf.np1_base && (rho_set || cf.last))) return FR_EINVAL;
  if (!rho_int && (!rf[0] || !rf[1] || !rf[2])) return FR_EINVAL;
  const T* f[3] = {rho_int ? rho_int : rf[0], rho_int ? nullptr : rf[1],
                   rho_int ? nullptr : rf[2]};
  auto go = [&](auto kern, const T* r) {
    hipLaunchKernelGGL(kern, grid3(d), dim3(64, 4), 0, s, d, cf, ru_int[0], ru_int[1],
                       ru_int[2], r, rho_np1, rho_new, f[0], f[1], f[2], u[0], u[1], u[2],
                       rho_set);
  };
  auto pick = [&](auto fi) {
    constexpr bool FI = decltype(fi)::value;
    if (rho_set && cf.lean)
      go(k_update_rho<T, RHO_SET_LEAN, FI>, nullptr);
    else if (rho_set)
      go(k_update_rho<T, RHO_SET, FI>, nullptr);
    else if (cf.np1_base)
      go(k_update_rho<T, RHO_BASE, FI>, rho);
    else
      go(k_update_rho<T, RHO_PLAIN, FI>, rho);
  };
  if (rho_int)
    pick(std::true_type{});
  else
    pick(std::false_type{});
  FR_LAUNCH_CHECK();
  return FR_OK;
}

template <typename T>
int box_faces_of(const Dev& d, const T* rn, T* const rf[3], hipStream_t s) {
  hipLaunchKernelGGL(k_faces_of<T>, grid3(d), dim3(64, 4), 0, s, d, rn, rf[0], rf[1], rf[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

template <typename T>
int box_update_ru_wop(const Dev& d, const BoxCoef& cf, const WopCoef& w, const T* const u[3],
                      const T* const ru_int[3], const T* const S[3], const T* const ru[3],
                      T* const ru_np1[3], T* const ru_new[3], const T* const rf[3],
                      hipStream_t s) {
  // z-marching kernel over the interior, then the periodic ghosts of the outputs
  const dim3 g(nblocks(d.Nx - 2, WX), nblocks(d.Ny - 2, WY), nblocks(d.Nz - 2, WKZ));
  if ((cf.lean && !cf.last) || (cf.np1_base && cf.last)) return FR_EINVAL;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, dim3(WX, WY), 0, s, d, cf, w, u[0], u[1], u[2], ru_int[0],
                       ru_int[1], ru_int[2], S[0], S[1], S[2], ru[0], ru[1], ru[2], ru_np1[0],
                       ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2], rf[0], rf[1],
                       rf[2]);
  };
  const int mode = cf.lean ? WOP_LEAN : (cf.np1_base ? WOP_BASE : WOP_PLAIN);
  if (w.gravity) {
    if (mode == WOP_LEAN)
      go(k_update_ru_wop_z<1, WOP_LEAN, T>);
    else if (mode == WOP_BASE)
      go(k_update_ru_wop_z<1, WOP_BASE, T>);
    else
      go(k_update_ru_wop_z<1, WOP_PLAIN, T>);
  } else {
    if (mode == WOP_LEAN)
      go(k_update_ru_wop_z<0, WOP_LEAN, T>);
    else if (mode == WOP_BASE)
      go(k_update_ru_wop_z<0, WOP_BASE, T>);
    else
      go(k_update_ru_wop_z<0, WOP_PLAIN, T>);
  }
  FR_LAUNCH_CHECK();
  if (cf.lean) return ghost_fill_n(d, ru_new, 3, s);
  T* const out6[6] = {ru_np1[0], ru_new[0], ru_np1[1], ru_new[1], ru_np1[2], ru_new[2]};
  return ghost_fill_n(d, out6, 6, s);
}

template <typename T>
int box_div_rhs(const Dev& d, const DivCoef& q, const T* rho_new, const T* Q,
                const T* const ru_new[3], T* const rf[3], double* rhs, hipStream_t s,
                double* part, double* dsum) {
  const dim3 g = grid3(d);
  // rf[0] null: the face values are formed but not stored
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(g.x, g.y, nblocks(d.Nz, DKZ)), dim3(64, 4), 0, s, d, q,
                       rho_new, Q, ru_new[0], ru_new[1], ru_new[2], rf[0], rf[1], rf[2], rhs,
                       part);
  };
  if (rf[0])
    go(k_div_rhs<T, true>);
  else
    go(k_div_rhs<T, false>);
  FR_LAUNCH_CHECK();
  // dsum[0]: the interior sum; dsum[1]: minus the interior mean
  if (part && dsum)
    return sum_parts(part, (long)g.x * g.y * g.z, dsum, s,
                     (double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2));
  return FR_OK;
}

long box_div_parts(const Dev& d) {
  const dim3 g = grid3(d);
  return (long)g.x * g.y * g.z;
}

template <typename T>
int box_update_ru_wp(const Dev& d, const BoxCoef& cf, const double* P, T* const ru_np1[3],
                     T* const ru_new[3], hipStream_t s, double* part, double* sums) {
  if (cf.lean) {
    if (!part || !sums) return FR_EINVAL;
    const dim3 g(nblocks(d.Nx, 64), nblocks(d.Ny, 4), nblocks(d.Nz, PKZ));
    hipLaunchKernelGGL(k_update_ru_wp_lean<T>, g, dim3(64, 4), 0, s, d, cf, P, ru_new[0],
                       ru_new[1], ru_new[2], part);
    FR_LAUNCH_CHECK();
    // sums[2m], sums[2m+1]: component m's interior sum and minus its mean
    return sum_parts(part, (long)g.x * g.y * g.z, sums, s,
                     (double)(d.Nx - 2) * (d.Ny - 2) * (d.Nz - 2), 3);
  }
  hipLaunchKernelGGL(k_update_ru_wp<T>, grid3(d), dim3(64, 4), 0, s, d, cf, P, ru_np1[0],
                     ru_np1[1], ru_np1[2], ru_new[0], ru_new[1], ru_new[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

#define FR_BOX_INST(T)                                                                        \
  template int box_update_rho<T>(const Dev&, const BoxCoef&, const T* const[3], const T*, T*, \
                                 T*, const T* const[3], T* const[3], hipStream_t, T*,         \
                                 const T*);                                                   \
  template int box_faces_of<T>(const Dev&, const T*, T* const[3], hipStream_t);               \
  template int box_update_ru_wop<T>(const Dev&, const BoxCoef&, const WopCoef&,               \
                                    const T* const[3], const T* const[3], const T* const[3],  \
                                    const T* const[3], T* const[3], T* const[3],              \
                                    const T* const[3], hipStream_t);                          \
  template int box_div_rhs<T>(const Dev&, const DivCoef&, const T*, const T*,                 \
                              const T* const[3], T* const[3], double*, hipStream_t, double*,  \
                              double*);                                                       \
  template int box_update_ru_wp<T>(const Dev&, const BoxCoef&, const double*, T* const[3],    \
                                   T* const[3], hipStream_t, double*, double*);               \
  template int box_set_mean_free<T>(const Dev&, T* const[3], const T* const[3],               \
                                    const double*, double, hipStream_t);
FR_BOX_INST(double)
FR_BOX_INST(float)

}  // namespace fr
