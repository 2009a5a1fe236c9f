// Lagrangian particle kernels on gfx950 (particle.cpp:58-658, 745-865).
//
//  * advance : 1-D linear gather of the face velocity at x_int (truncating
//              cell index, particle.cpp:340-360), then the position, drag and
//              heat ODE updates of the RK4 stage (particle.cpp:314-399,480-496)
//  * temp    : trilinear gather of T = (P0/R)/rho on cell centres (rounded
//              index, particle.cpp:402-456) and the heat RHS (460-478)
//  * cells   : particles bucketed by rounded cell index, each bucket in
//              ascending particle order (count + scan + scatter + sort)
//  * deposit : per target cell, the contributions of the particles in the
//              neighbouring buckets summed in ascending particle order --
//              the order of the reference's serial scatter loop, so the
//              deposits are deterministic and equal the reference's bit for
//              bit; no floating-point atomics (particle.cpp:498-658)
//  * migrate : the one-rank swap-with-last removal + append order of
//              SortValid{Z,Y,X}/FromBuf (particle.cpp:150-311, 853-858),
//              computed in parallel from closed-form event ranks
//
// Storage order is not the reference order.  Every storage slot carries the
// particle's reference index (`rank`, with its inverse `inv`); the reference
// order only decides the order of deposit sums and the outcome of migration,
// which therefore rewrites ranks instead of moving particle data.  Every few
// time steps (4 by default, fr_sim_set_resort_interval) the slots are
// re-sorted by cell so gathers and bucket builds touch memory in nearly cell
// order.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include <hip/hip_ext.h>

#include "fr_internal.hpp"

namespace fr {

__device__ __forceinline__ long clampi(long q, long n) { return q < 0 ? 0 : (q >= n ? n - 1 : q); }

// Response time of the particle with flag f under the per-particle extension:
// the top 24 bits of the murmur3 32-bit finaliser of f as a fraction u in
// [0, 1), Tp = lo + span * u.  Flags travel with the particles, so each keeps
// its Tp through migration and storage re-sorts.
__device__ __forceinline__ double tp_of_flag(int f, double lo, double span) {
  unsigned h = (unsigned)f;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return lo + span * ((double)(h >> 8) * (1.0 / 16777216.0));
}

// particle.cpp:314-399, 480-496 (one RK4 stage) with the gather 340-360.
// PH selects the phases of one launch: PH_ALL is the fused stage the time
// step uses; a single phase is one member of the reference's surface
// (gas2part_velocity stores the gathered velocity in gu/gv/gw for
// update_velocity; update_position and update_Temp need no gather).
constexpr int PH_GATHER = 1, PH_POS = 2, PH_VEL = 4, PH_TEMP = 8, PH_ALL = 15;

// EQ (first stage of a step whose stage slots 1 and 3 equal slot 0, as a
// full step leaves them): slot 0 is read once in place of all three.
template <typename T, int PH, bool EQ = false>
__global__ void k_part_advance(Dev d, PartPtrs p, AdvCoef c, const T* __restrict__ ux,
                               const T* __restrict__ uy, const T* __restrict__ uz,
                               double* __restrict__ gu, double* __restrict__ gv,
                               double* __restrict__ gw) {
  const int i = xcd_lin() * blockDim.x + threadIdx.x;
  if (i >= p.np) return;
  double e0[7];
  if constexpr (EQ) {
#pragma unroll
    for (int q = 0; q < 7; q++) e0[q] = p.s[q][0][i];
  }
  auto S0 = [&](int q) -> double {
    if constexpr (EQ) return e0[q]; else return p.s[q][0][i];
  };
  auto S1 = [&](int q) -> double {
    if constexpr (EQ) return e0[q]; else return p.s[q][1][i];
  };
  auto S3 = [&](int q) -> double {
    if constexpr (EQ) return e0[q]; else return p.s[q][3][i];
  };
  // T = float (precision form): the gas-side arithmetic (interpolation
  // weights, interpolated velocity, drag right-hand side) runs in FP32; the
  // particle state and its RK4 accumulation stay FP64.  T = double is the
  // reference arithmetic.
  using Cm = T;
  Cm ug = 0, vg = 0, wg = 0;
  if (PH & PH_GATHER) {
    const double xi = S1(0), yi = S1(1), zi = S1(2);
    const int ti = (int)(xi * d.idx), tj = (int)(yi * d.idy), tk = (int)(zi * d.idz);
    const Cm x_l = (Cm)(xi * d.idx - ((int)(xi * d.idx)));
    const Cm y_l = (Cm)(yi * d.idy - ((int)(yi * d.idy)));
    const Cm z_l = (Cm)(zi * d.idz - ((int)(zi * d.idz)));
    const int ii = ti + 1, jj = tj + 1, kk = tk + 1;
    const long q = (long)kk * d.NxNy + (long)jj * d.Nx + ii;
    ug = x_l * (Cm)ux[clampi(q + 1, d.size)] + (1 - x_l) * (Cm)ux[clampi(q, d.size)];
    vg = y_l * (Cm)uy[clampi(q + d.Nx, d.size)] + (1 - y_l) * (Cm)uy[clampi(q, d.size)];
    wg = z_l * (Cm)uz[clampi(q + d.NxNy, d.size)] + (1 - z_l) * (Cm)uz[clampi(q, d.size)];
    if (!(PH & PH_VEL)) {
      gu[i] = (double)ug;
      gv[i] = (double)vg;
      gw[i] = (double)wg;
    }
  }
  const double ui = S1(3), vi = S1(4), wi = S1(5);
  if (PH & PH_POS) {
    const double X[3] = {ui, vi, wi};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double np1 = S3(a) + c.dt_post * X[a];
      p.s[a][3][i] = np1;
      p.s[a][2][i] = c.last ? np1 : S0(a) + c.dt_pre * X[a];
      if (c.set0) p.s[a][0][i] = np1;
    }
  }
  if (PH & PH_VEL) {
    if (!(PH & PH_GATHER)) {
      ug = (Cm)gu[i];
      vg = (Cm)gv[i];
      wg = (Cm)gw[i];
    }
    // per-particle extension: 1/Tp of the particle's own flag
    const double itp = c.per_particle ? 1. / tp_of_flag(p.flag[i], c.tp_lo, c.tp_span) : c.iTp;
    const double ru = (double)((ug - (Cm)ui) * (Cm)itp);
    const double rv = (double)((vg - (Cm)vi) * (Cm)itp);
    const double rw = (double)((wg - (Cm)wi) * (Cm)itp);
    p.rhs_u[i] = ru;
    p.rhs_v[i] = rv;
    p.rhs_w[i] = rw;
    const double R[3] = {ru, rv, rw};
    const double G[3] = {c.gx, c.gy, c.gz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const double np1 = S3(3 + a) + c.dt_post * (R[a] + G[a]);
      p.s[3 + a][3][i] = np1;
      p.s[3 + a][2][i] = c.last ? np1 : S0(3 + a) + c.dt_pre * (R[a] + G[a]);
      if (c.set0) p.s[3 + a][0][i] = np1;
    }
  }
  if (PH & PH_TEMP) {  // RHS_T from the previous stage
    const double rt = p.rhs_T[i];
    const double tn = S3(6) + c.dtpost_iS * (rt + c.alpha);
    p.s[6][3][i] = tn;
    p.s[6][2][i] = c.last ? tn : S0(6) + c.dtpre_iS * (rt + c.alpha);
    if (c.set0) p.s[6][0][i] = tn;
  }
}

// Trilinear gas temperature at position (X, Y, Z) from the rounded index of
// the upper cell centre (particle.cpp:402-441), Tc(a, b, e) the cell values.
template <typename Cm, typename F>
__device__ __forceinline__ Cm interp_T(const Dev& d, double X, double Y, double Z, F Tc) {
  int ii = (int)(round(X * d.idx));
  int jj = (int)(round(Y * d.idy));
  int kk = (int)(round(Z * d.idz));
  const Cm x_l = (Cm)(X * d.idx - ii + 0.5);
  const Cm y_l = (Cm)(Y * d.idy - jj + 0.5);
  const Cm z_l = (Cm)(Z * d.idz - kk + 0.5);
  ii += 1;
  jj += 1;
  kk += 1;
  return x_l * y_l * z_l * Tc(ii, jj, kk) + (1 - x_l) * y_l * z_l * Tc(ii - 1, jj, kk) +
         x_l * (1 - y_l) * z_l * Tc(ii, jj - 1, kk) + x_l * y_l * (1 - z_l) * Tc(ii, jj, kk - 1) +
         (1 - x_l) * (1 - y_l) * z_l * Tc(ii - 1, jj - 1, kk) +
         (1 - x_l) * y_l * (1 - z_l) * Tc(ii - 1, jj, kk - 1) +
         x_l * (y_l - 1) * (z_l - 1) * Tc(ii, jj - 1, kk - 1) +
         (1 - x_l) * (1 - y_l) * (1 - z_l) * Tc(ii - 1, jj - 1, kk - 1);
}

// particle.cpp:402-478: Tg at the stage `st` position (1 = _int, 2 = _new),
// T = c0/rho (grid.cpp:676, c0 = P0/R), RHS_T = (Tg - T_stage)*iTth.
template <typename F>
__global__ void k_part_temp(Dev d, PartPtrs p, int st, double c0, const F* __restrict__ rho,
                            double iTth) {
  const int i = xcd_lin() * blockDim.x + threadIdx.x;
  if (i >= p.np) return;
  // F = float (precision form): gas temperature interpolation and the heat
  // right-hand side in FP32; F = double is the reference arithmetic.
  using Cm = F;
  const Cm cc0 = (Cm)c0;
  auto Tc = [&](int a, int b, int e) {
    return cc0 / (Cm)rho[clampi((long)e * d.NxNy + (long)b * d.Nx + a, d.size)];
  };
  const Cm Tg = interp_T<Cm>(d, p.s[0][st][i], p.s[1][st][i], p.s[2][st][i], Tc);
  p.rhs_T[i] = (double)((Tg - (Cm)p.s[6][st][i]) * (Cm)iTth);
}

// The same in the reference's two members: gas2part_Temp_{int,new} (mode 0:
// Tg from a gas temperature field into tg) and Compute_RHS_Temp_{int,new}
// (mode 1: RHS_T = (Tg - T_stage)*iTth).
__global__ void k_part_temp_split(Dev d, PartPtrs p, int st, int mode,
                                  const double* __restrict__ T, double* __restrict__ tg,
                                  double iTth) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.np) return;
  if (mode == 0) {
    auto Tc = [&](int a, int b, int e) {
      return T[clampi((long)e * d.NxNy + (long)b * d.Nx + a, d.size)];
    };
    tg[i] = interp_T<double>(d, p.s[0][st][i], p.s[1][st][i], p.s[2][st][i], Tc);
  } else {
    p.rhs_T[i] = (tg[i] - p.s[6][st][i]) * iTth;
  }
}

// ---- device-wide exclusive scan of int32 ------------------------------------
// A tile is STILE = 4096 ints: SV chunks of SB*4, each read as one int4 per
// thread (lane-contiguous 16-byte loads), so every load instruction of a
// wave covers 1 KB of consecutive ints.  Sums within a wave use shuffles;
// integer sums are exact in any order.
constexpr int SB = 256, SV = 4, STILE = SB * 4 * SV;

// EQ0: scan the indicator (in[e] == 0) instead of the values.
template <bool EQ0>
__device__ __forceinline__ int4 load_i4(const int* __restrict__ in, long e, long n) {
  int4 v = make_int4(0, 0, 0, 0);
  if (e + 4 <= n) {
    v = *reinterpret_cast<const int4*>(in + e);
  } else {
    if (e < n) v.x = in[e];
    if (e + 1 < n) v.y = in[e + 1];
    if (e + 2 < n) v.z = in[e + 2];
    if (EQ0) {  // past the end counts nothing
      if (e >= n) v.x = 1;
      if (e + 1 >= n) v.y = 1;
      if (e + 2 >= n) v.z = 1;
      v.w = 1;
    }
  }
  if (EQ0) {
    v.x = v.x == 0;
    v.y = v.y == 0;
    v.z = v.z == 0;
    v.w = v.w == 0;
  }
  return v;
}

__device__ __forceinline__ void store_i4(int* __restrict__ out, long e, long n, int4 v) {
  if (e + 4 <= n) {
    *reinterpret_cast<int4*>(out + e) = v;
    return;
  }
  if (e < n) out[e] = v.x;
  if (e + 1 < n) out[e + 1] = v.y;
  if (e + 2 < n) out[e + 2] = v.z;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__device__ __forceinline__ int wave_incl_scan(int x, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

template <bool EQ0>
__global__ void __launch_bounds__(SB) k_scan_tiles(const int* __restrict__ in, long n,
                                                   int* __restrict__ tsum) {
  __shared__ int wt[SB / 64];
  const long base = (long)blockIdx.x * STILE + threadIdx.x * 4;
  int s = 0;
#pragma unroll
  for (int q = 0; q < SV; q++) {
    const int4 v = load_i4<EQ0>(in, base + (long)q * SB * 4, n);
    s += v.x + v.y + v.z + v.w;
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) wt[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tsum[blockIdx.x] = wt[0] + wt[1] + wt[2] + wt[3];
}

// Exclusive scan of in[0..n) into out[0..n) (in == out allowed) in one
// workgroup of NT threads, in rounds of NT consecutive elements: each round
// is one coalesced load per thread (the next round's issued before this
// round's scan, so reads stay ahead of the writes), wave scans, wave totals
// in LDS, and a running carry.  Returns the total (to every thread).  Kept
// to small workgroups: a single workgroup of many waves waits for a whole
// CU to drain when other streams' kernels fill the GPU.
template <int NT>
__device__ __forceinline__ int block_scan_rounds(const int* in, int n, int* out, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int carry = 0;
  int next = tid < n ? in[tid] : 0;
  for (int base = 0; base < n; base += NT) {
    const int v = next;
    const int i = base + tid, j = i + NT;
    next = j < n ? in[j] : 0;
    const int inc = wave_incl_scan(v, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int off = carry, tot = 0;
#pragma unroll
    for (int u = 0; u < NT / 64; u++) {
      const int t = wsum[u];
      off += u < w ? t : 0;
      tot += t;
    }
    if (i < n) out[i] = off + inc - v;
    carry += tot;
    __syncthreads();
  }
  return carry;
}

// The tile sums are scanned in two levels: blocks of MIDN tile sums in place
// (each thread four consecutive ones), whose totals go to top[]; then top[]
// in one small workgroup, which also stores the grand total.  k_scan_apply
// adds both offsets.
constexpr int MIDN = 1024, ST = 256;
__global__ void __launch_bounds__(ST) k_scan_mid(int* __restrict__ tsum, int nt,
                                                 int* __restrict__ top) {
  __shared__ int wt[ST / 64];
  const int base = blockIdx.x * MIDN + threadIdx.x * 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int v[4];
#pragma unroll
  for (int q = 0; q < 4; q++) v[q] = base + q < nt ? tsum[base + q] : 0;
  const int s4 = v[0] + v[1] + v[2] + v[3];
  const int inc = wave_incl_scan(s4, lane);
  if (lane == 63) wt[w] = inc;
  __syncthreads();
  int run = inc - s4;
  for (int u = 0; u < w; u++) run += wt[u];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (base + q < nt) tsum[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 0) top[blockIdx.x] = wt[0] + wt[1] + wt[2] + wt[3];
}

__global__ void __launch_bounds__(ST) k_scan_top(int* top, int nb, int* __restrict__ out_total) {
  __shared__ int wsum[ST / 64];
  const int total = block_scan_rounds<ST>(top, nb, top, wsum);
  if (threadIdx.x == 0) *out_total = total;
}

// Chunk q of the tile precedes chunk q+1; within a chunk thread t's four
// ints precede thread t+1's.  All SV wave scans are formed before the one
// barrier that publishes the wave totals.
template <bool EQ0>
__global__ void __launch_bounds__(SB) k_scan_apply(const int* __restrict__ in, long n,
                                                   const int* __restrict__ tsum,
                                                   const int* __restrict__ top,
                                                   int* __restrict__ out) {
  __shared__ int wt[SV][SB / 64];
  const long base = (long)blockIdx.x * STILE + threadIdx.x * 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int4 v[SV];
  int inc[SV];
#pragma unroll
  for (int q = 0; q < SV; q++) {
    v[q] = load_i4<EQ0>(in, base + (long)q * SB * 4, n);
    inc[q] = wave_incl_scan(v[q].x + v[q].y + v[q].z + v[q].w, lane);
    if (lane == 63) wt[q][w] = inc[q];
  }
  __syncthreads();
  int run = tsum[blockIdx.x] + top[blockIdx.x / MIDN];
#pragma unroll
  for (int q = 0; q < SV; q++) {
    int x = run;
    for (int u = 0; u < w; u++) x += wt[q][u];
    x += inc[q] - (v[q].x + v[q].y + v[q].z + v[q].w);
    int4 o;
    o.x = x;
    x += v[q].x;
    o.y = x;
    x += v[q].y;
    o.z = x;
    x += v[q].z;
    o.w = x;
    store_i4(out, base + (long)q * SB * 4, n, o);
    run += wt[q][0] + wt[q][1] + wt[q][2] + wt[q][3];
  }
}

// Scratch ints scan_ex_t needs for n elements: the tile sums and the totals
// of their blocks.
inline long scan_tmp_ints(long n) {
  const long nt = (n + STILE - 1) / STILE;
  return nt + (nt + MIDN - 1) / MIDN + 4;
}

// out[0..n) = exclusive scan of in (EQ0: of in[e] == 0), out[n] = total.
// tmp holds scan_tmp_ints(n) ints; in and out are 16-byte aligned (whole
// allocations).
template <bool EQ0>
int scan_ex_t(const int* in, long n, int* out, int* tmp, hipStream_t s) {
  if (((uintptr_t)in | (uintptr_t)out) & 15) return FR_EINVAL;
  const int nt = (int)((n + STILE - 1) / STILE);
  if (nt == 0) {
    hipMemsetAsync(out, 0, sizeof(int), s);
    return FR_OK;
  }
  const int nb = (nt + MIDN - 1) / MIDN;
  int* top = tmp + nt;
  hipLaunchKernelGGL(k_scan_tiles<EQ0>, dim3(nt), dim3(SB), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_mid, dim3(nb), dim3(ST), 0, s, tmp, nt, top);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(ST), 0, s, top, nb, out + n);
  hipLaunchKernelGGL(k_scan_apply<EQ0>, dim3(nt), dim3(SB), 0, s, in, n, tmp, top, out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int scan_ex(const int* in, long n, int* out, int* tmp, hipStream_t s) {
  return scan_ex_t<false>(in, n, out, tmp, s);
}

// Exclusive scan of in[0..n), n = np - *pK read on the device, in one
// workgroup; out[n] = total.  For the per-axis leaver lists, whose length
// the host does not wait for.
__global__ void __launch_bounds__(ST) k_scan_small(const int* __restrict__ in, int np,
                                                   const int* __restrict__ pK,
                                                   int* __restrict__ out) {
  __shared__ int wsum[ST / 64];
  const int n = np - *pK;
  const int total = block_scan_rounds<ST>(in, n, out, wsum);
  if (threadIdx.x == 0) out[n > 0 ? n : 0] = total;
}

// ---- cell buckets ---------------------------------------------------------------
// mode 0: rounded index of the upper cell centre (x2i_round, particle.cpp:
// 121-134), used by the trilinear deposits; mode 1: truncating index (x2i,
// particle.cpp:106-119), used by the 1-D momentum deposit.
__device__ __forceinline__ long bucket_cell(const Dev& d, int mode, double x, double y,
                                            double z) {
  int ii, jj, kk;
  if (mode == 0) {
    ii = (int)(round(x * d.idx)) + 1;
    jj = (int)(round(y * d.idy)) + 1;
    kk = (int)(round(z * d.idz)) + 1;
  } else {
    ii = (int)(x * d.idx) + 1;
    jj = (int)(y * d.idy) + 1;
    kk = (int)(z * d.idz) + 1;
  }
  return clampi((long)kk * d.NxNy + (long)jj * d.Nx + ii, d.size);
}

// Storage is in near cell order (re-sorted every few steps), so neighbouring lanes
// mostly share a cell: each run of equal cells in consecutive lanes takes its
// slots with one atomic (its first lane's), the others offset from it.  Slots
// only have to be distinct within a cell; the placement orders by reference
// index.  256-thread blocks of 64-lane waves; the valid lanes of a wave are a
// prefix (np ends the last wave).
__global__ void __launch_bounds__(256) k_bucket_count(Dev d, PartPtrs p, int st, int mode,
                                                      int* __restrict__ cnt,
                                                      int* __restrict__ key,
                                                      int* __restrict__ slot) {
  const int i = xcd_lin() * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = i < p.np;
  int c = -1;
  if (valid) c = (int)bucket_cell(d, mode, p.s[0][st][i], p.s[1][st][i], p.s[2][st][i]);
  const int prev = __shfl_up(c, 1);
  const bool head = valid && (lane == 0 || prev != c);
  const unsigned long long heads = __ballot(head);
  const int nvalid = __popcll(__ballot(valid));
  const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
  const int leader = valid ? 63 - __clzll(heads & upto) : 0;
  int base = 0;
  if (head) {
    const unsigned long long after = heads & ~upto;
    const int end = min(after ? __ffsll((long long)after) - 1 : 64, nvalid);
    base = atomicAdd(&cnt[c], end - lane);
  }
  base = __shfl(base, leader);
  if (!valid) return;
  key[i] = c;
  slot[i] = base + (lane - leader);
}

// Buckets hold reference indices (ranks).
__global__ void k_bucket_scatter(int np, const int* __restrict__ key, const int* __restrict__ slot,
                                 const int* __restrict__ start, const int* __restrict__ rank,
                                 int* __restrict__ list) {
  const int i = xcd_lin() * blockDim.x + threadIdx.x;
  if (i >= np) return;
  list[start[key[i]] + slot[i]] = rank[i];
}

// Place every particle in its bucket by reference index (slots from atomics
// come in any order): its position is the bucket start plus the number of
// smaller reference indices in the bucket's unsorted list; recorded per
// storage slot in `pos` (the atomic-slot array reused).  One thread per
// particle, so the cost follows the particles, not the cells.
__global__ void k_bucket_place(int np, const int* __restrict__ key, const int* __restrict__ start,
                               const int* __restrict__ rank, const int* __restrict__ unsorted,
                               int* __restrict__ list, int* __restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int c = key[i], r = rank[i];
  const int b = start[c], e = start[c + 1];
  int n = 0;
  for (int u = b; u < e; u++) n += unsorted[u] < r;
  list[b + n] = r;
  pos[i] = b + n;
}

// Per-particle deposit operands written to the particle's bucket position, so
// the deposit kernels read them contiguously (reads here are in particle
// order; the scattered side is the write).
// mode 0 (particle.cpp:536-542): x_l = x*idx - round(x*idx) + 0.5, w = RHS_T.
// mode 1 (particle.cpp:505-511): x_l = x*idx - trunc(x*idx), w = RHS_u,v,w.
__global__ void k_bucket_gather(int np, int mode, PartPtrs p, int st,
                                const int* __restrict__ posv, double idx, double idy,
                                double idz, double* __restrict__ s0, double* __restrict__ s1,
                                double* __restrict__ s2, double* __restrict__ s3,
                                double* __restrict__ s4, double* __restrict__ s5) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int pos = posv[i];
  const double X = p.s[0][st][i], Y = p.s[1][st][i], Z = p.s[2][st][i];
  if (mode == 0) {
    const int ii = (int)(round(X * idx));
    const int jj = (int)(round(Y * idy));
    const int kk = (int)(round(Z * idz));
    s0[pos] = X * idx - ii + 0.5;
    s1[pos] = Y * idy - jj + 0.5;
    s2[pos] = Z * idz - kk + 0.5;
    s3[pos] = p.rhs_T[i];
  } else {
    s0[pos] = (X * idx - ((int)(X * idx)));
    s1[pos] = (Y * idy - ((int)(Y * idy)));
    s2[pos] = (Z * idz - ((int)(Z * idz)));
    s3[pos] = p.rhs_u[i];
    s4[pos] = p.rhs_v[i];
    s5[pos] = p.rhs_w[i];
  }
}

// Split-form builds: k_bucket_place and the operand gather in one pass,
// each particle's operand record (mode 0: x_l, y_l, z_l, RHS_T; mode 1:
// x_l, RHS_u, y_l, RHS_v, z_l, RHS_w; the arithmetic of k_bucket_gather)
// written at its placed position.  Without OPS only the placement.
template <bool OPS>
__global__ void k_bucket_place_rec(int np, int mode, PartPtrs p, int st,
                                   const int* __restrict__ key, const int* __restrict__ start,
                                   const int* __restrict__ rank, const int* __restrict__ unsorted,
                                   int* __restrict__ list, double idx, double idy, double idz,
                                   double* __restrict__ rec) {
  const int i = xcd_lin() * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int c = key[i], r = rank[i];
  const int b = start[c], e = start[c + 1];
  int n = 0;
  for (int u = b; u < e; u++) n += unsorted[u] < r;
  const long pos = b + n;
  list[pos] = r;
  if (!OPS) return;
  const double X = p.s[0][st][i], Y = p.s[1][st][i], Z = p.s[2][st][i];
  if (mode == 0) {
    const int ii = (int)(round(X * idx));
    const int jj = (int)(round(Y * idy));
    const int kk = (int)(round(Z * idz));
    double2* q = reinterpret_cast<double2*>(rec + 4 * pos);
    q[0] = make_double2(X * idx - ii + 0.5, Y * idy - jj + 0.5);
    q[1] = make_double2(Z * idz - kk + 0.5, p.rhs_T[i]);
  } else {
    double2* q = reinterpret_cast<double2*>(rec + 6 * pos);
    q[0] = make_double2(X * idx - ((int)(X * idx)), p.rhs_u[i]);
    q[1] = make_double2(Y * idy - ((int)(Y * idy)), p.rhs_v[i]);
    q[2] = make_double2(Z * idz - ((int)(Z * idz)), p.rhs_w[i]);
  }
}

// ---- deposits -----------------------------------------------------------------
// Target tiles of DTX x DTY x 1 cells.  The bucket starts of the source cells
// a tile reads (DTX+1 cells and their end, DTY+1 rows, 2 planes) are staged in
// LDS once per workgroup.
constexpr int DTX = 64, DTY = 4, DSR = DTX + 2, DROWS = (DTY + 1) * 2;

// Stage start[] of rows (b0 + r%(DTY+1), c0 + r/(DTY+1)), columns a0 + j,
// j < DSR.  Columns are clamped to [0, Nx] (column Nx is the row's end, the
// start of the next row) and rows outside the grid are all zero, so every
// staged cell outside the grid is an empty bucket: ss[r][j+1] - ss[r][j] is
// the particle count of source cell j with no bounds tests.  Cell indices fit
// in 32 bits (at most 1026^3 ghosted cells).
__device__ __forceinline__ void stage_starts(const Dev& d, const int* __restrict__ start,
                                             int a0, int b0, int c0, int (*ss)[DSR]) {
  const int tid = threadIdx.y * DTX + threadIdx.x;
  for (int t = tid; t < DROWS * DSR; t += DTX * DTY) {
    const int r = t / DSR, j = t - r * DSR;
    const int B = b0 + r % (DTY + 1), Cz = c0 + r / (DTY + 1);
    const int A = min(max(a0 + j, 0), d.Nx);
    int v = 0;
    if (B >= 0 && B < d.Ny && Cz >= 0 && Cz < d.Nz) v = start[(Cz * d.Ny + B) * d.Nx + A];
    ss[r][j] = v;
  }
  __syncthreads();
}

// K-way merge over bucket ranges in ascending particle order; the head
// particle index of every range is kept in a register.
template <int NB>
struct Merge {
  int cur[NB], end[NB], head[NB];
  __device__ __forceinline__ void init(const int* __restrict__ list) {
#pragma unroll
    for (int q = 0; q < NB; q++) head[q] = cur[q] < end[q] ? list[cur[q]] : INT_MAX;
  }
  // returns the range of the next particle (-1 when all are done), its position
  __device__ __forceinline__ int next(const int* __restrict__ list, int& pos) {
    int best = INT_MAX, wb = -1;
#pragma unroll
    for (int q = 0; q < NB; q++)
      if (head[q] < best) {
        best = head[q];
        wb = q;
      }
    // the chosen range's cursor by selects, so that each step issues one
    // load (a load per range under divergent lanes would serialise them)
    int c = 0, e = 0;
#pragma unroll
    for (int q = 0; q < NB; q++)
      if (q == wb) {
        c = cur[q];
        e = end[q];
      }
    const int h = c + 1 < e ? list[c + 1] : INT_MAX;
#pragma unroll
    for (int q = 0; q < NB; q++)
      if (q == wb) {
        cur[q] = c + 1;
        head[q] = h;
      }
    pos = c;
    return wb;
  }
};

// Capacity of a tile's (particle, target) contribution list in LDS; a tile
// with more contributions uses the per-target merge instead.
constexpr int PCAP = 1024;

// Exclusive scan of one int per thread over a 256-thread (4-wave) block;
// returns the thread's offset, *total gets the sum.  wtot: 4 ints of LDS.
__device__ __forceinline__ int block_excl_scan(int v, int* wtot, int* total) {
  const int lane = threadIdx.x & 63, w = (threadIdx.y * blockDim.x + threadIdx.x) >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  int base = 0;
  for (int k = 0; k < w; k++) base += wtot[k];
  *total = wtot[0] + wtot[1] + wtot[2] + wtot[3];
  return base + inc - v;
}

// Per-target form of the trilinear deposit: an 8-way merge of the target's
// source buckets (a..a+1, b..b+1, c..c+1) in ascending particle order.
__device__ __forceinline__ double dep_cic_merge(const int (*ss)[DSR], const int* __restrict__ list,
                                                const double* __restrict__ sxl,
                                                const double* __restrict__ syl,
                                                const double* __restrict__ szl,
                                                const double* __restrict__ sw) {
  Merge<8> m;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int r = threadIdx.y + ((q >> 1) & 1) + (DTY + 1) * (q >> 2), j = threadIdx.x + (q & 1);
    m.cur[q] = ss[r][j];
    m.end[q] = ss[r][j + 1];
  }
  m.init(list);
  double acc = 0.0;
  for (;;) {
    int pos;
    const int q = m.next(list, pos);
    if (q < 0) break;
    const double x_l = sxl[pos], y_l = syl[pos], z_l = szl[pos];
    const double wx = (q & 1) ? (1 - x_l) : x_l;
    const double wy = ((q >> 1) & 1) ? (1 - y_l) : y_l;
    const double wz = (q >> 2) ? (1 - z_l) : z_l;
    double t = wx * wy * wz;
    if (sw) t = t * sw[pos];
    acc = acc + t;
  }
  return acc;
}

// Trilinear deposit of particle.cpp:526-559 (sw == nullptr, concentration)
// or 586-658 (weight RHS_T): raw per-cell sums for every ghosted cell before
// the cumulative halo, in ascending particle order (the reference's scatter
// order), from the rounded-index buckets.  Target (a,b,c) collects from the
// bucket of cell (a+ox, b+oy, c+oz) the weight with (1-x_l) where ox = 1.
//
// A 64x4 target tile reads the particles of 10 source rows; each row's
// buckets are one contiguous range of the bucket arrays.  Particle-parallel:
// every source particle writes its (reference index, weight) for each of its
// in-tile targets into that target's segment of an LDS list (segment sizes
// from the bucket starts, offsets by a block scan); each contribution's
// place in its segment is the number of smaller reference indices there;
// each target then adds its segment in that order.  The same sum as the
// reference's loop, with no per-target merge.  A tile whose list exceeds
// PCAP is left to k_dep_cic_dense (its id is appended to dl, dl[0] counts)
// or, with dl null, merges per target.
template <typename T>
__global__ void __launch_bounds__(256) k_dep_cic(Dev d, const int* __restrict__ start,
                                                 const int* __restrict__ list,
                                                 const double* __restrict__ sxl,
                                                 const double* __restrict__ syl,
                                                 const double* __restrict__ szl,
                                                 const double* __restrict__ sw,
                                                 T* __restrict__ out, int* __restrict__ dl) {
  __shared__ int ss[DROWS][DSR];
  __shared__ int seg[DTX * DTY + 1];
  __shared__ int cur[DTX * DTY];
  __shared__ int wtot[4];
  __shared__ int rcnt[DROWS];
  __shared__ int rk[PCAP];
  __shared__ unsigned char tg[PCAP];
  __shared__ double wv[PCAP];
  __shared__ unsigned short ord[PCAP];
  const uint3 bk = xcd_block();
  const int a0 = bk.x * DTX, b0 = bk.y * DTY, c = bk.z;
  stage_starts(d, start, a0, b0, c, ss);
  const int tid = threadIdx.y * DTX + threadIdx.x;
  const int a = a0 + threadIdx.x, b = b0 + threadIdx.y;
  const bool live = a < d.Nx && b < d.Ny;
  // contributions to this target: buckets (a..a+1, b..b+1, c..c+1); the two
  // x neighbours of a staged row telescope (cells outside the grid are empty,
  // so a target outside it counts zero)
  int nt = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int r = threadIdx.y + (q & 1) + (DTY + 1) * (q >> 1);
    nt += ss[r][threadIdx.x + 2] - ss[r][threadIdx.x];
  }
  if (tid < DROWS) rcnt[tid] = ss[tid][DSR - 1] - ss[tid][0];  // source cells a0 .. a0+DTX
  int total;
  const int off = block_excl_scan(nt, wtot, &total);  // its barrier also publishes rcnt
  if (total <= PCAP) {
    seg[tid] = off;
    cur[tid] = off;
    if (tid == 0) seg[DTX * DTY] = total;
    int rpre[DROWS + 1];
    rpre[0] = 0;
#pragma unroll
    for (int r = 0; r < DROWS; r++) rpre[r + 1] = rpre[r] + rcnt[r];
    __syncthreads();
    const int npart = rpre[DROWS];
    for (int f = tid; f < npart; f += DTX * DTY) {
      int r = 0, base = 0;  // source row of contribution f (unrolled: rpre stays in registers)
#pragma unroll
      for (int k = 1; k < DROWS; k++) r += f >= rpre[k];
#pragma unroll
      for (int k = 0; k < DROWS; k++)
        if (k == r) base = rpre[k];
      const int u = ss[r][0] + (f - base);
      // source cell: the last of columns 0..DTX whose bucket starts at or
      // before u (empty buckets before it share its start, later ones start
      // after u); a fixed 7-step search over the 65 columns
      int lo = 0;
#pragma unroll
      for (int w = 64; w > 0; w >>= 1)
        if (lo + w <= DTX && ss[r][lo + w] <= u) lo += w;
      const int sa = a0 + lo, sb = b0 + r % (DTY + 1), oz = r / (DTY + 1);
      const double x_l = sxl[u], y_l = syl[u], z_l = szl[u];
      const int rank = list[u];
#pragma unroll
      for (int oy = 0; oy < 2; oy++) {
        const int tb = sb - oy - b0;
        if (tb < 0 || tb >= DTY) continue;
#pragma unroll
        for (int ox = 0; ox < 2; ox++) {
          const int ta = sa - ox - a0;
          if (ta < 0 || ta >= DTX) continue;
          const double wx = ox ? (1 - x_l) : x_l;
          const double wy = oy ? (1 - y_l) : y_l;
          const double wz = oz ? (1 - z_l) : z_l;
          double t = wx * wy * wz;
          if (sw) t = t * sw[u];
          const int tl = tb * DTX + ta;
          const int k = atomicAdd(&cur[tl], 1);
          if (k < seg[tl + 1]) {
            rk[k] = rank;
            tg[k] = (unsigned char)tl;
            wv[k] = t;
          }
        }
      }
    }
    __syncthreads();
    for (int j = tid; j < total; j += DTX * DTY) {
      const int tl = tg[j], r = rk[j];
      const int s0 = seg[tl], s1 = seg[tl + 1];
      int n = 0;
      for (int k = s0; k < s1; k++) n += rk[k] < r;
      ord[s0 + n] = (unsigned short)j;
    }
    __syncthreads();
    if (live) {
      double acc = 0.0;
      for (int k = off; k < off + nt; k++) acc = acc + wv[ord[k]];
      out[at(d, a, b, c)] = (T)acc;
    }
    return;
  }
  // a dense tile: its id goes to the list of k_dep_cic_dense, or without a
  // list every target merges its buckets from global memory
  if (dl) {
    if (tid == 0) dl[1 + atomicAdd(dl, 1)] = (int)(bk.x + gridDim.x * (bk.y + gridDim.y * bk.z));
    return;
  }
  if (live) out[at(d, a, b, c)] = (T)dep_cic_merge(ss, list, sxl, syl, szl, sw);
}

// Dense tiles of k_dep_cic (ids in dl[1..dl[0]], grid-stride): the particles
// of the tile's ten source rows are staged in LDS in bucket order (reference
// index and operands), and every target merges its eight source buckets
// there in ascending reference index.  A tile with more than DCAP source
// particles merges from global memory.
constexpr int DCAP = 2048;

template <typename T>
__global__ void __launch_bounds__(256) k_dep_cic_dense(Dev d, const int* __restrict__ start,
                                                       const int* __restrict__ list,
                                                       const double* __restrict__ sxl,
                                                       const double* __restrict__ syl,
                                                       const double* __restrict__ szl,
                                                       const double* __restrict__ sw,
                                                       T* __restrict__ out,
                                                       const int* __restrict__ dl, int nbx,
                                                       int nby) {
  __shared__ int ss[DROWS][DSR];
  __shared__ int prk[DCAP];
  __shared__ double px[DCAP], py[DCAP], pz[DCAP], pw[DCAP];
  const int ntile = dl[0];
  const int tid = threadIdx.y * DTX + threadIdx.x;
  for (int it = blockIdx.x; it < ntile; it += gridDim.x) {
    const int id = dl[1 + it];
    const int a0 = (id % nbx) * DTX, b0 = ((id / nbx) % nby) * DTY, c = id / (nbx * nby);
    __syncthreads();  // the previous tile is done with ss and the staged particles
    stage_starts(d, start, a0, b0, c, ss);
    const int a = a0 + threadIdx.x, b = b0 + threadIdx.y;
    const bool live = a < d.Nx && b < d.Ny;
    // LDS position of bucket entry u of row r: u + boff[r]
    int rpre[DROWS + 1], boff[DROWS];
    rpre[0] = 0;
#pragma unroll
    for (int r = 0; r < DROWS; r++) {
      rpre[r + 1] = rpre[r] + (ss[r][DSR - 1] - ss[r][0]);
      boff[r] = rpre[r] - ss[r][0];
    }
    const int npart = rpre[DROWS];
    if (npart > DCAP) {  // block-uniform
      if (live) out[at(d, a, b, c)] = (T)dep_cic_merge(ss, list, sxl, syl, szl, sw);
      continue;
    }
    // SU particles per thread per round, all loads issued before the stores
    constexpr int SU = 8;
    for (int f0 = tid; f0 < npart; f0 += SU * DTX * DTY) {
      int uu[SU];
#pragma unroll
      for (int e = 0; e < SU; e++) {
        const int f = f0 + e * DTX * DTY;
        int r = 0, base = 0;
#pragma unroll
        for (int k = 1; k < DROWS; k++) r += f >= rpre[k];
#pragma unroll
        for (int k = 0; k < DROWS; k++)
          if (k == r) base = rpre[k];
        uu[e] = f < npart ? ss[r][0] + (f - base) : -1;
      }
      int vr[SU];
      double vx[SU], vy[SU], vz[SU], vw[SU];
#pragma unroll
      for (int e = 0; e < SU; e++) {
        if (uu[e] < 0) continue;
        vr[e] = list[uu[e]];
        vx[e] = sxl[uu[e]];
        vy[e] = syl[uu[e]];
        vz[e] = szl[uu[e]];
        if (sw) vw[e] = sw[uu[e]];
      }
#pragma unroll
      for (int e = 0; e < SU; e++) {
        if (uu[e] < 0) continue;
        const int f = f0 + e * DTX * DTY;
        prk[f] = vr[e];
        px[f] = vx[e];
        py[f] = vy[e];
        pz[f] = vz[e];
        if (sw) pw[f] = vw[e];
      }
    }
    __syncthreads();
    if (!live) continue;
    // 8-way merge with packed state: key[q] = reference index * 8 + q of the
    // bucket's head (INT_MAX when done; indices stay below 2^28), ce[q] =
    // cursor | end << 16 (staged positions are below DCAP)
    int key[8], ce[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int r = threadIdx.y + ((q >> 1) & 1) + (DTY + 1) * (q >> 2), j = threadIdx.x + (q & 1);
      int o = 0;
#pragma unroll
      for (int k = 0; k < DROWS; k++)
        if (k == r) o = boff[k];
      const int c0 = ss[r][j] + o, e0 = ss[r][j + 1] + o;
      ce[q] = c0 | (e0 << 16);
      key[q] = c0 < e0 ? ((prk[c0] << 3) | q) : INT_MAX;
    }
    double acc = 0.0;
    for (;;) {
      int best = key[0];
#pragma unroll
      for (int q = 1; q < 8; q++) best = min(best, key[q]);
      if (best == INT_MAX) break;
      const int q = best & 7;
      int cq = 0;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k == q) cq = ce[k];
      const int pos = cq & 0xffff, nx = pos + 1;
      const int h = nx < (cq >> 16) ? ((prk[nx] << 3) | q) : INT_MAX;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k == q) {
          ce[k] = cq + 1;
          key[k] = h;
        }
      const double x_l = px[pos], y_l = py[pos], z_l = pz[pos];
      const double wx = (q & 1) ? (1 - x_l) : x_l;
      const double wy = ((q >> 1) & 1) ? (1 - y_l) : y_l;
      const double wz = (q >> 2) ? (1 - z_l) : z_l;
      double t = wx * wy * wz;
      if (sw) t = t * pw[pos];
      acc = acc + t;
    }
    out[at(d, a, b, c)] = (T)acc;
  }
}

// 1-D linear momentum deposit (particle.cpp:498-524) from the truncating-
// index buckets: component M of target t collects (1-x_l)*R from bucket t and
// x_l*R from bucket t - e_M, in ascending particle order.  LDS rows hold
// y = b0-1.., z = c-1..c, x = a0-1.. (the lower neighbours).
template <int M>
__device__ __forceinline__ double dep_vel_comp(const int (*ss)[DSR], const int* __restrict__ list,
                                               const double* __restrict__ l,
                                               const double* __restrict__ R, int cm) {
  const int r0 = threadIdx.y + 1 + (DTY + 1), j0 = threadIdx.x + 1;
  const int r1 = M == 1 ? r0 - 1 : (M == 2 ? r0 - (DTY + 1) : r0);
  const int j1 = M == 0 ? j0 - 1 : j0;
  Merge<2> m;
  m.cur[0] = ss[r0][j0];
  m.end[0] = ss[r0][j0 + 1];
  if (cm > 0) {
    m.cur[1] = ss[r1][j1];
    m.end[1] = ss[r1][j1 + 1];
  } else {
    m.cur[1] = m.end[1] = 0;
  }
  m.init(list);
  double acc = 0.0;
  for (;;) {
    int pos;
    const int q = m.next(list, pos);
    if (q < 0) break;
    acc = acc + (q == 0 ? (1 - l[pos]) * R[pos] : l[pos] * R[pos]);
  }
  return acc;
}

template <typename T>
__global__ void __launch_bounds__(256) k_dep_vel(Dev d, const int* __restrict__ start,
                                                 const int* __restrict__ list,
                                                 const double* __restrict__ sxl,
                                                 const double* __restrict__ syl,
                                                 const double* __restrict__ szl,
                                                 const double* __restrict__ sru,
                                                 const double* __restrict__ srv,
                                                 const double* __restrict__ srw,
                                                 T* __restrict__ ox, T* __restrict__ oy,
                                                 T* __restrict__ oz) {
  // Particle-parallel form as in k_dep_cic: 3 segments per target (one per
  // component); a source particle in cell s adds (1-l_M) R_M to target s and
  // l_M R_M to target s + e_M.
  constexpr int VCAP = 896, NT = DTX * DTY;
  __shared__ int ss[DROWS][DSR];
  __shared__ int seg[3 * NT + 1];
  __shared__ int cur[3 * NT];
  __shared__ int wtot[4];
  __shared__ int rcnt[DROWS];
  __shared__ int rk[VCAP];
  __shared__ unsigned short tg[VCAP];
  __shared__ double wv[VCAP];
  __shared__ unsigned short ord[VCAP];
  const uint3 bk = xcd_block();
  const int a0 = bk.x * DTX, b0 = bk.y * DTY, c = bk.z;
  stage_starts(d, start, a0 - 1, b0 - 1, c - 1, ss);
  const int tid = threadIdx.y * DTX + threadIdx.x;
  const int a = a0 + threadIdx.x, b = b0 + threadIdx.y;
  const bool live = a < d.Nx && b < d.Ny;
  // per-component contribution counts of this target (buckets t and t - e_M)
  // (staged cells outside the grid are empty buckets, so the neighbours
  // below a = 0, b = 0, c = 0 count zero)
  int nm[3] = {0, 0, 0};
  if (live) {
    const int r0 = threadIdx.y + 1 + (DTY + 1), j0 = threadIdx.x + 1;
    const int own = ss[r0][j0 + 1] - ss[r0][j0];
    nm[0] = own + (ss[r0][j0] - ss[r0][j0 - 1]);
    nm[1] = own + (ss[r0 - 1][j0 + 1] - ss[r0 - 1][j0]);
    nm[2] = own + (ss[r0 - (DTY + 1)][j0 + 1] - ss[r0 - (DTY + 1)][j0]);
  }
  // source cells a0-1 .. a0+63 of each row (j = 0 .. DTX); row 0 (plane c-1,
  // row b0-1) reaches no target of the tile and is skipped
  if (tid < DROWS) rcnt[tid] = tid == 0 ? 0 : ss[tid][DSR - 1] - ss[tid][0];
  int total;
  const int off = block_excl_scan(nm[0] + nm[1] + nm[2], wtot, &total);
  if (total <= VCAP) {
    seg[3 * tid] = off;
    seg[3 * tid + 1] = off + nm[0];
    seg[3 * tid + 2] = off + nm[0] + nm[1];
    cur[3 * tid] = off;
    cur[3 * tid + 1] = off + nm[0];
    cur[3 * tid + 2] = off + nm[0] + nm[1];
    if (tid == 0) seg[3 * NT] = total;
    int rpre[DROWS + 1];
    rpre[0] = 0;
#pragma unroll
    for (int r = 0; r < DROWS; r++) rpre[r + 1] = rpre[r] + rcnt[r];
    __syncthreads();
    const int npart = rpre[DROWS];
    const double* L[3] = {sxl, syl, szl};
    const double* R[3] = {sru, srv, srw};
    for (int f = tid; f < npart; f += NT) {
      int r = 0, base = 0;
#pragma unroll
      for (int k = 1; k < DROWS; k++) r += f >= rpre[k];
#pragma unroll
      for (int k = 0; k < DROWS; k++)
        if (k == r) base = rpre[k];
      const int u = ss[r][0] + (f - base);
      int lo = 0;  // source cell: last of columns 0..DTX starting at or before u
#pragma unroll
      for (int w = 64; w > 0; w >>= 1)
        if (lo + w <= DTX && ss[r][lo + w] <= u) lo += w;
      const int sa = lo - 1, sb = r % (DTY + 1) - 1, sc = r / (DTY + 1) - 1;  // tile-relative
      const int rank = list[u];
      // components with a target in the tile: x and y need sc = 0, x and z
      // need sb >= 0
      const bool need[3] = {sc == 0 && sb >= 0, sc == 0, sb >= 0};
#pragma unroll
      for (int M = 0; M < 3; M++) {
        if (!need[M]) continue;
        const double l = L[M][u], Rv = R[M][u];
#pragma unroll
        for (int up = 0; up < 2; up++) {  // up = 0: target s, weight (1-l) R; 1: s + e_M, l R
          const int ta = sa + (M == 0 ? up : 0), tb = sb + (M == 1 ? up : 0),
                    tc = sc + (M == 2 ? up : 0);
          if (tc != 0 || ta < 0 || ta >= DTX || tb < 0 || tb >= DTY) continue;
          if (a0 + ta >= d.Nx || b0 + tb >= d.Ny) continue;
          const int sgi = 3 * (tb * DTX + ta) + M;
          const int k = atomicAdd(&cur[sgi], 1);
          if (k < seg[sgi + 1]) {
            rk[k] = rank;
            tg[k] = (unsigned short)sgi;
            wv[k] = up == 0 ? (1 - l) * Rv : l * Rv;
          }
        }
      }
    }
    __syncthreads();
    for (int j = tid; j < total; j += NT) {
      const int sgi = tg[j], rr = rk[j];
      const int s0 = seg[sgi], s1 = seg[sgi + 1];
      int n = 0;
      for (int k = s0; k < s1; k++) n += rk[k] < rr;
      ord[s0 + n] = (unsigned short)j;
    }
    __syncthreads();
    if (live) {
      const long n = at(d, a, b, c);
      T* o[3] = {ox, oy, oz};
      int k = off;
#pragma unroll
      for (int M = 0; M < 3; M++) {
        double acc = 0.0;
        for (int e = 0; e < nm[M]; e++, k++) acc = acc + wv[ord[k]];
        o[M][n] = (T)acc;
      }
    }
    return;
  }
  if (!live) return;
  const long n = at(d, a, b, c);
  ox[n] = (T)dep_vel_comp<0>(ss, list, sxl, sru, a);
  oy[n] = (T)dep_vel_comp<1>(ss, list, syl, srv, b);
  oz[n] = (T)dep_vel_comp<2>(ss, list, szl, srw, c);
}

// ---- pull-form deposits ----------------------------------------------------------
// One thread per target cell, which merges its source buckets in ascending
// reference index straight from the bucket arrays (bucket order is cell
// order, so the lanes of a wave read neighbouring ranges).  No LDS lists, no
// atomics, no barriers: at the box's densities (well under one particle per
// cell) a target has a handful of contributions and the kernel streams its
// output like a stencil.  The sums are the reference's, bit for bit.

// Two-way merges of the target's own bucket (weight (1-l) R) with the
// bucket below it along each component (weight l R), the three components
// advanced together: each round takes one step of every unfinished merge,
// so their loads are in flight at once and a wave waits for about as many
// memory round trips as its longest merge, not their sum.  A merge whose
// lower bucket is empty (most of them at box densities) reads no reference
// indices.
struct Merge2 {
  int i0, e0, i1, e1, k0, k1;
  double acc;
};

__device__ __forceinline__ void merge2_start(Merge2& m, const int* __restrict__ list) {
  const bool both = m.i0 < m.e0 && m.i1 < m.e1;
  // with one bucket empty the order is the other bucket's own order: keys
  // 0 (own) or 1 (lower) stand in for the reference indices
  m.k0 = m.i0 < m.e0 ? (both ? list[m.i0] : 0) : INT_MAX;
  m.k1 = m.i1 < m.e1 ? (both ? list[m.i1] : 1) : INT_MAX;
  m.acc = 0.0;
}

__device__ __forceinline__ void merge2_step(Merge2& m, const int* __restrict__ list,
                                            const double* __restrict__ L,
                                            const double* __restrict__ R) {
  const bool own = m.k0 < m.k1;  // reference indices are distinct
  const int pos = own ? m.i0 : m.i1;
  const double l = L[pos], r = R[pos];
  m.acc = m.acc + (own ? (1 - l) * r : l * r);
  const int nx = pos + 1, en = own ? m.e0 : m.e1;
  // the key of the next entry: its reference index while the other bucket
  // still has entries, else a placeholder that keeps this bucket chosen
  const int other = own ? m.k1 : m.k0;
  const int kn = nx < en ? (other != INT_MAX ? list[nx] : (own ? 0 : 1)) : INT_MAX;
  if (own) {
    m.i0 = nx;
    m.k0 = kn;
  } else {
    m.i1 = nx;
    m.k1 = kn;
  }
}

// particle.cpp:498-524 (before the cumulative halo): target t of component M
// collects (1-l_M) R_M from the particles of bucket t and l_M R_M from those
// of bucket t - e_M (truncating-index buckets).  Every ghosted cell is
// written.
// The three components of momentum target n = (a, b, c): two-way merges
// (own bucket, bucket below along the component), advanced together.
template <typename T>
__device__ __forceinline__ void dep_vel_target(const Dev& d, const int* __restrict__ start,
                                               const int* __restrict__ list,
                                               const double* const L[3],
                                               const double* const R[3], T* const o[3], int a,
                                               int b, int c) {
  const int n = (c * d.Ny + b) * d.Nx + a;
  const int s0 = start[n], e0 = start[n + 1];
  Merge2 m[3];
#pragma unroll
  for (int M = 0; M < 3; M++) {
    m[M].i0 = s0;
    m[M].e0 = e0;
    m[M].i1 = m[M].e1 = 0;
  }
  if (a > 0) {
    m[0].i1 = start[n - 1];
    m[0].e1 = s0;
  }
  if (b > 0) {
    m[1].i1 = start[n - d.Nx];
    m[1].e1 = start[n - d.Nx + 1];
  }
  if (c > 0) {
    m[2].i1 = start[n - d.NxNy];
    m[2].e1 = start[n - d.NxNy + 1];
  }
#pragma unroll
  for (int M = 0; M < 3; M++) merge2_start(m[M], list);
  for (;;) {
    bool any = false;
#pragma unroll
    for (int M = 0; M < 3; M++) {
      if (m[M].k0 != INT_MAX || m[M].k1 != INT_MAX) {
        merge2_step(m[M], list, L[M], R[M]);
        any = true;
      }
    }
    if (!any) break;
  }
#pragma unroll
  for (int M = 0; M < 3; M++) o[M][n] = (T)m[M].acc;
}

// particle.cpp:498-524 (before the cumulative halo): target t of component M
// collects (1-l_M) R_M from the particles of bucket t and l_M R_M from those
// of bucket t - e_M (truncating-index buckets).  Every ghosted cell is
// written.
template <typename T>
__global__ void __launch_bounds__(256) k_dep_vel_pull(Dev d, const int* __restrict__ start,
                                                      const int* __restrict__ list,
                                                      const double* __restrict__ sxl,
                                                      const double* __restrict__ syl,
                                                      const double* __restrict__ szl,
                                                      const double* __restrict__ sru,
                                                      const double* __restrict__ srv,
                                                      const double* __restrict__ srw,
                                                      T* __restrict__ ox, T* __restrict__ oy,
                                                      T* __restrict__ oz) {
  const uint3 bk = xcd_block();
  const int a = bk.x * DTX + threadIdx.x, b = bk.y * DTY + threadIdx.y, c = bk.z;
  if (a >= d.Nx || b >= d.Ny) return;
  const double* L[3] = {sxl, syl, szl};
  const double* R[3] = {sru, srv, srw};
  T* o[3] = {ox, oy, oz};
  dep_vel_target(d, start, list, L, R, o, a, b, c);
}

// The eight bucket ranges of trilinear target (a, b, c): range 2*rq + ox is
// source cell (a+ox, b+(rq&1), c+(rq>>1)).  Buckets outside the grid are
// empty: columns clamp to the row end (column Nx is the start of the next
// row), rows and planes past the grid get no range.
__device__ __forceinline__ void cic_ranges(const Dev& d, const int* __restrict__ start, int a,
                                           int b, int c, int* cur, int* end) {
  const int a1 = min(a + 1, d.Nx), a2 = min(a + 2, d.Nx);
#pragma unroll
  for (int rq = 0; rq < 4; rq++) {
    const int B = b + (rq & 1), Cz = c + (rq >> 1);
    int v0 = 0, v1 = 0, v2 = 0;
    if (B < d.Ny && Cz < d.Nz) {
      const int rb = (Cz * d.Ny + B) * d.Nx;
      v0 = start[rb + a];
      v1 = start[rb + a1];
      v2 = start[rb + a2];
    }
    cur[2 * rq] = v0;
    end[2 * rq] = v1;
    cur[2 * rq + 1] = v1;
    end[2 * rq + 1] = v2;
  }
}

// Weight of the particle at bucket position pos for the target it reaches
// through range q (particle.cpp:536-556 / 597-655).
__device__ __forceinline__ double cic_weight(int q, int pos, const double* __restrict__ sxl,
                                             const double* __restrict__ syl,
                                             const double* __restrict__ szl,
                                             const double* __restrict__ sw) {
  const double x_l = sxl[pos], y_l = syl[pos], z_l = szl[pos];
  const double wx = (q & 1) ? (1 - x_l) : x_l;
  const double wy = ((q >> 1) & 1) ? (1 - y_l) : y_l;
  const double wz = (q >> 2) ? (1 - z_l) : z_l;
  double t = wx * wy * wz;
  if (sw) t = t * sw[pos];
  return t;
}

// The ordered sum of trilinear target (a, b, c): an 8-way merge of its
// source buckets in ascending reference index.
__device__ __forceinline__ double dep_cic_target(const Dev& d, const int* __restrict__ start,
                                                 const int* __restrict__ list,
                                                 const double* __restrict__ sxl,
                                                 const double* __restrict__ syl,
                                                 const double* __restrict__ szl,
                                                 const double* __restrict__ sw, int a, int b,
                                                 int c) {
  Merge<8> m;
  cic_ranges(d, start, a, b, c, m.cur, m.end);
  m.init(list);
  double acc = 0.0;
  for (;;) {
    int pos;
    const int q = m.next(list, pos);
    if (q < 0) break;
    acc = acc + cic_weight(q, pos, sxl, syl, szl, sw);
  }
  return acc;
}

// particle.cpp:526-559 / 586-658 (before the cumulative halo): target
// (a,b,c) collects from the rounded-index bucket of cell (a+ox, b+oy, c+oz)
// the weight with (1-x_l) where ox = 1 and x_l where ox = 0 (likewise y, z),
// times RHS_T when sw is given; an 8-way merge in ascending reference index.
// Every ghosted cell is written.
template <typename T>
__global__ void __launch_bounds__(256) k_dep_cic_pull(Dev d, const int* __restrict__ start,
                                                      const int* __restrict__ list,
                                                      const double* __restrict__ sxl,
                                                      const double* __restrict__ syl,
                                                      const double* __restrict__ szl,
                                                      const double* __restrict__ sw,
                                                      T* __restrict__ out) {
  const uint3 bk = xcd_block();
  const int a = bk.x * DTX + threadIdx.x, b = bk.y * DTY + threadIdx.y, c = bk.z;
  if (a >= d.Nx || b >= d.Ny) return;
  out[(c * d.Ny + b) * d.Nx + a] = (T)dep_cic_target(d, start, list, sxl, syl, szl, sw, a, b, c);
}

// ---- split-form deposits --------------------------------------------------------
// The reference adds a target's contributions to 0.0 in ascending particle
// order.  With at most two contributions that sum does not depend on the
// order: 0.0 + t is t except that -0 becomes +0, and (0.0 + t1) + t2 equals
// (0.0 + t2) + t1 in every case, signed zeros included.  The split form sums
// such targets straight from their bucket ranges with no reference index
// read; the targets with three or more contributions (at 0.12 particles per
// cell about 7% of the trilinear and 0.2% of the momentum targets: Poisson
// counts of mean 0.96 and 0.24) are compacted per workgroup in LDS and merged
// in ascending reference index by the workgroup's first threads.
//
// These kernels are bound by instruction issue, not bytes, so the layout
// serves the instruction count: a thread owns one (a, b) column of SKZ
// consecutive target planes (the bucket starts of a source plane serve two
// target planes from registers; cells a, a+1, a+2 of a row come from one
// address with immediate offsets), and the bucket gather stores each
// particle's operands as one record (array of structures) so that one
// contribution is one or two 16-byte loads from one address.
constexpr int SKZ = 4;    // target planes per thread
constexpr int SNT = 256;  // threads per workgroup, consecutive cells of an (x, y) plane

// Trilinear operand record: x_l, y_l, z_l, RHS_T (32 B).  Weight through
// range q = ox + 2 oy + 4 oz (particle.cpp:536-556 / 597-655; the same
// expression as cic_weight).
template <bool W>
__device__ __forceinline__ double cic_weight_rec(int q, const double* __restrict__ ops, int pos) {
  const double* r = ops + 4 * (long)pos;
  const double2 xy = *reinterpret_cast<const double2*>(r);
  const double2 zw = *reinterpret_cast<const double2*>(r + 2);
  const double wx = (q & 1) ? (1 - xy.x) : xy.x;
  const double wy = ((q >> 1) & 1) ? (1 - xy.y) : xy.y;
  const double wz = (q >> 2) ? (1 - zw.x) : zw.x;
  double t = wx * wy * wz;
  if (W) t = t * zw.y;
  return t;
}

template <bool W>
__device__ __forceinline__ double dep_cic_target_rec(const Dev& d, const int* __restrict__ start,
                                                     const int* __restrict__ list,
                                                     const double* __restrict__ ops, int a, int b,
                                                     int c) {
  Merge<8> m;
  cic_ranges(d, start, a, b, c, m.cur, m.end);
  m.init(list);
  double acc = 0.0;
  for (;;) {
    int pos;
    const int q = m.next(list, pos);
    if (q < 0) break;
    acc = acc + cic_weight_rec<W>(q, ops, pos);
  }
  return acc;
}

// Bucket starts of source rows b and b+1 of plane cz for column (a, b):
// v[0..2] = start of cells a, a+1, a+2 of row b (a+2 clamped to the row end),
// v[3..5] the same of row b+1; rows or planes outside the grid are empty.
// start[] holds ncell + 2 entries, so the unclamped a+2 load of the last cell
// stays inside it.
__device__ __forceinline__ void cic_plane_starts(const Dev& d, const int* __restrict__ start,
                                                 int ij, int cz, bool col, bool row1,
                                                 bool lastx, int* v) {
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = 0;
  if (col && cz < d.Nz) {
    const int* p = start + ((long)cz * d.NxNy + ij);
    v[0] = p[0];
    v[1] = p[1];
    v[2] = p[2];
    if (row1) {
      v[3] = p[d.Nx];
      v[4] = p[d.Nx + 1];
      v[5] = p[d.Nx + 2];
    }
    if (lastx) {
      v[2] = v[1];
      v[5] = v[4];
    }
  }
}

// Trilinear deposit of particle.cpp:526-559 / 586-658 (before the cumulative
// halo), split form; every ghosted cell is written.  ops: trilinear records
// in bucket order.
// part (W only, may be null): per workgroup, the sum of its outputs (each
// thread over its planes, then a fixed tree) at part[bk.y * gridDim.x + bk.x].
// Every deposited value lands in exactly one interior cell after the
// cumulative ghost fold, so these partials sum to the field's interior sum
// after the fold (in another order).
template <typename T, bool W>
__global__ void __launch_bounds__(SNT) k_dep_cic_split(Dev d, const int* __restrict__ start,
                                                       const int* __restrict__ list,
                                                       const double* __restrict__ ops,
                                                       T* __restrict__ out,
                                                       double* __restrict__ part) {
  __shared__ int wt[4];
  __shared__ double wsum[SNT / 64];
  __shared__ int hl[SNT * SKZ];
  // the tile's outputs, stored to memory together once its merged targets
  // are in (whole cache lines, not lines with the merged targets as holes)
  __shared__ T ob[SKZ][SNT];
  const uint3 bk = xcd_block();
  const int tid = threadIdx.x;
  const int ij = bk.x * SNT + tid, c0 = bk.y * SKZ;
  const bool col = ij < d.NxNy;
  const int b = col ? ij / d.Nx : 0, a = ij - b * d.Nx;
  const bool row1 = col && b + 1 < d.Ny, lastx = a == d.Nx - 1;
  int lo[6], hi[6];
  cic_plane_starts(d, start, ij, c0, col, row1, lastx, lo);
  unsigned hard = 0;
#pragma unroll
  for (int k = 0; k < SKZ; k++) {
    const int c = c0 + k;
    cic_plane_starts(d, start, ij, c + 1, col, row1, lastx, hi);
    if (col && c < d.Nz) {
      // source rows rq = oy + 2 oz: (b, c), (b+1, c), (b, c+1), (b+1, c+1);
      // each is the contiguous range [v0, v2) of cells a (ox = 0, below v1)
      // and a+1 (ox = 1).  The first two contributions in range order:
      int tot = 0, p1 = 0, p2 = 0, r1 = 0, r2 = 0, m1 = 0, m2 = 0;
#pragma unroll
      for (int rq = 3; rq >= 0; rq--) {
        const int* v = rq < 2 ? lo + 3 * rq : hi + 3 * (rq - 2);
        const int n = v[2] - v[0];
        tot += n;
        if (n > 0) {
          p2 = n >= 2 ? v[0] + 1 : p1;
          r2 = n >= 2 ? rq : r1;
          m2 = n >= 2 ? v[1] : m1;
          p1 = v[0];
          r1 = rq;
          m1 = v[1];
        }
      }
      if (tot <= 2) {
        double acc = 0.0;
        if (tot > 0) {
          const bool two = tot == 2;
          const int pb = two ? p2 : p1, rb = two ? r2 : r1, mb = two ? m2 : m1;
          const double t1 = cic_weight_rec<W>(2 * r1 + (p1 >= m1 ? 1 : 0), ops, p1);
          const double t2 = cic_weight_rec<W>(2 * rb + (pb >= mb ? 1 : 0), ops, pb);
          acc = acc + t1;
          if (two) acc = acc + t2;
        }
        ob[k][tid] = (T)acc;
      } else {
        hard |= 1u << k;
      }
    }
#pragma unroll
    for (int q = 0; q < 6; q++) lo[q] = hi[q];
  }
  int nh;
  int o = block_excl_scan(__popc(hard), wt, &nh);
#pragma unroll
  for (int k = 0; k < SKZ; k++)
    if (hard & (1u << k)) hl[o++] = (k << 8) | tid;
  __syncthreads();
  for (int i = tid; i < nh; i += SNT) {
    const int e = hl[i];
    const int ij2 = bk.x * SNT + (e & 255), c = c0 + (e >> 8);
    const int b2 = ij2 / d.Nx, a2 = ij2 - b2 * d.Nx;
    ob[e >> 8][e & 255] = (T)dep_cic_target_rec<W>(d, start, list, ops, a2, b2, c);
  }
  __syncthreads();
  double ts = 0.0;
  if (col) {
#pragma unroll
    for (int k = 0; k < SKZ; k++)
      if (c0 + k < d.Nz) {
        const T v = ob[k][tid];
        out[(long)(c0 + k) * d.NxNy + ij] = v;
        ts += (double)v;
      }
  }
  if (W && part) {
    for (int o = 32; o > 0; o >>= 1) ts += __shfl_xor(ts, o);
    if ((tid & 63) == 0) wsum[tid >> 6] = ts;
    __syncthreads();
    if (tid == 0) part[(long)bk.y * gridDim.x + bk.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
  }
}

// Momentum operand record: (x_l, RHS_u), (y_l, RHS_v), (z_l, RHS_w) (48 B);
// component M's pair at 16 M bytes.
__device__ __forceinline__ double2 vel_pair(const double* __restrict__ ops, int pos, int M) {
  return *reinterpret_cast<const double2*>(ops + 6 * (long)pos + 2 * M);
}

// merge2_step on the operand records
__device__ __forceinline__ void merge2_step_rec(Merge2& m, const int* __restrict__ list,
                                                const double* __restrict__ ops, int M) {
  const bool own = m.k0 < m.k1;
  const int pos = own ? m.i0 : m.i1;
  const double2 lr = vel_pair(ops, pos, M);
  m.acc = m.acc + (own ? (1 - lr.x) * lr.y : lr.x * lr.y);
  const int nx = pos + 1, en = own ? m.e0 : m.e1;
  const int other = own ? m.k1 : m.k0;
  const int kn = nx < en ? (other != INT_MAX ? list[nx] : (own ? 0 : 1)) : INT_MAX;
  if (own) {
    m.i0 = nx;
    m.k0 = kn;
  } else {
    m.i1 = nx;
    m.k1 = kn;
  }
}

template <typename T>
__device__ __forceinline__ void dep_vel_target_rec(const Dev& d, const int* __restrict__ start,
                                                   const int* __restrict__ list,
                                                   const double* __restrict__ ops,
                                                   T* const o[3], int a, int b, int c) {
  const long n = (long)c * d.NxNy + (long)b * d.Nx + a;
  const int s0 = start[n], e0 = start[n + 1];
  Merge2 m[3];
#pragma unroll
  for (int M = 0; M < 3; M++) {
    m[M].i0 = s0;
    m[M].e0 = e0;
    m[M].i1 = m[M].e1 = 0;
  }
  if (a > 0) {
    m[0].i1 = start[n - 1];
    m[0].e1 = s0;
  }
  if (b > 0) {
    m[1].i1 = start[n - d.Nx];
    m[1].e1 = start[n - d.Nx + 1];
  }
  if (c > 0) {
    m[2].i1 = start[n - d.NxNy];
    m[2].e1 = start[n - d.NxNy + 1];
  }
#pragma unroll
  for (int M = 0; M < 3; M++) merge2_start(m[M], list);
  for (;;) {
    bool any = false;
#pragma unroll
    for (int M = 0; M < 3; M++) {
      if (m[M].k0 != INT_MAX || m[M].k1 != INT_MAX) {
        merge2_step_rec(m[M], list, ops, M);
        any = true;
      }
    }
    if (!any) break;
  }
#pragma unroll
  for (int M = 0; M < 3; M++) o[M][n] = (T)m[M].acc;
}

// Momentum deposit of particle.cpp:498-524 (before the cumulative halo),
// split form: component M of target t takes (1-l_M) R_M from bucket t and
// l_M R_M from bucket t - e_M; a target whose components each have at most
// two contributions is summed directly, the rest merge all three.  The own
// range of a plane is the z-lower range of the next.  ops: momentum records
// in bucket order.
template <typename T>
__global__ void __launch_bounds__(SNT) k_dep_vel_split(Dev d, const int* __restrict__ start,
                                                       const int* __restrict__ list,
                                                       const double* __restrict__ ops,
                                                       T* __restrict__ ox, T* __restrict__ oy,
                                                       T* __restrict__ oz) {
  __shared__ int wt[4];
  __shared__ int hl[SNT * SKZ];
  const uint3 bk = xcd_block();
  const int tid = threadIdx.x;
  const int ij = bk.x * SNT + tid, c0 = bk.y * SKZ;
  const bool col = ij < d.NxNy;
  const int b = col ? ij / d.Nx : 0, a = ij - b * d.Nx;
  T* const o[3] = {ox, oy, oz};
  int zs = 0, ze = 0;  // own range of plane c - 1
  if (col && c0 > 0) {
    const int* p = start + ((long)(c0 - 1) * d.NxNy + ij);
    zs = p[0];
    ze = p[1];
  }
  unsigned hard = 0;
#pragma unroll
  for (int k = 0; k < SKZ; k++) {
    const int c = c0 + k;
    if (col && c < d.Nz) {
      const long n = (long)c * d.NxNy + ij;
      const int* p = start + n;
      const int s0 = p[0], e0 = p[1];
      int i1[3] = {0, 0, zs}, e1[3] = {0, 0, ze};
      if (a > 0) {
        i1[0] = p[-1];
        e1[0] = s0;
      }
      if (b > 0) {
        i1[1] = p[-d.Nx];
        e1[1] = p[-d.Nx + 1];
      }
      const int no = e0 - s0;
      int nl[3];
      bool hk = false;
#pragma unroll
      for (int M = 0; M < 3; M++) {
        nl[M] = e1[M] - i1[M];
        hk = hk || no + nl[M] > 2;
      }
      if (!hk) {
        const int pb = no > 1 ? s0 + 1 : s0;
#pragma unroll
        for (int M = 0; M < 3; M++) {
          double2 w0 = make_double2(0, 0), w1 = w0, u0 = w0, u1 = w0;
          if (no > 0) {
            w0 = vel_pair(ops, s0, M);
            w1 = vel_pair(ops, pb, M);
          }
          if (nl[M] > 0) {
            u0 = vel_pair(ops, i1[M], M);
            u1 = vel_pair(ops, nl[M] > 1 ? i1[M] + 1 : i1[M], M);
          }
          double acc = 0.0;
          if (no > 0) acc = acc + (1 - w0.x) * w0.y;
          if (no > 1) acc = acc + (1 - w1.x) * w1.y;
          if (nl[M] > 0) acc = acc + u0.x * u0.y;
          if (nl[M] > 1) acc = acc + u1.x * u1.y;
          o[M][n] = (T)acc;
        }
      } else {
        hard |= 1u << k;
      }
      zs = s0;
      ze = e0;
    }
  }
  int nh;
  int off = block_excl_scan(__popc(hard), wt, &nh);
#pragma unroll
  for (int k = 0; k < SKZ; k++)
    if (hard & (1u << k)) hl[off++] = (k << 8) | tid;
  __syncthreads();
  for (int i = tid; i < nh; i += SNT) {
    const int e = hl[i];
    const int ij2 = bk.x * SNT + (e & 255), c = c0 + (e >> 8);
    const int b2 = ij2 / d.Nx, a2 = ij2 - b2 * d.Nx;
    dep_vel_target_rec(d, start, list, ops, o, a2, b2, c);
  }
}

// ---- migration (one rank) -------------------------------------------------------
// Flags live in reference order: flg[rank] = 2 if P > hi, 1 if P < lo, for the
// leavers only (flg is cleared first).
__global__ void k_mig_flag(int np, const double* __restrict__ P, const int* __restrict__ rank,
                           double lo, double hi, int* __restrict__ flg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double v = P[i];
  const int f = v > hi ? 2 : (v < lo ? 1 : 0);
  if (f) flg[rank[i]] = f;  // flg is all zero on entry (k_mig_event clears it)
}

// Front holes (leavers at i < K) get rank m = i - pNL[i]; tail non-leavers
// (i >= K) get rank from the end m = K - pNL[i] - 1.
__global__ void k_mig_rank(int np, const int* __restrict__ flg, const int* __restrict__ pNL,
                           int* __restrict__ holepos, int* __restrict__ tailnl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int K = pNL[np];
  if (i < K && flg[i] != 0) holepos[i - pNL[i]] = i;
  if (i >= K && flg[i] == 0) tailnl[K - pNL[i] - 1] = i;
}

// Event order of the serial scan: E = np - x with x = q_m for the front hole
// of rank m (q_0 = np, q_m = position of the tail non-leaver of rank m-1),
// x = q_h for a leaver sitting at K (h = number of front holes), x = i for
// every other tail leaver.
__global__ void k_mig_event(int np, int* __restrict__ flg, const int* __restrict__ pNL,
                            const int* __restrict__ tailnl, int* __restrict__ evt,
                            int* __restrict__ islo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np || flg[i] == 0) return;
  const int K = pNL[np];
  const int h = K - pNL[K < np ? K : np];
  int x;
  if (i < K) {
    const int m = i - pNL[i];
    x = m == 0 ? np : tailnl[m - 1];
  } else if (i == K) {
    x = h == 0 ? np : tailnl[h - 1];
  } else {
    x = i;
  }
  const int e = np - x;
  evt[e] = i;
  islo[e] = flg[i] == 1;
  flg[i] = 0;  // leaves flg all zero for the next axis
}

struct MigAxis {
  int axis;
  double L;
};

// The moved particles: tail non-leaver s < h goes to the front hole s; leaver
// of event e goes to K + (its rank among the lo leavers, or nlo + its rank
// among the hi leavers) -- the append order lo then hi.  Records the storage
// slot (read from inv before any rank changes), the new reference index and
// whether the particle is a leaver.
__global__ void k_mig_moves(int np, const int* __restrict__ pNL,
                            const int* __restrict__ holepos, const int* __restrict__ tailnl,
                            const int* __restrict__ evt, const int* __restrict__ plo,
                            const int* __restrict__ inv, int* __restrict__ mslot,
                            int* __restrict__ mto, int* __restrict__ mlv) {
  const int K = pNL[np], nleave = np - K;
  const int h = K - pNL[K < np ? K : np];
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < h + nleave;
       s += gridDim.x * blockDim.x) {
    int src, to, lv;
    if (s < h) {
      src = tailnl[s];
      to = holepos[s];
      lv = 0;
    } else {
      const int e = s - h;
      src = evt[e];
      const int nlo = plo[nleave];
      const int r = plo[e];
      const bool lo = (plo[e + 1] - r) != 0;
      to = K + (lo ? r : nlo + (e - r));
      lv = 1;
    }
    mslot[s] = inv[src];
    mto[s] = to;
    mlv[s] = lv;
  }
}

// New reference indices of the moved slots; leavers are wrapped by MODE_L*
// (particle.cpp:58-104): _new, base and _np1 of this axis shift by -+L when
// _new > L or < 0.
__global__ void k_mig_apply(int np, PartPtrs p, MigAxis ax, const int* __restrict__ pNL,
                            const int* __restrict__ mslot, const int* __restrict__ mto,
                            const int* __restrict__ mlv, int* __restrict__ rank,
                            int* __restrict__ inv) {
  const int K = pNL[np], nleave = np - K;
  const int h = K - pNL[K < np ? K : np];
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < h + nleave;
       s += gridDim.x * blockDim.x) {
    const int slot = mslot[s], to = mto[s];
    rank[slot] = to;
    inv[to] = slot;
    if (mlv[s]) {
      const int a = ax.axis;
      double vn = p.s[a][2][slot], v0 = p.s[a][0][slot], v1 = p.s[a][3][slot];
      if (vn > ax.L) {
        vn -= ax.L;
        v0 -= ax.L;
        v1 -= ax.L;
      } else if (vn < 0) {
        vn += ax.L;
        v0 += ax.L;
        v1 += ax.L;
      }
      p.s[a][2][slot] = vn;
      p.s[a][0][slot] = v0;
      p.s[a][3][slot] = v1;
    }
  }
}

// ---- duct inlet/outlet ----------------------------------------------------------------
__global__ void k_live(int np, const int* __restrict__ rank, int K, int* __restrict__ live) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) live[i] = rank[i] < K;
}

// Slots with live[i] move to slot pos[i] of the second storage set.
__global__ void k_compact(int np, const int* __restrict__ live, const int* __restrict__ pos,
                          PartPtrs a, const int* __restrict__ ra, PartPtrs b,
                          int* __restrict__ rb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np || !live[i]) return;
  const int u = pos[i];
#pragma unroll
  for (int q = 0; q < 7; q++)
#pragma unroll
    for (int t = 0; t < 4; t++) b.s[q][t][u] = a.s[q][t][i];
  b.rhs_u[u] = a.rhs_u[i];
  b.rhs_v[u] = a.rhs_v[i];
  b.rhs_w[u] = a.rhs_w[i];
  b.rhs_T[u] = a.rhs_T[i];
  b.flag[u] = a.flag[i];
  rb[u] = ra[i];
}

// Inlet particle t (ToBuf order: 4 stages of x,y,z,u,v,w,T, then the flag as
// a double) becomes slot and reference index np0 + t; its RHS values start
// at zero, as the reference's resized vectors do.
__global__ void k_inject(int n, int np0, const double* __restrict__ buf, PartPtrs p,
                         int* __restrict__ rank, int* __restrict__ inv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int u = np0 + t;
  const double* b = buf + (long)t * 29;
#pragma unroll
  for (int q = 0; q < 7; q++)
#pragma unroll
    for (int st = 0; st < 4; st++) p.s[q][st][u] = b[q * 4 + st];
  p.flag[u] = (int)b[28];
  p.rhs_u[u] = 0.0;
  p.rhs_v[u] = 0.0;
  p.rhs_w[u] = 0.0;
  p.rhs_T[u] = 0.0;
  rank[u] = u;
  inv[u] = u;
}

// The box's '>hi' list of its x pass, which the duct of a coupled run
// receives as its Rbuf1 (particle.cpp:821, 836-838).  After the x pass the
// leavers hold reference indices K..np-1, the '<lo' list first, so the '>hi'
// list is the last n_hi = (np - K) - n_lo indices in packing order, with
// K = pNL[np] and n_lo = plo[np - K] as the x pass left them.  Each row is
// the ToBuf layout (particle.cpp:252-271): 7 quantities x 4 stages, then the
// flag.  Grid-stride over the rows; block 0 lane 0 stores the count.
__global__ void __launch_bounds__(256) k_gather_hi(int np, const int* __restrict__ pNL,
                                                   const int* __restrict__ plo,
                                                   const int* __restrict__ inv, PartPtrs p,
                                                   double* __restrict__ buf, int cap,
                                                   int* __restrict__ n_out) {
  int nhi = 0;
  if (np > 0) {
    const int nleave = np - pNL[np];
    nhi = nleave - plo[nleave];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = nhi;
  const int n = nhi < cap ? nhi : cap;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int u = inv[np - nhi + t];
    double* b = buf + (long)t * 29;
#pragma unroll
    for (int q = 0; q < 7; q++)
#pragma unroll
      for (int st = 0; st < 4; st++) b[q * 4 + st] = p.s[q][st][u];
    b[28] = (double)p.flag[u];
  }
}

// ---- storage order ------------------------------------------------------------------
__global__ void k_iota2(int n, int* __restrict__ a, int* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i] = i;
}

__global__ void k_inv(int n, const int* __restrict__ rank, int* __restrict__ inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) inv[rank[i]] = i;
}

// Reference order <-> storage order of one array.
template <typename T>
__global__ void k_to_ref(int n, const int* __restrict__ rank, const T* __restrict__ src,
                         T* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[rank[i]] = src[i];
}

template <typename T>
__global__ void k_from_ref(int n, const int* __restrict__ rank, const T* __restrict__ src,
                           T* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[rank[i]];
}

// Whether stage slots 1 and 3 equal slot 0 bit for bit for every particle
// (diff stays 0 if so).  At the top of a time step they do on every path of
// the step (the last stage's advance writes its result to slots 0 (in the
// box's step), 2 and 3, migration wraps slots 0, 2 and 3 alike, and the stage
// swap moves slot 2 to slot 1), but a state set from outside need not.  Reads
// are in slot order.
__global__ void __launch_bounds__(256) k_stage_eq(int np, PartPtrs a, int* __restrict__ diff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool same = true;
  if (i < np) {
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const long long v0 = __double_as_longlong(a.s[q][0][i]);
      same = same && __double_as_longlong(a.s[q][1][i]) == v0 &&
             __double_as_longlong(a.s[q][3][i]) == v0;
    }
  }
  if (__ballot(!same) && (threadIdx.x & 63) == 0) atomicOr(diff, 1);
}

// New slot u takes the particle of reference index r = list[u] (cell
// order) from slot inv[r]; inv[r] becomes u (list is a permutation, so each
// thread reads and writes only its own inv entry).  The reads are scattered
// (a particle that changed its cell row or plane since the last re-sort
// comes from far away, and each value read costs a cache line): with *diff
// clear, slot 0 is read once and written to slots 0, 1 and 3.  (Reading in
// slot order and scattering the writes instead measured slower: 16.9 against
// 10.3 ms at 512^3 / 16M.)
// Runs only at the top of the first RK stage's particle update, right before
// the advance.
__global__ void k_resort(int np, const int* __restrict__ list, int* __restrict__ inv,
                         PartPtrs a, PartPtrs b, int* __restrict__ rb,
                         const int* __restrict__ diff) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= np) return;
  const int r = list[u];
  const int o = inv[r];
  // Stage slot 2 and rhs_u,v,w are written by the advance that follows
  // before anything reads them; slots 0, 1, 3 and rhs_T carry over.
  if (*diff == 0) {
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const double v = a.s[q][0][o];
      b.s[q][0][u] = v;
      b.s[q][1][u] = v;
      b.s[q][3][u] = v;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 7; q++) {
      b.s[q][0][u] = a.s[q][0][o];
      b.s[q][1][u] = a.s[q][1][o];
      b.s[q][3][u] = a.s[q][3][o];
    }
  }
  b.rhs_T[u] = a.rhs_T[o];
  b.flag[u] = a.flag[o];
  rb[u] = r;
  inv[r] = u;
}

}  // namespace fr

// ---- host side of the particle set ---------------------------------------------
namespace fr {

static int alloc_buckets(Buckets& B, long ncell, int cap, int nops) {
  FR_CHECK(hipMalloc(&B.cnt, sizeof(int) * (ncell + 1)));
  // two entries past the scanned ncell + 1: the split-form deposits load the
  // start of cell a+2 before clamping it to the row end
  FR_CHECK(hipMalloc(&B.start, sizeof(int) * (ncell + 2)));
  FR_CHECK(hipMemset(B.start, 0, sizeof(int) * (ncell + 2)));
  FR_CHECK(hipMalloc(&B.rec, sizeof(double) * nops * (size_t)cap));
  FR_CHECK(hipMalloc(&B.list, sizeof(int) * cap));
  FR_CHECK(hipMalloc(&B.list2, sizeof(int) * cap));
  FR_CHECK(hipMalloc(&B.key, sizeof(int) * cap));
  FR_CHECK(hipMalloc(&B.slot, sizeof(int) * cap));
  for (int q = 0; q < nops; q++) FR_CHECK(hipMalloc(&B.s[q], sizeof(double) * cap));
  FR_CHECK(hipMalloc(&B.scan_tmp, sizeof(int) * scan_tmp_ints(ncell)));
  return FR_OK;
}

static void free_buckets(Buckets& B) {
  void* all[] = {B.cnt, B.start, B.list, B.list2, B.key, B.slot, B.scan_tmp, B.rec};
  for (void* p : all) (void)hipFree(p);
  for (double* p : B.s)
    if (p) (void)hipFree(p);
}

// Default forms by the mean particles per ghosted cell.  Up to
// SPARSE_MAX_DENSITY both deposits take the split form: at 512^3 / 16M
// particles (0.12 per cell) its kernels run alone in 1.20 ms (trilinear, tiled
// form 1.90 ms) and 2.08 ms (momentum, pull form 2.42 ms, tiled 7.9 ms), and
// the full step is 0.6% faster (interleaved runs on one box: 808.3, 808.7
// against 802.7, 803.6 Mcell*steps/s with the momentum deposit pulled and the
// trilinear ones tiled).  Denser sets (the duct's inlet third holds about 3
// particles per cell, where nearly every target merges) take the tiled form.
constexpr double SPARSE_MAX_DENSITY = 0.5;

int PartDev::dep_kind(bool trilinear) const {
  if (dep_form == 3) return trilinear ? DEP_TILE : DEP_PULL;
  if (dep_form == 1) return DEP_PULL;
  if (dep_form == 2) return DEP_TILE;
  if (dep_form == 4) return DEP_SPLIT;
  (void)trilinear;
  return (double)np <= SPARSE_MAX_DENSITY * (double)ncell ? DEP_SPLIT : DEP_TILE;
}

int PartDev::alloc(int capacity, const Dev& d) {
  flg_zero = false;
  if (const char* e = getenv("FR_DEPOSIT"))
    dep_form = !strcmp(e, "pull")    ? 1
               : !strcmp(e, "tile")  ? 2
               : !strcmp(e, "mixed") ? 3
               : !strcmp(e, "split") ? 4
                                     : 0;
  // FR_HEAT_SUM=field: the heat sum always from a reduction of the folded
  // field (the split deposit's own partials add the same values in another
  // order; the switch lets the deposit forms be compared bit for bit)
  if (const char* e = getenv("FR_HEAT_SUM")) heat_sum_field = !strcmp(e, "field");
  cap = capacity < 1 ? 1 : capacity;
  ncell = d.size;
  for (int q = 0; q < 7; q++)
    for (int t = 0; t < 4; t++) {
      FR_CHECK(hipMalloc(&s[q][t], sizeof(double) * cap));
      FR_CHECK(hipMalloc(&alt.s[q][t], sizeof(double) * cap));
    }
  FR_CHECK(hipMalloc(&flag, sizeof(int) * cap));
  FR_CHECK(hipMalloc(&alt.flag, sizeof(int) * cap));
  double** r[8] = {&rhs_u, &rhs_v, &rhs_w, &rhs_T, &alt.rhs_u, &alt.rhs_v, &alt.rhs_w, &alt.rhs_T};
  for (auto pp : r) {
    FR_CHECK(hipMalloc(pp, sizeof(double) * cap));
    FR_CHECK(hipMemset(*pp, 0, sizeof(double) * cap));
  }
  FR_CHECK(hipMalloc(&rank, sizeof(int) * cap));
  FR_CHECK(hipMalloc(&rank_alt, sizeof(int) * cap));
  FR_CHECK(hipMalloc(&inv, sizeof(int) * cap));
  int rc = alloc_buckets(R, ncell, cap, 4);
  if (!rc) rc = alloc_buckets(T, ncell, cap, 6);
  if (rc) return rc;
  const long big = ncell > cap ? ncell : cap;
  FR_CHECK(hipMalloc(&mig_tmp, sizeof(int) * scan_tmp_ints(big)));
  FR_CHECK(hipMalloc(&rank_dep, sizeof(int) * cap));
  int** m[11] = {&flg, &nl, &pNL, &holepos, &tailnl, &evt, &islo, &plo, &dst, &mslot, &mlv};
  for (auto pp : m) FR_CHECK(hipMalloc(pp, sizeof(int) * (cap + 1)));
  FR_CHECK(hipHostMalloc(&hbuf, sizeof(int) * 4, hipHostMallocDefault));
  FR_CHECK(hipMalloc(&dflag, sizeof(int) * 4));
  const long ntile = (long)nblocks(d.Nx, DTX) * nblocks(d.Ny, DTY) * d.Nz;
  FR_CHECK(hipMalloc(&dense_c, sizeof(int) * (ntile + 1)));
  FR_CHECK(hipMalloc(&dense_w, sizeof(int) * (ntile + 1)));
  cic_nparts = (long)nblocks(d.NxNy, SNT) * nblocks(d.Nz, SKZ);
  FR_CHECK(hipMalloc(&cic_part, sizeof(double) * cic_nparts));
  FR_RET(t_vel.init());
  FR_RET(t_cic.init());
  return FR_OK;
}

void PartDev::release() {
  t_vel.release();
  t_cic.release();
  for (int q = 0; q < 7; q++)
    for (int t = 0; t < 4; t++) {
      (void)hipFree(s[q][t]);
      (void)hipFree(alt.s[q][t]);
    }
  void* all[] = {flag, rhs_u, rhs_v, rhs_w, rhs_T, alt.flag, alt.rhs_u, alt.rhs_v, alt.rhs_w,
                 alt.rhs_T, rank, rank_alt, inv, mig_tmp, rank_dep, flg, nl, pNL,
                 holepos, tailnl, evt, islo, plo, dst, mslot, mlv, inj_buf, dense_c, dense_w,
                 gas[0], gas[1], gas[2], gas[3], dflag, cic_part};
  for (void* p : all) (void)hipFree(p);
  for (auto& g : gas) g = nullptr;
  gas_cap = 0;
  free_buckets(R);
  free_buckets(T);
  (void)hipHostFree(hbuf);
}

PartPtrs PartDev::ptrs() const {
  PartPtrs p;
  for (int q = 0; q < 7; q++)
    for (int t = 0; t < 4; t++) p.s[q][t] = s[q][t];
  p.flag = flag;
  p.rhs_u = rhs_u;
  p.rhs_v = rhs_v;
  p.rhs_w = rhs_w;
  p.rhs_T = rhs_T;
  p.np = np;
  return p;
}

template <typename T>
int PartDev::advance(const Dev& d, const AdvCoef& c, const T* const u[3], hipStream_t st,
                     bool eq) {
  if (np == 0) return FR_OK;
  if (eq && (c.last || c.set0)) return FR_EINVAL;
  if (eq)
    hipLaunchKernelGGL((k_part_advance<T, PH_ALL, true>), dim3(nblocks(np, 256)), dim3(256), 0,
                       st, d, ptrs(), c, u[0], u[1], u[2], nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_part_advance<T, PH_ALL>), dim3(nblocks(np, 256)), dim3(256), 0, st, d,
                       ptrs(), c, u[0], u[1], u[2], nullptr, nullptr, nullptr);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int PartDev::ensure_gas() {
  if (gas_cap >= cap && gas[0]) return FR_OK;
  for (auto& g : gas) {
    (void)hipFree(g);
    g = nullptr;
  }
  gas_cap = 0;
  for (auto& g : gas) FR_CHECK(hipMalloc(&g, sizeof(double) * cap));
  gas_cap = cap;
  return FR_OK;
}

int PartDev::advance_phase(const Dev& d, const AdvCoef& c, const double* const u[3], int phase,
                           hipStream_t st) {
  if (phase != PH_GATHER && phase != PH_POS && phase != PH_VEL && phase != PH_TEMP)
    return FR_EINVAL;
  FR_RET(ensure_gas());
  if (np == 0) return FR_OK;
  const dim3 g(nblocks(np, 256)), b(256);
  const double* uu[3] = {u ? u[0] : nullptr, u ? u[1] : nullptr, u ? u[2] : nullptr};
  if (phase == PH_GATHER && !(uu[0] && uu[1] && uu[2])) return FR_EINVAL;
  if (phase == PH_GATHER)
    hipLaunchKernelGGL((k_part_advance<double, PH_GATHER>), g, b, 0, st, d, ptrs(), c, uu[0],
                       uu[1], uu[2], gas[0], gas[1], gas[2]);
  else if (phase == PH_POS)
    hipLaunchKernelGGL((k_part_advance<double, PH_POS>), g, b, 0, st, d, ptrs(), c, uu[0], uu[1],
                       uu[2], gas[0], gas[1], gas[2]);
  else if (phase == PH_VEL)
    hipLaunchKernelGGL((k_part_advance<double, PH_VEL>), g, b, 0, st, d, ptrs(), c, uu[0], uu[1],
                       uu[2], gas[0], gas[1], gas[2]);
  else
    hipLaunchKernelGGL((k_part_advance<double, PH_TEMP>), g, b, 0, st, d, ptrs(), c, uu[0],
                       uu[1], uu[2], gas[0], gas[1], gas[2]);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int PartDev::temp_split(const Dev& d, int stage, int mode, const double* T, double iTth,
                        hipStream_t st) {
  if ((stage != 1 && stage != 2) || (mode != 0 && mode != 1) || (mode == 0 && !T))
    return FR_EINVAL;
  FR_RET(ensure_gas());
  if (np == 0) return FR_OK;
  hipLaunchKernelGGL(k_part_temp_split, dim3(nblocks(np, 256)), dim3(256), 0, st, d, ptrs(),
                     stage, mode, T, gas[3], iTth);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

template <typename T>
int PartDev::temp(const Dev& d, int stage, double c0, const T* rho, double iTth,
                  hipStream_t st) {
  if (np == 0) return FR_OK;
  hipLaunchKernelGGL(k_part_temp<T>, dim3(nblocks(np, 256)), dim3(256), 0, st, d, ptrs(), stage,
                     c0, rho, iTth);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int PartDev::build(const Dev& d, int stage, int mode, hipStream_t st, bool operands,
                   const int* rk) {
  FR_RET(build_head(d, stage, mode, st, rk));
  return build_tail(d, stage, mode, st, operands, rk);
}

int PartDev::build_head(const Dev& d, int stage, int mode, hipStream_t st, const int* rk) {
  Buckets& B = mode == 0 ? R : T;
  if (!rk) rk = rank;
  // the operand layout follows the deposit form chosen now; the deposit
  // reads whichever layout its buckets were built with
  B.rec_form = dep_kind(mode == 0) == DEP_SPLIT;
  B.valid = false;
  FR_CHECK(hipMemsetAsync(B.cnt, 0, sizeof(int) * ncell, st));
  if (np > 0)
    hipLaunchKernelGGL(k_bucket_count, dim3(nblocks(np, 256)), dim3(256), 0, st, d, ptrs(), stage,
                       mode, B.cnt, B.key, B.slot);
  int rc = scan_ex(B.cnt, ncell, B.start, B.scan_tmp, st);
  if (rc) return rc;
  if (np > 0)
    hipLaunchKernelGGL(k_bucket_scatter, dim3(nblocks(np, 256)), dim3(256), 0, st, np, B.key,
                       B.slot, B.start, rk, B.list2);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int PartDev::build_tail(const Dev& d, int stage, int mode, hipStream_t st, bool operands,
                        const int* rk) {
  Buckets& B = mode == 0 ? R : T;
  if (!rk) rk = rank;
  if (np > 0 && B.rec_form) {
    if (operands)
      hipLaunchKernelGGL(k_bucket_place_rec<true>, dim3(nblocks(np, 256)), dim3(256), 0, st, np,
                         mode, ptrs(), stage, B.key, B.start, rk, B.list2, B.list, d.idx, d.idy,
                         d.idz, B.rec);
    else
      hipLaunchKernelGGL(k_bucket_place_rec<false>, dim3(nblocks(np, 256)), dim3(256), 0, st, np,
                         mode, ptrs(), stage, B.key, B.start, rk, B.list2, B.list, d.idx, d.idy,
                         d.idz, B.rec);
  } else if (np > 0) {
    hipLaunchKernelGGL(k_bucket_place, dim3(nblocks(np, 256)), dim3(256), 0, st, np, B.key,
                       B.start, rk, B.list2, B.list, B.slot);
    if (operands)
      hipLaunchKernelGGL(k_bucket_gather, dim3(nblocks(np, 256)), dim3(256), 0, st, np, mode,
                         ptrs(), stage, B.slot, d.idx, d.idy, d.idz, B.s[0], B.s[1], B.s[2],
                         B.s[3], B.s[4], B.s[5]);
  }
  FR_LAUNCH_CHECK();
  B.valid = true;
  return FR_OK;
}

static dim3 grid3p(const Dev& d) { return dim3(nblocks(d.Nx, 64), nblocks(d.Ny, 4), d.Nz); }

template <typename T>
int PartDev::deposit_cic(const Dev& d, bool weighted, T* out, bool halo, hipStream_t st,
                         double* sum, bool* summed) {
  if (!R.valid) return FR_EINVAL;
  if (summed) *summed = false;
  // Dense tiles (over PCAP contributions) get their own kernel when the mean
  // tile holds more than PCAP/2 contributions; at lower mean densities the
  // few dense tiles merge in place, which spares every call a launch that
  // reserves 76 KB of LDS per workgroup.
  const dim3 g = grid3p(d);
  const int kind = R.rec_form ? DEP_SPLIT : dep_kind(true) == DEP_TILE ? DEP_TILE : DEP_PULL;
  if (kind != DEP_TILE) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    t_cic.next(&e0, &e1);
    const dim3 gs(nblocks(d.NxNy, SNT), nblocks(d.Nz, SKZ));
    // the weighted deposit's interior sum after the fold, from its partials
    const bool part = kind == DEP_SPLIT && weighted && sum && summed && !heat_sum_field &&
                      (long)gs.x * gs.y <= cic_nparts;
    if (kind == DEP_SPLIT && weighted)
      hipExtLaunchKernelGGL((k_dep_cic_split<T, true>), gs, dim3(SNT), 0, st, e0, e1, 0, d,
                            R.start, R.list, R.rec, out, part ? cic_part : nullptr);
    else if (kind == DEP_SPLIT)
      hipExtLaunchKernelGGL((k_dep_cic_split<T, false>), gs, dim3(SNT), 0, st, e0, e1, 0, d,
                            R.start, R.list, R.rec, out, nullptr);
    else
      hipExtLaunchKernelGGL(k_dep_cic_pull<T>, g, dim3(DTX, DTY), 0, st, e0, e1, 0, d, R.start,
                            R.list, R.s[0], R.s[1], R.s[2], weighted ? R.s[3] : nullptr, out);
    FR_LAUNCH_CHECK();
    if (part) {
      FR_RET(sum_parts(cic_part, (long)gs.x * gs.y, sum, st));
      *summed = true;
    }
    int rc = cum_fold(d, out, st);
    if (!rc && halo) rc = ghost_fill(d, out, st);
    return rc;
  }
  const double cells = (double)d.Nx * d.Ny * d.Nz;
  const bool defer = (double)np * 8 * DTX * DTY > 0.5 * PCAP * cells && np < (1 << 28);
  int* dl = defer ? (weighted ? dense_w : dense_c) : nullptr;
  if (dl) FR_CHECK(hipMemsetAsync(dl, 0, sizeof(int), st));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  t_cic.next(&e0, &e1);
  hipExtLaunchKernelGGL(k_dep_cic<T>, g, dim3(64, 4), 0, st, e0, e1, 0, d, R.start, R.list,
                        R.s[0], R.s[1], R.s[2], weighted ? R.s[3] : nullptr, out, dl);
  FR_LAUNCH_CHECK();
  if (dl) {
    // two workgroups of k_dep_cic_dense fit a CU (LDS): one resident round
    const int nt = (int)(g.x * g.y * g.z);
    hipLaunchKernelGGL(k_dep_cic_dense<T>, dim3(nt < 512 ? nt : 512), dim3(64, 4), 0, st, d,
                       R.start, R.list, R.s[0], R.s[1], R.s[2], weighted ? R.s[3] : nullptr, out,
                       dl, (int)g.x, (int)g.y);
    FR_LAUNCH_CHECK();
  }
  int rc = cum_fold(d, out, st);
  if (!rc && halo) rc = ghost_fill(d, out, st);
  return rc;
}

template <typename F>
int PartDev::deposit_vel(const Dev& d, F* const out[3], hipStream_t st, bool halo) {
  if (!T.valid) return FR_EINVAL;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  t_vel.next(&e0, &e1);
  const int kind = T.rec_form ? DEP_SPLIT : dep_kind(false) == DEP_TILE ? DEP_TILE : DEP_PULL;
  if (kind == DEP_SPLIT)
    hipExtLaunchKernelGGL(k_dep_vel_split<F>, dim3(nblocks(d.NxNy, SNT), nblocks(d.Nz, SKZ)),
                          dim3(SNT), 0, st, e0, e1, 0, d, T.start, T.list, T.rec, out[0], out[1],
                          out[2]);
  else if (kind == DEP_PULL)
    hipExtLaunchKernelGGL(k_dep_vel_pull<F>, grid3p(d), dim3(DTX, DTY), 0, st, e0, e1, 0, d,
                          T.start, T.list, T.s[0], T.s[1], T.s[2], T.s[3], T.s[4], T.s[5], out[0],
                          out[1], out[2]);
  else
    hipExtLaunchKernelGGL(k_dep_vel<F>, grid3p(d), dim3(64, 4), 0, st, e0, e1, 0, d, T.start,
                          T.list, T.s[0], T.s[1], T.s[2], T.s[3], T.s[4], T.s[5], out[0], out[1],
                          out[2]);
  FR_LAUNCH_CHECK();
  int rc = cum_fold_n(d, out, 3, st);  // the three components, one launch per phase
  return rc || !halo ? rc : ghost_fill_n(d, out, 3, st);
}

// particle::Send_Recv on one rank: axes z, y, x (particle.cpp:853-858).  The
// event order is computed on reference indices; only the reference indices of
// the moved particles change, and the leavers' coordinates are wrapped.
int PartDev::migrate(const double lo[3], const double hi[3], const double L[3], hipStream_t st,
                     bool duct, const double* inj, int n_in, bool inj_dev) {
  // The leaver count K of each axis stays on the device: the kernels read it
  // from pNL[np] and loop over the moves, so the host does not wait for it
  // (the duct's outflow, which shrinks the set, is the exception).
  if (!flg_zero) FR_CHECK(hipMemsetAsync(flg, 0, sizeof(int) * (cap + 1), st));
  flg_zero = false;  // until every axis below has cleared its flags
  for (int ax = 2; ax >= 0 && np > 0; ax--) {
    const unsigned nb = nblocks(np, 256);
    hipLaunchKernelGGL(k_mig_flag, dim3(nb), dim3(256), 0, st, np, (const double*)s[ax][2], rank,
                       lo[ax], hi[ax], flg);
    int rc = scan_ex_t<true>(flg, np, pNL, mig_tmp, st);  // pNL: stayers before each index
    if (rc) return rc;
    hipLaunchKernelGGL(k_mig_rank, dim3(nb), dim3(256), 0, st, np, flg, pNL, holepos, tailnl);
    hipLaunchKernelGGL(k_mig_event, dim3(nb), dim3(256), 0, st, np, flg, pNL, tailnl, evt, islo);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(ST), 0, st, islo, np, pNL + np, plo);
    MigAxis m{ax, L[ax]};
    // h + nleave <= np moves (h <= K front holes, nleave = np - K leavers)
    const unsigned ng = std::min(nblocks(np, 256), 1024u);
    hipLaunchKernelGGL(k_mig_moves, dim3(ng), dim3(256), 0, st, np, pNL, holepos, tailnl, evt,
                       plo, inv, mslot, dst, mlv);
    hipLaunchKernelGGL(k_mig_apply, dim3(ng), dim3(256), 0, st, np, ptrs(), m, pNL, mslot, dst,
                       mlv, rank, inv);
    FR_LAUNCH_CHECK();
    if (duct && ax == 0) {
      FR_CHECK(hipMemcpyAsync(hbuf, pNL + np, sizeof(int), hipMemcpyDeviceToHost, st));
      FR_CHECK(hipStreamSynchronize(st));
      const int K = hbuf[0];
      if (K == np) continue;
      // the leavers now hold reference indices K..np-1 and are not received
      // back: keep the slots with rank < K, in storage order
      hipLaunchKernelGGL(k_live, dim3(nb), dim3(256), 0, st, np, rank, K, nl);
      rc = scan_ex(nl, np, pNL, mig_tmp, st);
      if (rc) return rc;
      PartPtrs a = ptrs(), b = alt;
      hipLaunchKernelGGL(k_compact, dim3(nb), dim3(256), 0, st, np, nl, pNL, a, rank, b,
                         rank_alt);
      FR_LAUNCH_CHECK();
      swap_storage();
      np = K;
      if (np > 0)
        hipLaunchKernelGGL(k_inv, dim3(nblocks(np, 256)), dim3(256), 0, st, np, rank, inv);
      FR_LAUNCH_CHECK();
    }
  }
  flg_zero = true;
  if (duct && n_in > 0) {  // FromBuf of the inlet particles (particle.cpp:286-301)
    if (np + n_in > cap) {
      fprintf(stderr, "fr_amd: duct particle capacity %d exceeded (%d + %d inlet)\n", cap, np,
              n_in);
      return FR_EOVERFLOW;
    }
    const double* src = inj;
    if (!inj_dev) {
      if (n_in > inj_cap) {
        (void)hipFree(inj_buf);
        inj_buf = nullptr;
        FR_CHECK(hipMalloc(&inj_buf, sizeof(double) * 29 * (size_t)n_in));
        inj_cap = n_in;
      }
      FR_CHECK(hipMemcpyAsync(inj_buf, inj, sizeof(double) * 29 * (size_t)n_in,
                              hipMemcpyHostToDevice, st));
      src = inj_buf;
    }
    hipLaunchKernelGGL(k_inject, dim3(nblocks(n_in, 256)), dim3(256), 0, st, n_in, np, src,
                       ptrs(), rank, inv);
    FR_LAUNCH_CHECK();
    np += n_in;
    if (!inj_dev) FR_CHECK(hipStreamSynchronize(st));  // the host buffer may be reused
  }
  R.valid = false;
  T.valid = false;
  return FR_OK;
}

int PartDev::gather_hi(double* buf, int cap, int* n_out, hipStream_t st) {
  const unsigned nb = std::max(1u, std::min(nblocks(np, 256), 64u));
  hipLaunchKernelGGL(k_gather_hi, dim3(nb), dim3(256), 0, st, np, (const int*)pNL,
                     (const int*)plo, (const int*)inv, ptrs(), buf, cap, n_out);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

void PartDev::swap_storage() {
  for (int q = 0; q < 7; q++)
    for (int t = 0; t < 4; t++) std::swap(s[q][t], alt.s[q][t]);
  std::swap(flag, alt.flag);
  std::swap(rhs_u, alt.rhs_u);
  std::swap(rhs_v, alt.rhs_v);
  std::swap(rhs_w, alt.rhs_w);
  std::swap(rhs_T, alt.rhs_T);
  std::swap(rank, rank_alt);
}

void PartDev::swap_stages(int a, int b) {
  for (int q = 0; q < 7; q++) std::swap(s[q][a], s[q][b]);
}

int PartDev::copy_stage(int dst_stage, int src_stage, hipStream_t st) {
  if (np == 0) return FR_OK;
  for (int q = 0; q < 7; q++)
    FR_CHECK(hipMemcpyAsync(s[q][dst_stage], s[q][src_stage], sizeof(double) * np,
                            hipMemcpyDeviceToDevice, st));
  return FR_OK;
}

int PartDev::reset_order(hipStream_t st) {
  if (np > 0)
    hipLaunchKernelGGL(k_iota2, dim3(nblocks(np, 256)), dim3(256), 0, st, np, rank, inv);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

// Storage slots re-sorted by truncating cell index of the base positions,
// ties by reference index (the truncating buckets of stage 0).  Particle
// values, reference indices and buckets of the rounded index are unchanged.
int PartDev::resort(const Dev& d, hipStream_t st) {
  if (np < 2) return FR_OK;
  int rc = build(d, 0, 1, st, false);
  if (rc) return rc;
  PartPtrs a = ptrs(), b = alt;
  b.np = np;
  int* diff = dflag;
  FR_CHECK(hipMemsetAsync(diff, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_stage_eq, dim3(nblocks(np, 256)), dim3(256), 0, st, np, a, diff);
  hipLaunchKernelGGL(k_resort, dim3(nblocks(np, 256)), dim3(256), 0, st, np, T.list, inv, a, b,
                     rank_alt, diff);
  FR_LAUNCH_CHECK();
  swap_storage();
  T.valid = false;
  return FR_OK;
}

// Host transfer in reference order; `alt` storage is the staging buffer.
int PartDev::get(double* host, const double* src, hipStream_t st) {
  if (np == 0) return FR_OK;
  double* tmp = alt.s[0][0];
  hipLaunchKernelGGL(k_to_ref<double>, dim3(nblocks(np, 256)), dim3(256), 0, st, np, rank, src,
                     tmp);
  FR_LAUNCH_CHECK();
  FR_CHECK(hipMemcpyAsync(host, tmp, sizeof(double) * np, hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

int PartDev::set(const double* host, double* dst, hipStream_t st) {
  if (np == 0) return FR_OK;
  double* tmp = alt.s[0][0];
  FR_CHECK(hipMemcpyAsync(tmp, host, sizeof(double) * np, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_from_ref<double>, dim3(nblocks(np, 256)), dim3(256), 0, st, np, rank, tmp,
                     dst);
  FR_LAUNCH_CHECK();
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

int PartDev::get_flags(int* host, hipStream_t st) {
  if (np == 0) return FR_OK;
  hipLaunchKernelGGL(k_to_ref<int>, dim3(nblocks(np, 256)), dim3(256), 0, st, np, rank,
                     (const int*)flag, alt.flag);
  FR_LAUNCH_CHECK();
  FR_CHECK(hipMemcpyAsync(host, alt.flag, sizeof(int) * np, hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

int PartDev::set_flags(const int* host, hipStream_t st) {
  if (np == 0) return FR_OK;
  FR_CHECK(hipMemcpyAsync(alt.flag, host, sizeof(int) * np, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_from_ref<int>, dim3(nblocks(np, 256)), dim3(256), 0, st, np, rank,
                     (const int*)alt.flag, flag);
  FR_LAUNCH_CHECK();
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

template int PartDev::advance<double>(const Dev&, const AdvCoef&, const double* const[3],
                                      hipStream_t, bool);
template int PartDev::advance<float>(const Dev&, const AdvCoef&, const float* const[3],
                                     hipStream_t, bool);
template int PartDev::temp<double>(const Dev&, int, double, const double*, double, hipStream_t);
template int PartDev::temp<float>(const Dev&, int, double, const float*, double, hipStream_t);
template int PartDev::deposit_cic<double>(const Dev&, bool, double*, bool, hipStream_t, double*,
                                          bool*);
template int PartDev::deposit_cic<float>(const Dev&, bool, float*, bool, hipStream_t, double*,
                                         bool*);
template int PartDev::deposit_vel<double>(const Dev&, double* const[3], hipStream_t, bool);
template int PartDev::deposit_vel<float>(const Dev&, float* const[3], hipStream_t, bool);

}  // namespace fr
