// Box pressure solve on gfx950: the defect-correction loop of poisson::FFT
// (poisson.cpp:391-411) around the constant-coefficient spectral solve
// poisson::CCP_FFT (poisson.cpp:413-462).
//
// The residual is real, so the x transform is real-to-complex and only the
// nx/2+1 non-negative x frequencies are carried through the y and z passes.
// The spectrum is stored (kz, kx, y)-major, element ((k*nxh + kx)*ny + j):
// the y passes read and write whole contiguous pencils, the z pass rows of W
// consecutive y.
//   pass X : residual (computed inline) -> real FFT along x         (rows)
//   pass Y : FFT along y                                             (strided)
//   pass Z : FFT along z, divide by lambda(k), inverse FFT along z
//   pass Y': inverse FFT along y
//   pass X': inverse real FFT along x, dp = -Re/N, P += dp, max|dp|
// Division by the real, even lambda(k) keeps the spectrum Hermitian, so the
// result equals the full complex transform of the reference's solve.
//
// 1-D transforms: in-place radix-2 decimation in time on bit-reversed input
// with host twiddles (cos(2 pi k/n), sign*sin(2 pi k/n)); the real transform
// of length N packs x_2m + i x_2m+1 into an N/2-point complex transform and
// unpacks with the length-N twiddles.  The CPU oracle uses the same
// algorithm and operation order, so the two agree bit for bit.  A workgroup
// holds W pencils of length L in LDS, padded by one complex per pencil.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include <hip/hip_ext.h>

#include "fft_lds.hpp"

namespace fr {

struct PoisDev {
  Dev d;
  int nx, ny, nz;      // interior sizes
  int nxh;             // nx/2 + 1 retained x frequencies
  int lxm, ly, lz;     // log2(nx/2), log2(ny), log2(nz)
  int wx, wy, wz;      // pencils per workgroup for each pass (powers of two)
  int lwx, lwy, lwz;   // their log2
  int lxz;             // forward x pass: log2 of the z planes of its (y, z) row tile
  int wxi, lwxi;       // inverse x pass: rows per workgroup, log2
  int lgx, lgxi;       // log2 of the neighbouring tiles per XCD group (xcd_group), fwd/inv
  double tot;          // nx*ny*nz
  double itot;         // 1/tot, exact (tot is a power of two): x*itot == x/tot
};

// Pass X: per row (j,k) the residual, packed as z_m = r_2m + i r_2m+1, an
// nx/2-point FFT, then the real-transform unpacking to X_k, k = 0..nx/2:
//   E = (Z_k + conj Z_{M-k})/2, O = (Z_k - conj Z_{M-k})/(2i), X = E + W^k O.
// R: real type of the spectrum (double; float for the precision form's inner
// transform, where the residual is still formed in double and rounded once).
template <typename R>
struct Cx;
template <>
struct Cx<double> {
  using T = double2;
};
template <>
struct Cx<float> {
  using T = float2;
};

// CF, the coefficient form: CF_C (cx, cy, cz hold C), CF_FACE (they hold
// Rho_face; C = 1./Rho_face), CF_CELL (cx holds the cell-centred density the
// face values are interpolated from; cy and cz are not read).
enum { CF_C = 0, CF_FACE = 1, CF_CELL = 2 };
template <int CF, typename TC, typename R>
__global__ void __launch_bounds__(FT) k_pass_x_fwd(PoisDev q, const double* __restrict__ P,
                                                   const TC* __restrict__ cx,
                                                   const TC* __restrict__ cy,
                                                   const TC* __restrict__ cz,
                                                   const double* __restrict__ RHS,
                                                   const typename Cx<R>::T* __restrict__ twM,
                                                   const typename Cx<R>::T* __restrict__ twN,
                                                   typename Cx<R>::T* __restrict__ out,
                                                   const int* __restrict__ gate,
                                                   const double* __restrict__ rsum) {
  using C2 = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  C2* sm = reinterpret_cast<C2*>(smraw);
  if (*gate) return;  // the solve has converged: a queued iteration does nothing
  // rsum: the right-hand side's interior sum and minus its mean (rsum[1],
  // the value make_mean_zero adds), the mean not yet removed from the field;
  // the shift is read where it is added, after the residual's loads
  const bool sh = rsum != nullptr;
  const Dev& d = q.d;
  const int M = q.nx >> 1, W = q.wx, ps = M + 1;
  C2* buf = sm;
  C2* twm = sm + W * ps;
  C2* twn = twm + (M / 2 + 1);
  load_tw(twm, twM, M);
  load_tw(twn, twN, q.nx);
  const int lm = q.lxm, lw = q.lwx;
  // The workgroup's W rows are a tile of 2^lwy_ consecutive y in 2^lxz
  // consecutive planes: the residual's neighbour rows in y and z are shared
  // inside the tile.  Plain dispatch order: the neighbouring tiles run at
  // the same time across the XCDs (measured: an XCD-contiguous order fetches
  // 16% more and runs 8% slower here).
  const int lwy_ = q.lwx - q.lxz;
  const unsigned bid = xcd_group(blockIdx.x, q.lgx, gridDim.x);
  const int j0 = (int)(bid & ((1u << (q.ly - lwy_)) - 1)) << lwy_;
  const int k0 = (int)(bid >> (q.ly - lwy_)) << q.lxz;
  for (int t = threadIdx.x; t < W * M; t += FT) {
    const int p = t >> lm, m = t & (M - 1);
    const int jj = j0 + (p & ((1 << lwy_) - 1)), kk = k0 + (p >> lwy_);
    double r0 = 0.0, r1 = 0.0;
    if (kk < q.nz) {
      const long c = at(d, 2 * m + 1, jj + 1, kk + 1);
      const double shift = sh ? rsum[1] : 0.0;
      if constexpr (CF == CF_CELL) {
        r0 = residual_cell<TC>(d, P, cx, RHS, c, sh, shift);
        r1 = residual_cell<TC>(d, P, cx, RHS, c + 1, sh, shift);
      } else {
        r0 = residual_t<CF == CF_FACE, TC>(d, P, cx, cy, cz, RHS, c, sh, shift);
        r1 = residual_t<CF == CF_FACE, TC>(d, P, cx, cy, cz, RHS, c + 1, sh, shift);
      }
    }
    C2 v;
    v.x = (R)r0;
    v.y = (R)r1;
    buf[p * ps + bitrev(m, q.lxm)] = v;
  }
  fft_lds(buf, M, q.lxm, W, ps, twm);
  for (int t = threadIdx.x; t < W * (M + 1); t += FT) {
    const int k = t >> lw, p = t & (W - 1);  // pencil fastest: consecutive y in the spectrum
    const int jj = j0 + (p & ((1 << lwy_) - 1)), kk = k0 + (p >> lwy_);
    if (kk >= q.nz) continue;
    const C2 A = buf[p * ps + (k == M ? 0 : k)];
    const C2 B = buf[p * ps + (k == 0 ? 0 : M - k)];
    const R Er = (R)0.5 * (A.x + B.x), Ei = (R)0.5 * (A.y - B.y);
    const R Dr = A.x - B.x, Di = A.y + B.y;
    const R Or = (R)0.5 * Di, Oi = (R)-0.5 * Dr;
    const C2 w = twn[k];
    const R Tr = Or * w.x - Oi * w.y, Ti = Or * w.y + Oi * w.x;
    C2 o;
    o.x = Er + Tr;
    o.y = Ei + Ti;
    out[((long)kk * (M + 1) + k) * q.ny + jj] = o;
  }
}

// Strided pass along y (axis 1) or z (axis 2) over a tile of W consecutive x
// frequencies.  For axis 2 with `divide` set: forward FFT, divide by lambda
// except the zero mode (poisson.cpp:438-451), then inverse FFT (tw2) in LDS.
template <typename R>
__global__ void __launch_bounds__(FT) k_pass_strided(PoisDev q, int axis, int divide,
                                                     const typename Cx<R>::T* __restrict__ tw,
                                                     const typename Cx<R>::T* __restrict__ tw2,
                                                     const double* __restrict__ lamx,
                                                     const double* __restrict__ lamy,
                                                     const double* __restrict__ lamz,
                                                     typename Cx<R>::T* __restrict__ data,
                                                     const int* __restrict__ gate) {
  using C2 = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  C2* sm = reinterpret_cast<C2*>(smraw);
  if (*gate) return;
  const int L = axis == 1 ? q.ny : q.nz;
  const int lg = axis == 1 ? q.ly : q.lz;
  const int W = axis == 1 ? q.wy : q.wz;
  const int lw = axis == 1 ? q.lwy : q.lwz;
  const int ps = L + 1;
  const int nxh = q.nxh;
  C2* buf = sm;
  C2* tws = sm + W * ps;
  C2* tws2 = tws + (L / 2 + 1);
  load_tw(tws, tw, L);
  if (divide) load_tw(tws2, tw2, L);
  // axis 1: W pencils (kx = i0 + p) of plane k = other, each contiguous in y;
  // axis 2: W pencils (y = i0 + p) of frequency kx = other, stride nxh*ny in z
  const int span = axis == 1 ? nxh : q.ny;  // range of the tile index
  const int tiles = (span + W - 1) >> lw;
  const unsigned bid = xcd_block().x;
  const int other = (int)(bid / (unsigned)tiles);
  const int i0 = ((int)bid - other * tiles) << lw;
  if (axis == 1) {
    const long base = ((long)other * nxh + i0) * q.ny;
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int p = t >> lg, e = t & (L - 1);
      if (i0 + p < nxh) buf[p * ps + bitrev(e, lg)] = data[base + (long)p * q.ny + e];
    }
  } else {
    const long stride = (long)nxh * q.ny, base = (long)other * q.ny + i0;
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int e = t >> lw, p = t & (W - 1);
      if (i0 + p < q.ny) buf[p * ps + bitrev(e, lg)] = data[base + e * stride + p];
    }
  }
  fft_lds(buf, L, lg, W, ps, tws);
  if (divide) {
    // lambda = (ax + ay) + az, host tables from the reference formula
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int e = t >> lw, p = t & (W - 1);
      const int ix = other, jy = i0 + p, kz = e;  // divide runs on axis 2 only
      if (jy < q.ny && ix + jy + kz > 0) {
        const R lam = (R)(lamx[ix] + lamy[jy] + lamz[kz]);
        C2 v = buf[p * ps + e];
        v.y /= lam;
        v.x /= lam;
        buf[p * ps + e] = v;
      }
    }
    __syncthreads();
    // bit-reverse permutation in place for the inverse transform
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int p = t >> lg, e = t & (L - 1);
      const int r = bitrev(e, lg);
      if (e < r) {
        C2 a = buf[p * ps + e];
        buf[p * ps + e] = buf[p * ps + r];
        buf[p * ps + r] = a;
      }
    }
    fft_lds(buf, L, lg, W, ps, tws2);
  }
  if (axis == 1) {
    const long base = ((long)other * nxh + i0) * q.ny;
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int p = t >> lg, e = t & (L - 1);
      if (i0 + p < nxh) data[base + (long)p * q.ny + e] = buf[p * ps + e];
    }
  } else {
    const long stride = (long)nxh * q.ny, base = (long)other * q.ny + i0;
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int e = t >> lw, p = t & (W - 1);
      if (i0 + p < q.ny) data[base + e * stride + p] = buf[p * ps + e];
    }
  }
}

// Passes Y, Y' and Z for pencils of 256 or 512 points: one pencil per
// 64-lane wave held in registers (wave_fft), no workgroup barrier inside the
// transforms.  The y pencils are contiguous, so each register slot's loads
// and stores are 64 consecutive elements; the z pass takes WZ_WAVES
// consecutive y pencils of one x frequency per workgroup, so its strided
// rows are WZ_WAVES elements wide.  Bit-identical to k_pass_strided.
constexpr int WY_WAVES = 4, WZ_WAVES = 8;

template <typename R, int LG>
__global__ void __launch_bounds__(64 * WY_WAVES) k_pass_y_wave(
    PoisDev q, const typename Cx<R>::T* __restrict__ tw, typename Cx<R>::T* __restrict__ data,
    const int* __restrict__ gate) {
  using C2 = typename Cx<R>::T;
  using P = WavePlan<LG>;
  __shared__ C2 tws[P::L / 2 + 1];
  __shared__ C2 exs[WY_WAVES][P::XS];
  if (*gate) return;
  for (int t = threadIdx.x; t <= P::L / 2; t += 64 * WY_WAVES) tws[t] = tw[t];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long pen = (long)blockIdx.x * WY_WAVES + wave;  // (kz, kx) pencil, P::L == ny
  if (pen >= (long)q.nz * q.nxh) return;
  C2* p = data + pen * P::L;
  C2 x[P::E];
#pragma unroll
  for (int k = 0; k < P::E; k++) x[k] = p[wave_in_index<LG>(lane, k)];
  wave_fft<LG>(x, exs[wave], lane, tws);
#pragma unroll
  for (int k = 0; k < P::E; k++) p[wave_out_index<LG>(lane, k)] = x[k];
}

// The z pass stages its WZ_WAVES pencils through the waves' exchange areas:
// the workgroup loads and stores (z, y) tiles whose WZ_WAVES consecutive y
// elements are one contiguous 128-byte row segment (one memory request per
// segment instead of one per element), and each wave then reads its pencil
// from LDS.  The area pitch is XS + 2 elements, so the WZ_WAVES pencils of a
// staged z start on different LDS banks.  The inverse transform reads the
// forward twiddles conjugated (the host's inverse table is exactly that), so
// one table and the areas fit two workgroups per CU.
template <typename R, int LG>
__global__ void __launch_bounds__(64 * WZ_WAVES) k_pass_z_wave(
    PoisDev q, const typename Cx<R>::T* __restrict__ tw, const double* __restrict__ lamx,
    const double* __restrict__ lamy, const double* __restrict__ lamz,
    typename Cx<R>::T* __restrict__ data, const int* __restrict__ gate) {
  using C2 = typename Cx<R>::T;
  using P = WavePlan<LG>;
  constexpr int XP = P::XS + 2;
  __shared__ C2 twf[P::L / 2 + 1];
  __shared__ C2 exs[WZ_WAVES][XP];
  if (*gate) return;
  for (int t = threadIdx.x; t <= P::L / 2; t += 64 * WZ_WAVES) twf[t] = tw[t];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles = q.ny / WZ_WAVES;
  const unsigned bid = xcd_block().x;
  const int kx = (int)(bid / (unsigned)tiles);
  const int j0 = ((int)bid - kx * tiles) * WZ_WAVES, jy = j0 + wave;
  const long stride = (long)q.nxh * q.ny;  // P::L == nz
  C2* base = data + (long)kx * q.ny + j0;
  static_assert(WZ_WAVES == 8, "staging maps 8 consecutive y per z");
  for (int t = threadIdx.x; t < WZ_WAVES * P::L; t += 64 * WZ_WAVES) {
    const int y = t & (WZ_WAVES - 1), z = t >> 3;
    exs[y][wave_xi(z)] = base[(long)z * stride + y];
  }
  __syncthreads();
  C2 x[P::E];
#pragma unroll
  for (int k = 0; k < P::E; k++) x[k] = exs[wave][wave_xi(wave_in_index<LG>(lane, k))];
  wave_sync();
  wave_fft<LG>(x, exs[wave], lane, twf);
  // divide by lambda except the zero mode (poisson.cpp:438-451)
  const double lxy = lamx[kx] + lamy[jy];
#pragma unroll
  for (int k = 0; k < P::E; k++) {
    const int kz = wave_out_index<LG>(lane, k);
    if (kx + jy + kz > 0) {
      const R lam = (R)(lxy + lamz[kz]);
      x[k].y /= lam;
      x[k].x /= lam;
    }
  }
  wave_to_bitrev<LG>(x, exs[wave], lane);
  wave_fft<LG, true>(x, exs[wave], lane, twf);
#pragma unroll
  for (int k = 0; k < P::E; k++) exs[wave][wave_xi(wave_out_index<LG>(lane, k))] = x[k];
  __syncthreads();
  for (int t = threadIdx.x; t < WZ_WAVES * P::L; t += 64 * WZ_WAVES) {
    const int y = t & (WZ_WAVES - 1), z = t >> 3;
    base[(long)z * stride + y] = exs[y][wave_xi(z)];
  }
}

// Pass X': inverse real FFT along x: Z_k = (X_k + conj X_{M-k}) +
// i w^k (X_k - conj X_{M-k}), k < M, an nx/2-point inverse FFT, x_2m = Re z_m,
// x_2m+1 = Im z_m; dp = (-1)*x/N (poisson.cpp:455-461); P += dp at the
// interior; per-block max|dp| (tensor0::max_abs).
template <typename R>
__global__ void __launch_bounds__(FT) k_pass_x_inv(PoisDev q,
                                                   const typename Cx<R>::T* __restrict__ in,
                                                   const typename Cx<R>::T* __restrict__ twM,
                                                   const typename Cx<R>::T* __restrict__ twN,
                                                   double* __restrict__ P,
                                                   double* __restrict__ bmax,
                                                   const int* __restrict__ gate) {
  using C2 = typename Cx<R>::T;
  extern __shared__ __align__(16) unsigned char smraw[];
  C2* sm = reinterpret_cast<C2*>(smraw);
  __shared__ double red[FT];
  if (*gate) return;
  const Dev& d = q.d;
  const int M = q.nx >> 1, W = q.wxi, ps = M + 1;
  C2* buf = sm;
  C2* twm = sm + W * ps;
  C2* twn = twm + (M / 2 + 1);
  load_tw(twm, twM, M);
  load_tw(twn, twN, q.nx);
  __syncthreads();
  const int rows = q.ny * q.nz, lm = q.lxm, lw = q.lwxi;
  const unsigned bid = xcd_group(blockIdx.x, q.lgxi, gridDim.x);
  const int row0 = (int)bid * W;
  for (int t = threadIdx.x; t < W * M; t += FT) {
    const int k = t >> lw, p = t & (W - 1);  // pencil fastest: consecutive y in the spectrum
    const int row = row0 + p;
    C2 z;
    z.x = (R)0;
    z.y = (R)0;
    if (row < rows) {
      const int jj = row & (q.ny - 1), kk = row >> q.ly;
      const long b = (long)kk * (M + 1) * q.ny + jj;
      const C2 A = in[b + (long)k * q.ny], B = in[b + (long)(M - k) * q.ny];
      const R Sr = A.x + B.x, Si = A.y - B.y;
      const R Dr = A.x - B.x, Di = A.y + B.y;
      const C2 w = twn[k];
      const R Tr = Dr * w.x - Di * w.y, Ti = Dr * w.y + Di * w.x;
      z.x = Sr - Ti;
      z.y = Si + Tr;
    }
    buf[p * ps + bitrev(k, q.lxm)] = z;
  }
  fft_lds(buf, M, q.lxm, W, ps, twm);
  double m = 0.0;
  for (int t = threadIdx.x; t < W * M; t += FT) {
    const int p = t >> lm, e = t & (M - 1);
    const int row = row0 + p;
    if (row < rows) {
      const int jj = row & (q.ny - 1), kk = row >> q.ly;
      const C2 z = buf[p * ps + e];
      const double dp0 = (-1) * (double)z.x * q.itot;
      const double dp1 = (-1) * (double)z.y * q.itot;
      const long c = at(d, 2 * e + 1, jj + 1, kk + 1);
      P[c] = P[c] + dp0;
      P[c + 1] = P[c + 1] + dp1;
      const double a0 = dp0 >= 0 ? dp0 : -dp0, a1 = dp1 >= 0 ? dp1 : -dp1;
      m = a0 > m ? a0 : m;
      m = a1 > m ? a1 : m;
    }
  }
  block_max_store(m, red, bmax + bid);
}

// One workgroup of MT threads, eight independent loads in flight per thread
// (the per-block maxima of an x pass: 65536 at 512^3), then wave shuffles.
// The max is exact in any order.
constexpr int MT = 1024;
__global__ void __launch_bounds__(MT) k_max_final(const double* __restrict__ b, int n,
                                                  double* __restrict__ out,
                                                  int* __restrict__ gate, double eps) {
  __shared__ double red[MT / 64];
  if (gate && gate[0]) return;
  double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int t = threadIdx.x;
  for (; t + 7 * MT < n; t += 8 * MT) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = b[t + q * MT];
#pragma unroll
    for (int q = 0; q < 8; q++) m[q] = fmax(m[q], v[q]);
  }
  for (; t < n; t += MT) m[0] = fmax(m[0], b[t]);
  double r = fmax(fmax(fmax(m[0], m[1]), fmax(m[2], m[3])),
                  fmax(fmax(m[4], m[5]), fmax(m[6], m[7])));
  for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x >= 64) return;
  r = threadIdx.x < MT / 64 ? red[threadIdx.x] : 0.0;
  for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o));
  if (threadIdx.x == 0) {
    out[0] = r;
    if (gate) {
      gate[1] += 1;
      if (!(r > eps)) gate[0] = 1;
    }
  }
}

int max_final(const double* b, int n, double* out, hipStream_t st, int* gate, double eps) {
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(MT), 0, st, b, n, out, gate, eps);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

}  // namespace fr

using namespace fr;

struct fr_box_poisson {
  fr_geom g;
  PoisDev q;
  // float spectrum, twiddles and pencil widths of the precision form's inner
  // transform (allocated at its first use)
  PoisDev q32;
  float2* data32 = nullptr;
  float2 *twf32[3] = {}, *twi32[3] = {}, *twfm32 = nullptr, *twim32 = nullptr;
  size_t lds32[4] = {};
  int nbx32 = 0, nbxi32 = 0;
  double2* data = nullptr;  // (nx/2+1)*ny*nz complex
  double2 *twf[3] = {}, *twi[3] = {};  // length nx, ny, nz
  double2 *twfm = nullptr, *twim = nullptr;  // length nx/2
  double* lam[3] = {};
  double* bmax = nullptr;
  double* dres = nullptr;
  double* hres = nullptr;
  int nbx = 0, nbxi = 0;  // workgroups of the forward and inverse x passes
  size_t lds[4] = {};  // dynamic LDS bytes of the x, y, z and inverse x passes
  // HIP events bound to each launch of the forward x pass (residual + real
  // FFT) of a batch of queued iterations, read after the batch's host
  // synchronisation for the iterations that ran: summed ms, count
  static constexpr int QB = 16;  // most iterations queued per synchronisation
  hipEvent_t tx0[QB] = {}, tx1[QB] = {};
  double tx_ms = 0;
  long tx_n = 0;
  // device convergence gate (done flag, iterations run) and its host copy
  int* gate = nullptr;
  int* hgate = nullptr;
  int last_iters = 1;  // iterations of the previous solve: the next batch size
};

// One defect-correction iteration; leaves max|dp| in s->dres[0].  cf: the
// coefficient form of k_pass_x_fwd.  R is the real type of the spectral
// solve (the residual and P stay double).
template <typename TC, typename R>
static int box_poisson_iter(fr_box_poisson* s, const TC* cx, const TC* cy, const TC* cz,
                            int cf, double* P, const double* RHS, double eps, int slot,
                            hipStream_t st, const double* rsum) {
  using C2 = typename Cx<R>::T;
  const bool f = sizeof(R) == 4;
  PoisDev& q = f ? s->q32 : s->q;
  C2* data = f ? (C2*)s->data32 : (C2*)s->data;
  C2* const* twf = f ? (C2* const*)s->twf32 : (C2* const*)s->twf;
  C2* const* twi = f ? (C2* const*)s->twi32 : (C2* const*)s->twi;
  const C2* twfm = f ? (const C2*)s->twfm32 : (const C2*)s->twfm;
  const C2* twim = f ? (const C2*)s->twim32 : (const C2*)s->twim;
  const size_t* lds = f ? s->lds32 : s->lds;
  const int nbx = f ? s->nbx32 : s->nbx, nbxi = f ? s->nbxi32 : s->nbxi;
  // the events are bound to the launch (start and end of the kernel itself)
  const int* g = s->gate;
  auto xf = [&](auto kern) {
    hipExtLaunchKernelGGL(kern, dim3(nbx), dim3(FT), lds[0], st, s->tx0[slot], s->tx1[slot], 0,
                          q, P, cx, cy, cz, RHS, twfm, twf[0], data, g, rsum);
  };
  if (cf == CF_CELL)
    xf(k_pass_x_fwd<CF_CELL, TC, R>);
  else if (cf == CF_FACE)
    xf(k_pass_x_fwd<CF_FACE, TC, R>);
  else
    xf(k_pass_x_fwd<CF_C, TC, R>);
  const int ty = (q.nxh + q.wy - 1) / q.wy, tz = (q.ny + q.wz - 1) / q.wz;
  // 256- and 512-point y and z transforms one pencil per wave, other lengths
  // in LDS workgroup tiles
  const dim3 gyw(nblocks((long)q.nz * q.nxh, WY_WAVES)), byw(64 * WY_WAVES);
  auto ypass = [&](const C2* tw) {
    if (q.ly == 9)
      hipLaunchKernelGGL((k_pass_y_wave<R, 9>), gyw, byw, 0, st, q, tw, data, g);
    else if (q.ly == 8)
      hipLaunchKernelGGL((k_pass_y_wave<R, 8>), gyw, byw, 0, st, q, tw, data, g);
    else
      hipLaunchKernelGGL(k_pass_strided<R>, dim3(ty * q.nz), dim3(FT), lds[1], st, q, 1, 0, tw,
                         tw, s->lam[0], s->lam[1], s->lam[2], data, g);
  };
  ypass(twf[1]);
  const bool zw = (q.lz == 9 || q.lz == 8) && q.ny % WZ_WAVES == 0;
  const dim3 gzw(q.ny / WZ_WAVES * q.nxh), bzw(64 * WZ_WAVES);
  if (zw && q.lz == 9)
    hipLaunchKernelGGL((k_pass_z_wave<R, 9>), gzw, bzw, 0, st, q, twf[2], s->lam[0], s->lam[1],
                       s->lam[2], data, g);
  else if (zw)
    hipLaunchKernelGGL((k_pass_z_wave<R, 8>), gzw, bzw, 0, st, q, twf[2], s->lam[0], s->lam[1],
                       s->lam[2], data, g);
  else
    hipLaunchKernelGGL(k_pass_strided<R>, dim3(tz * q.nxh), dim3(FT), lds[2], st, q, 2, 1,
                       twf[2], twi[2], s->lam[0], s->lam[1], s->lam[2], data, g);
  ypass(twi[1]);
  hipLaunchKernelGGL(k_pass_x_inv<R>, dim3(nbxi), dim3(FT), lds[3], st, q, data, twim, twi[0],
                     P, s->bmax, g);
  FR_LAUNCH_CHECK();
  // the periodic ghost copy of P below is idempotent, so it needs no gate
  int rc = max_final(s->bmax, nbxi, s->dres, st, s->gate, eps);
  return rc ? rc : ghost_fill(q.d, P, st);
}

// Pencil widths and LDS bytes of the passes for complex elements of `csize`
// bytes: the y, z and inverse x passes within `want` bytes of LDS, the
// forward x pass within `want_f` (measured at 512^3: 8 rows per workgroup run
// the forward pass 10% faster than 4, while the inverse pass runs 8% slower;
// 16 rows slow both).  lxz: log2 of the forward pass's z planes per tile,
// clamped to the tile and the grid.
static void plan_passes(PoisDev& q, int csize, int want, int want_f, int lxz, size_t lds[4],
                        int* nbx, int* nbxi) {
  const int nx = q.nx, ny = q.ny, nz = q.nz, M = nx / 2;
  const int budget = want / csize - 160 * 16 / csize;  // room for the static reduction array
  const int budget_f = want_f / csize - 160 * 16 / csize;
  const int twx = (M / 2 + 1) + (nx / 2 + 1);
  q.wx = pencils(M, ny * nz, budget_f, twx);
  q.wxi = pencils(M, ny * nz, budget, twx);
  q.wy = pencils(ny, q.nxh, budget, 2 * (ny / 2 + 1));
  q.wz = pencils(nz, ny, budget, 2 * (nz / 2 + 1));
  q.lwx = log2i(q.wx);
  q.lwxi = log2i(q.wxi);
  q.lwy = log2i(q.wy);
  q.lwz = log2i(q.wz);
  q.lxz = std::max(std::min({lxz, q.lwx, q.lz}), q.lwx - q.ly);
  lds[0] = (size_t)csize * (q.wx * (M + 1) + twx);
  lds[1] = (size_t)csize * (q.wy * (ny + 1) + 2 * (ny / 2 + 1));
  lds[2] = (size_t)csize * (q.wz * (nz + 1) + 2 * (nz / 2 + 1));
  lds[3] = (size_t)csize * (q.wxi * (M + 1) + twx);
  *nbx = (int)(((long)ny * nz + q.wx - 1) / q.wx);
  *nbxi = (int)(((long)ny * nz + q.wxi - 1) / q.wxi);
}

// LDS budget per workgroup: 32 KB by default (several workgroups per CU);
// FR_FFT_LDS_BYTES overrides it for tuning, capped by the device limit.
static int lds_budget() {
  int dev = 0, maxsh = 65536;
  if (hipGetDevice(&dev) == hipSuccess)
    if (hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      maxsh = 65536;
  if (maxsh > 163840) maxsh = 163840;
  if (maxsh < 65536) maxsh = 65536;
  int want = 32768;
  if (const char* e = getenv("FR_FFT_LDS_BYTES")) want = atoi(e);
  if (want < 8192) want = 8192;
  if (want > maxsh) want = maxsh;
  return want;
}

// Forward x pass: 48 KB of LDS (8 rows of 512 points), all in one plane, by
// default; FR_FFT_XF_LDS_BYTES and FR_FFT_XZ (log2 of the planes per tile)
// override them for tuning.  Measured at 512^3 (8 rows per workgroup): one
// plane 1.517 ms per launch, two 1.735, four 2.080.  In plain dispatch order
// the tiles of neighbouring planes already run on the same XCD at the same
// time and share its L2, so the refetched rows are the tiles' y neighbours,
// which a taller tile in y halves.
static int lds_budget_f() {
  int want = 49152;
  if (const char* e = getenv("FR_FFT_XF_LDS_BYTES")) want = atoi(e);
  if (want < 8192) want = 8192;
  if (want > 163840) want = 163840;
  return want;
}

// log2 of the neighbouring x-pass tiles that share an XCD (xcd_group): the
// forward pass's tiles share their y-neighbour rows, the inverse pass's
// 4-row tiles the 128-byte segments of the spectrum (4 rows read half a
// segment each).  Measured at 512^3 (serialised launches): forward 1.529 ms
// ungrouped, 1.494 in groups of 4 tiles, 1.568 of 16; inverse 0.798 ms
// ungrouped, 0.717 in groups of 8, 0.701 of 32, 0.697 of 64.  FR_FFT_XG and
// FR_FFT_XGI override the defaults (2 and 5) for tuning.
static int env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  const int v = e ? atoi(e) : dflt;
  return v < 0 ? 0 : v > 6 ? 6 : v;
}

static int x_tile_planes() {
  int l = 0;
  if (const char* e = getenv("FR_FFT_XZ")) l = atoi(e);
  return l < 0 ? 0 : l;
}

template <typename R>
static int allow_lds(size_t need) {
  if (need <= 65536) return FR_OK;
  const void* xf[5] = {(const void*)k_pass_x_fwd<CF_C, double, R>,
                       (const void*)k_pass_x_fwd<CF_FACE, double, R>,
                       (const void*)k_pass_x_fwd<CF_FACE, float, R>,
                       (const void*)k_pass_x_fwd<CF_CELL, double, R>,
                       (const void*)k_pass_x_fwd<CF_CELL, float, R>};
  for (const void* f : xf)
    FR_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
  FR_CHECK(hipFuncSetAttribute((const void*)k_pass_strided<R>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
  FR_CHECK(hipFuncSetAttribute((const void*)k_pass_x_inv<R>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
  return FR_OK;
}

// Float resources of the precision form, at its first solve.
static int ensure_f32(fr_box_poisson* s) {
  if (s->data32) return FR_OK;
  PoisDev& q = s->q32;
  q = s->q;
  plan_passes(q, (int)sizeof(float2), lds_budget(), lds_budget_f(), x_tile_planes(), s->lds32,
              &s->nbx32, &s->nbxi32);
  q.lgx = env_int("FR_FFT_XG", 2);
  q.lgxi = env_int("FR_FFT_XGI", 5);
  FR_RET(allow_lds<float>(std::max({s->lds32[0], s->lds32[1], s->lds32[2], s->lds32[3]})));
  const int n[3] = {q.nx, q.ny, q.nz};
  for (int a = 0; a < 3; a++) {
    FR_RET(upload_tw(n[a], -1, &s->twf32[a]));
    FR_RET(upload_tw(n[a], 1, &s->twi32[a]));
  }
  FR_RET(upload_tw(n[0] / 2, -1, &s->twfm32));
  FR_RET(upload_tw(n[0] / 2, 1, &s->twim32));
  if (s->nbxi32 > s->nbxi) return FR_EINVAL;  // bmax holds nbxi entries
  FR_CHECK(hipMalloc(&s->data32, sizeof(float2) * (size_t)q.nxh * n[1] * n[2]));
  return FR_OK;
}

namespace fr {

template <typename TC>
int box_poisson_solve(fr_box_poisson* s, const TC* cx, const TC* cy, const TC* cz, int cf,
                      double* P, const double* RHS, int iterations, double epsilon, int* num_iter,
                      hipStream_t st, const double* rsum) {
  if (!s || !cx || !P || !RHS || cf < CF_C || cf > CF_CELL) return FR_EINVAL;
  if (cf != CF_CELL && (!cy || !cz)) return FR_EINVAL;
  // FP32 coefficients (the precision form): FP32 inner spectral solve, with
  // the FP64 residual and P of the defect correction around it
  constexpr bool f32 = sizeof(TC) == 4;
  if (f32) FR_RET(ensure_f32(s));
  // poisson::FFT's loop (poisson.cpp:391-411): while (i < iterations &&
  // error > epsilon).  Iterations are queued in batches (the first as long
  // as the previous solve); the device gate stops the queued ones after the
  // one whose max|dp| passes the test, so the iterations run and their
  // results are those of the loop, with one host synchronisation per batch.
  FR_CHECK(hipMemsetAsync(s->gate, 0, sizeof(int) * 2, st));
  int launched = 0, ran = 0, batch = s->last_iters;
  while (launched < iterations) {
    const int k = std::min({batch, iterations - launched, (int)fr_box_poisson::QB});
    for (int j = 0; j < k; j++) {
      int rc = f32 ? box_poisson_iter<TC, float>(s, cx, cy, cz, cf, P, RHS, epsilon, j, st,
                                                 rsum)
                   : box_poisson_iter<TC, double>(s, cx, cy, cz, cf, P, RHS, epsilon, j, st,
                                                  rsum);
      if (rc) return rc;
    }
    launched += k;
    FR_CHECK(hipMemcpyAsync(s->hgate, s->gate, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
    FR_CHECK(hipStreamSynchronize(st));
    const int ran_now = s->hgate[1] - ran;
    ran = s->hgate[1];
    for (int j = 0; j < ran_now; j++) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, s->tx0[j], s->tx1[j]) == hipSuccess) {
        s->tx_ms += ms;
        s->tx_n++;
      }
    }
    if (s->hgate[0]) break;
    batch = 1;
  }
  s->last_iters = ran > 0 ? ran : 1;
  if (num_iter) *num_iter = ran;
  return FR_OK;
}
template int box_poisson_solve<double>(fr_box_poisson*, const double*, const double*,
                                       const double*, int, double*, const double*, int, double,
                                       int*, hipStream_t, const double*);
template int box_poisson_solve<float>(fr_box_poisson*, const float*, const float*, const float*,
                                      int, double*, const double*, int, double, int*,
                                      hipStream_t, const double*);

void box_poisson_xtimer(fr_box_poisson* s, double* ms, long* n, bool reset) {
  if (!s) return;
  if (ms) *ms = s->tx_ms;
  if (n) *n = s->tx_n;
  if (reset) {
    s->tx_ms = 0;
    s->tx_n = 0;
  }
}

}  // namespace fr

extern "C" {

int fr_box_poisson_create(const fr_geom* g, fr_box_poisson** out) {
  if (!geom_ok(g) || !out || !g->is_box) return FR_EINVAL;
  const int n[3] = {g->Nx - 2, g->Ny - 2, g->Nz - 2};
  int lg[3];
  for (int a = 0; a < 3; a++) {
    lg[a] = ilog2(n[a]);
    if (lg[a] < 1 || n[a] > 1024) return FR_EINVAL;
  }
  if (n[0] < 4) return FR_EINVAL;
  fr_box_poisson* s = new fr_box_poisson();
  s->g = *g;
  PoisDev& q = s->q;
  q.d = make_dev(g);
  q.nx = n[0];
  q.ny = n[1];
  q.nz = n[2];
  q.nxh = n[0] / 2 + 1;
  q.lxm = lg[0] - 1;
  q.ly = lg[1];
  q.lz = lg[2];
  q.tot = (double)((long)n[0] * n[1] * n[2]);
  q.itot = 1.0 / q.tot;
  const double Ls[3] = {g->Lx, g->Ly, g->Lz};
  const double ds[3] = {g->dx, g->dy, g->dz};
  FR_CHECK(hipMalloc(&s->data, sizeof(double2) * (size_t)q.nxh * n[1] * n[2]));
  for (int a = 0; a < 3; a++) {
    int rc = upload_tw(n[a], -1, &s->twf[a]);
    if (!rc) rc = upload_tw(n[a], 1, &s->twi[a]);
    if (rc) return rc;
    // modified wavenumber term per axis (poisson.cpp:415-447):
    // (2/d^2)*(1-cos(kk*d)), kk = ((i<n/2)?i:i-n)*2pi/L
    double TWO_PI(2 * 3.141592653589793);
    double two_pi_over_L = TWO_PI / Ls[a];
    double d2 = 2. / (ds[a] * ds[a]);
    std::vector<double> lam(n[a]);
    for (int i = 0; i < n[a]; i++) {
      double kk = ((i < n[a] / 2) ? i : i - n[a]) * two_pi_over_L;
      lam[i] = (d2) * (1 - cos(kk * ds[a]));
    }
    FR_CHECK(hipMalloc(&s->lam[a], sizeof(double) * n[a]));
    FR_CHECK(hipMemcpy(s->lam[a], lam.data(), sizeof(double) * n[a], hipMemcpyHostToDevice));
  }
  {
    int rc = upload_tw(n[0] / 2, -1, &s->twfm);
    if (!rc) rc = upload_tw(n[0] / 2, 1, &s->twim);
    if (rc) return rc;
  }
  plan_passes(q, (int)sizeof(double2), lds_budget(), lds_budget_f(), x_tile_planes(), s->lds,
              &s->nbx, &s->nbxi);
  q.lgx = env_int("FR_FFT_XG", 2);
  q.lgxi = env_int("FR_FFT_XGI", 5);
  {
    const int rc = allow_lds<double>(std::max({s->lds[0], s->lds[1], s->lds[2], s->lds[3]}));
    if (rc) return rc;
  }
  FR_CHECK(hipMalloc(&s->bmax, sizeof(double) * (s->nbxi + 1)));
  FR_CHECK(hipMalloc(&s->dres, sizeof(double) * 2));
  FR_CHECK(hipHostMalloc(&s->hres, sizeof(double) * 2, hipHostMallocDefault));
  for (int j = 0; j < fr_box_poisson::QB; j++) {
    FR_CHECK(hipEventCreate(&s->tx0[j]));
    FR_CHECK(hipEventCreate(&s->tx1[j]));
  }
  FR_CHECK(hipMalloc(&s->gate, sizeof(int) * 2));
  FR_CHECK(hipHostMalloc(&s->hgate, sizeof(int) * 2, hipHostMallocDefault));
  *out = s;
  return FR_OK;
}

int fr_box_poisson_destroy(fr_box_poisson* s) {
  if (!s) return FR_OK;
  (void)hipFree(s->data);
  (void)hipFree(s->data32);
  for (int a = 0; a < 3; a++) {
    (void)hipFree(s->twf[a]);
    (void)hipFree(s->twi[a]);
    (void)hipFree(s->lam[a]);
    (void)hipFree(s->twf32[a]);
    (void)hipFree(s->twi32[a]);
  }
  (void)hipFree(s->twfm32);
  (void)hipFree(s->twim32);
  (void)hipFree(s->twfm);
  (void)hipFree(s->twim);
  (void)hipFree(s->bmax);
  (void)hipFree(s->dres);
  (void)hipHostFree(s->hres);
  for (int j = 0; j < fr_box_poisson::QB; j++) {
    if (s->tx0[j]) (void)hipEventDestroy(s->tx0[j]);
    if (s->tx1[j]) (void)hipEventDestroy(s->tx1[j]);
  }
  (void)hipFree(s->gate);
  (void)hipHostFree(s->hgate);
  delete s;
  return FR_OK;
}

int fr_box_poisson_solve(fr_box_poisson* s, const double* cx, const double* cy,
                         const double* cz, double* P, const double* RHS, int iterations,
                         double epsilon, int* num_iter, void* stream) {
  return fr::box_poisson_solve(s, cx, cy, cz, false, P, RHS, iterations, epsilon, num_iter,
                               S(stream), nullptr);
}

}  // extern "C"
