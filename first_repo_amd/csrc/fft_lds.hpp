// LDS radix-2 FFT building blocks shared by the box and duct pressure
// solves (gfx950).  A workgroup of FT threads holds W pencils of length L in
// LDS (stride L+1 complex) and transforms them in place.  C2 is double2 (the
// FP64 solves) or float2 (the inner transform of the FP32 precision form).
#pragma once
#include <cmath>
#include <vector>

#include "fr_internal.hpp"

namespace fr {

constexpr int FT = 256;  // threads per FFT workgroup

// Pencils per workgroup for length L within `budget` complex elements after
// `extra` complex of twiddles: W*(L+1) + extra <= budget, W the largest power
// of two that fits and does not exceed cap (at least 1).  Index arithmetic
// in the passes uses shifts by log2(W).
inline int pencils(int L, int cap, int budget, int extra) {
  int w = (budget - extra) / (L + 1);
  if (w > cap) w = cap;
  int p2 = 1;
  while (p2 * 2 <= w) p2 *= 2;
  return p2;
}

inline int log2i(int w) {
  int l = 0;
  while ((1 << (l + 1)) <= w) l++;
  return l;
}

__device__ __forceinline__ int bitrev(int x, int logn) {
  return logn == 0 ? 0 : (int)(__brev((unsigned)x) >> (32 - logn));
}

template <typename C2>
__device__ __forceinline__ void bfly(C2& u, C2& t, const C2 w) {
  const auto vr = t.x * w.x - t.y * w.y;
  const auto vi = t.x * w.y + t.y * w.x;
  const C2 a = u;
  u.x = a.x + vr;
  u.y = a.y + vi;
  t.x = a.x - vr;
  t.y = a.y - vi;
}

// In-place radix-2 DIT over W pencils (stride ps complex) already in
// bit-reversed order; W and L = 2^lg are powers of two.
// tw[k] = (cos(2 pi k/L), sign*sin(2 pi k/L)), k <= L/2.
// Stages are applied in pairs from registers (each thread owns the four
// elements two consecutive stages combine), with exactly the butterflies and
// twiddles of the stage-by-stage algorithm, so the result is bit-identical.
// Stage s (half = 2^s) uses twiddle index pos * 2^(lg-1-s).
template <typename C2>
__device__ inline void fft_lds(C2* buf, int L, int lg, int W, int ps, const C2* tw) {
  const int lw = __ffs(W) - 1;
  int s = 0;
  for (; s + 1 < lg; s += 2) {
    __syncthreads();
    const int half = 1 << s, lq = lg - 2, sa = lg - 1 - s, sb = lg - 2 - s;
    for (int b = threadIdx.x; b < (W << lq); b += FT) {
      // first pair of stages (groups of 4 adjacent elements): pencil index
      // fastest across lanes, so a wave's 64-byte groups from W pencils
      // (stride L+1) fall on distinct LDS banks; later stages: position
      // fastest (consecutive elements)
      const int p = s == 0 ? (b & (W - 1)) : (b >> lq);
      const int bb = s == 0 ? (b >> lw) : (b & ((1 << lq) - 1));
      const int pos = bb & (half - 1);
      const int i0 = p * ps + ((bb >> s) << (s + 2)) + pos;
      C2 x0 = buf[i0], x1 = buf[i0 + half], x2 = buf[i0 + 2 * half], x3 = buf[i0 + 3 * half];
      const C2 wa = tw[pos << sa];
      bfly(x0, x1, wa);
      bfly(x2, x3, wa);
      bfly(x0, x2, tw[pos << sb]);
      bfly(x1, x3, tw[(pos + half) << sb]);
      buf[i0] = x0;
      buf[i0 + half] = x1;
      buf[i0 + 2 * half] = x2;
      buf[i0 + 3 * half] = x3;
    }
  }
  if (s < lg) {  // odd number of stages: the last one alone
    __syncthreads();
    const int half = 1 << s, lh = lg - 1, sa = lg - 1 - s;
    for (int b = threadIdx.x; b < (W << lh); b += FT) {
      const int p = b >> lh;
      const int bb = b & ((1 << lh) - 1);
      const int pos = bb & (half - 1);
      const int i0 = p * ps + ((bb >> s) << (s + 1)) + pos;
      C2 u = buf[i0], t = buf[i0 + half];
      bfly(u, t, tw[pos << sa]);
      buf[i0] = u;
      buf[i0 + half] = t;
    }
  }
  __syncthreads();
}

template <typename C2>
__device__ inline void load_tw(C2* tws, const C2* tw, int L) {
  for (int t = threadIdx.x; t <= L / 2; t += FT) tws[t] = tw[t];
}

// ---- one pencil per 64-lane wave, in registers ---------------------------------
// Length L = 2^LG (LG 8..10): each lane holds E = L/64 elements.  The same
// radix-2 DIT butterflies and twiddles as fft_lds, grouped into passes of up
// to log2(E) consecutive stages: a pass's groups (2^g elements at stride
// 2^s0, closed under stages s0..s0+g-1) sit in the lanes' registers, and the
// passes exchange elements through a wave-private LDS area, so the transform
// has no workgroup barrier.  The result is bit-identical to fft_lds.
template <int LG>
struct WavePlan {
  static constexpr int L = 1 << LG, LE = LG - 6, E = 1 << LE;
  static constexpr int XS = L + (L >> 3);  // exchange entries per wave (padded)
  static constexpr int S0(int p) { return p * LE; }
  static constexpr int G(int p) { return LG - p * LE < LE ? LG - p * LE : LE; }
  static constexpr int NPASS = (LG + LE - 1) / LE;
};

// padded LDS index: one spare entry every 8 keeps the strided phases of an
// exchange off a single bank group
__device__ __forceinline__ int wave_xi(int p) { return p + (p >> 3); }

// Buffer position of register slot k of `lane` in the pass (S0, GG): slot
// k = r*2^GG + m is element m of the lane's group r; group ids run lane-major
// in the first pass (so the bit-reversed load is coalesced) and r-major
// after it (so the natural-order result is).
template <int LG, int S0, int GG>
__device__ __forceinline__ int wave_pos(int lane, int k) {
  constexpr int E = WavePlan<LG>::E, R = E >> GG;
  const int r = k >> GG, m = k & ((1 << GG) - 1);
  const int gid = S0 == 0 ? lane * R + r : r * 64 + lane;
  const int pos = gid & ((1 << S0) - 1);
  return ((gid >> S0) << (S0 + GG)) + pos + (m << S0);
}

// CJ: use the conjugates of tw (the inverse transform from the forward
// table; negating the sine is exact, so the result equals a conjugate table)
template <int LG, int S0, int GG, bool CJ, typename C2>
__device__ __forceinline__ void wave_stages(C2 (&x)[WavePlan<LG>::E], int lane, const C2* tw) {
  constexpr int E = WavePlan<LG>::E, G = 1 << GG, R = E >> GG;
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int gid = S0 == 0 ? lane * R + r : r * 64 + lane;
    const int pos = gid & ((1 << S0) - 1);
#pragma unroll
    for (int t = 0; t < GG; t++) {
      const int s = S0 + t, h = 1 << t;
#pragma unroll
      for (int m = 0; m < G; m++) {
        if (m & h) continue;
        C2 w = tw[(pos + ((m & (h - 1)) << S0)) << (LG - 1 - s)];
        if (CJ) w.y = -w.y;
        bfly(x[r * G + m], x[r * G + m + h], w);
      }
    }
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Registers of pass PA to those of pass PB through the wave's area ex.
template <int LG, int PA, int PB, typename C2>
__device__ __forceinline__ void wave_exchange(C2 (&x)[WavePlan<LG>::E], C2* ex, int lane) {
  using P = WavePlan<LG>;
#pragma unroll
  for (int k = 0; k < P::E; k++) ex[wave_xi(wave_pos<LG, P::S0(PA), P::G(PA)>(lane, k))] = x[k];
  wave_sync();
#pragma unroll
  for (int k = 0; k < P::E; k++) x[k] = ex[wave_xi(wave_pos<LG, P::S0(PB), P::G(PB)>(lane, k))];
  wave_sync();
}

template <int LG, int PP, bool CJ, typename C2>
__device__ __forceinline__ void wave_passes(C2 (&x)[WavePlan<LG>::E], C2* ex, int lane,
                                            const C2* tw) {
  using P = WavePlan<LG>;
  if constexpr (PP < P::NPASS) {
    if constexpr (PP > 0) wave_exchange<LG, PP - 1, PP>(x, ex, lane);
    wave_stages<LG, P::S0(PP), P::G(PP), CJ>(x, lane, tw);
    wave_passes<LG, PP + 1, CJ>(x, ex, lane, tw);
  }
}

// On entry x[k] is buffer element wave_pos<LG,0,LE>(lane,k) of the
// bit-reversed input (wave_in_index gives its natural index); on exit x[k]
// is output wave_out_index(lane, k), natural order.  CJ: twiddles conj(tw).
template <int LG, bool CJ = false, typename C2>
__device__ __forceinline__ void wave_fft(C2 (&x)[WavePlan<LG>::E], C2* ex, int lane,
                                         const C2* tw) {
  wave_passes<LG, 0, CJ>(x, ex, lane, tw);
}

// natural index loaded into slot k: bitrev(lane*E + k) = bitrev_LE(k)*64 + bitrev6(lane)
template <int LG>
__device__ __forceinline__ int wave_in_index(int lane, int k) {
  constexpr int LE = WavePlan<LG>::LE;
  return (int)((__brev((unsigned)k) >> (32 - LE)) << 6) | (int)(__brev((unsigned)lane) >> 26);
}

template <int LG>
__device__ __forceinline__ int wave_out_index(int lane, int k) {
  using P = WavePlan<LG>;
  return wave_pos<LG, P::S0(P::NPASS - 1), P::G(P::NPASS - 1)>(lane, k);
}

// Output registers (natural order) to the bit-reversed input registers of a
// second transform, through ex.
template <int LG, typename C2>
__device__ __forceinline__ void wave_to_bitrev(C2 (&x)[WavePlan<LG>::E], C2* ex, int lane) {
  using P = WavePlan<LG>;
#pragma unroll
  for (int k = 0; k < P::E; k++)
    ex[wave_xi(bitrev(wave_out_index<LG>(lane, k), LG))] = x[k];
  wave_sync();
#pragma unroll
  for (int k = 0; k < P::E; k++) x[k] = ex[wave_xi(lane * P::E + k)];
  wave_sync();
}

// Residual r = (div(C grad P) - RHS) * cc, cc = -1/I_F2C(C)
// (poisson.cpp:393-403, poisson_IO.cpp:90-97), at interior cell c.  With INV
// the arrays hold Rho_face and C = 1./Rho_face is formed inline, the value
// grid::Solve_Poisson passes (grid.cpp:641).
// TC is the type of the coefficient arrays (float in the FP32 precision form,
// promoted to double here); P and the right-hand side are always double.
// The residual from its operands: face coefficients C (lower, upper face
// per axis), P at the cell and its six neighbours, the right-hand side.
__device__ __forceinline__ double residual_v(const Dev& d, double Cx0, double Cx1, double Cy0,
                                             double Cy1, double Cz0, double Cz1, double pc,
                                             double pxp, double pxm, double pyp, double pym,
                                             double pzp, double pzm, double rhs) {
  const double OneThird = 1. / 3.;
  const double ccf = OneThird * ((Cx0 + Cx1) * 0.5 + (Cy0 + Cy1) * 0.5 + (Cz0 + Cz1) * 0.5);
  const double cc = -1. / ccf;
  const double D = (((pxp - pc) * d.idx) * Cx1 - ((pc - pxm) * d.idx) * Cx0) * d.idx +
                   (((pyp - pc) * d.idy) * Cy1 - ((pc - pym) * d.idy) * Cy0) * d.idy +
                   (((pzp - pc) * d.idz) * Cz1 - ((pc - pzm) * d.idz) * Cz0) * d.idz;
  return (D - rhs) * cc;
}

// sh: the right-hand side is RHS[c] + a (a = minus its interior mean, not
// yet applied to the field: the same value the field takes when it is).
template <bool INV, typename TC = double>
__device__ __forceinline__ double residual_t(const Dev& d, const double* __restrict__ P,
                                             const TC* __restrict__ cx,
                                             const TC* __restrict__ cy,
                                             const TC* __restrict__ cz,
                                             const double* __restrict__ RHS, long c,
                                             bool sh = false, double a = 0.0) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  auto ld = [](const TC* a, long q) { return INV ? 1. / (double)a[q] : (double)a[q]; };
  return residual_v(d, ld(cx, c), ld(cx, c + sx), ld(cy, c), ld(cy, c + sy), ld(cz, c),
                    ld(cz, c + sz), P[c], P[c + sx], P[c - sx], P[c + sy], P[c - sy], P[c + sz],
                    P[c - sz], sh ? RHS[c] + a : RHS[c]);
}

// The same residual from the cell-centred density rn instead of the face
// arrays: the face value at the lower face of q along e is the one the
// divergence kernel stores, I_C2F = (rn[q] + rn[q - e]) * 0.5 in double,
// rounded to TC, and C = 1./face.  Equal to residual_t<true> on faces formed
// from the same rn (ghost values of rn equal their images).
template <typename TC>
__device__ __forceinline__ double residual_cell(const Dev& d, const double* __restrict__ P,
                                                const TC* __restrict__ rn,
                                                const double* __restrict__ RHS, long c, bool sh,
                                                double a) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  auto v = [&](long q) { return (double)rn[q]; };
  auto C = [](double hi, double lo) { return 1. / (double)(TC)((hi + lo) * 0.5); };
  const double r = v(c);
  return residual_v(d, C(r, v(c - sx)), C(v(c + sx), r), C(r, v(c - sy)), C(v(c + sy), r),
                    C(r, v(c - sz)), C(v(c + sz), r), P[c], P[c + sx], P[c - sx], P[c + sy],
                    P[c - sy], P[c + sz], P[c - sz], sh ? RHS[c] + a : RHS[c]);
}

__device__ __forceinline__ double residual(const Dev& d, const double* __restrict__ P,
                                           const double* __restrict__ cx,
                                           const double* __restrict__ cy,
                                           const double* __restrict__ cz,
                                           const double* __restrict__ RHS, long c) {
  return residual_t<false>(d, P, cx, cy, cz, RHS, c);
}

// Block-wide max of one value per thread (FT threads); thread 0 stores it.
__device__ inline void block_max_store(double m, double* red, double* out) {
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = FT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace fr

namespace fr {

// Max of n per-block maxima into out[0] (one workgroup), on stream st.
// With a gate (int[2], device): nothing happens once gate[0] is set;
// otherwise gate[1] counts the iteration and gate[0] is set when the max is
// not above eps (the stopping test of the defect-correction loops), so later
// iterations already queued behind it do nothing.
int max_final(const double* b, int n, double* out, hipStream_t st, int* gate = nullptr,
              double eps = 0.0);

inline int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return (1 << l) == n ? l : -1;
}

// Twiddles exactly as the oracle forms them: a = 2.0*PI*k/n,
// (cos a, sign*sin a) for k = 0..n/2.
inline int upload_tw(int n, int sign, double2** out) {
  const double PI_ = 3.141592653589793;
  std::vector<double2> t(n / 2 + 1);
  for (int k = 0; k <= n / 2; k++) {
    double ang = 2.0 * PI_ * (double)k / (double)n;
    t[k] = make_double2(cos(ang), sign * sin(ang));
  }
  FR_CHECK(hipMalloc(out, sizeof(double2) * t.size()));
  FR_CHECK(hipMemcpy(*out, t.data(), sizeof(double2) * t.size(), hipMemcpyHostToDevice));
  return FR_OK;
}

// The same twiddles rounded to float.
inline int upload_tw(int n, int sign, float2** out) {
  const double PI_ = 3.141592653589793;
  std::vector<float2> t(n / 2 + 1);
  for (int k = 0; k <= n / 2; k++) {
    double ang = 2.0 * PI_ * (double)k / (double)n;
    t[k] = make_float2((float)cos(ang), (float)(sign * sin(ang)));
  }
  FR_CHECK(hipMalloc(out, sizeof(float2) * t.size()));
  FR_CHECK(hipMemcpy(*out, t.data(), sizeof(float2) * t.size(), hipMemcpyHostToDevice));
  return FR_OK;
}

}  // namespace fr
