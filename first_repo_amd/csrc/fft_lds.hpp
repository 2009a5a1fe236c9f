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

template <bool INV, typename TC = double>
__device__ __forceinline__ double residual_t(const Dev& d, const double* __restrict__ P,
                                             const TC* __restrict__ cx,
                                             const TC* __restrict__ cy,
                                             const TC* __restrict__ cz,
                                             const double* __restrict__ RHS, long c) {
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  auto ld = [](const TC* a, long q) { return INV ? 1. / (double)a[q] : (double)a[q]; };
  return residual_v(d, ld(cx, c), ld(cx, c + sx), ld(cy, c), ld(cy, c + sy), ld(cz, c),
                    ld(cz, c + sz), P[c], P[c + sx], P[c - sx], P[c + sy], P[c - sy], P[c + sz],
                    P[c - sz], RHS[c]);
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
