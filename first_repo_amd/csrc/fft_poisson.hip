// Box pressure solve on gfx950: the defect-correction loop of poisson::FFT
// (poisson.cpp:391-411) around the constant-coefficient spectral solve
// poisson::CCP_FFT (poisson.cpp:413-462).
//
// The 3-D complex transform is done as LDS-resident batched 1-D transforms:
//   pass X : residual (computed inline) -> FFT along x          (rows)
//   pass Y : FFT along y                                          (strided)
//   pass Z : FFT along z, divide by lambda(k), inverse FFT along z
//   pass Y': inverse FFT along y
//   pass X': inverse FFT along x, dp = -Re/N, P += dp, max|dp|
// Each 1-D transform is an in-place radix-2 decimation-in-time FFT on
// bit-reversed input with host-computed twiddles cos(2 pi k/n), sign*sin(..)
// (the same algorithm the CPU oracle uses, so the two agree bit for bit).
// A workgroup holds W pencils of length L in LDS, padded by one complex per
// pencil to spread the LDS banks, within a 64 KB budget (pencils + twiddles).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "fr_internal.hpp"

namespace fr {

constexpr int FT = 256;        // threads per FFT workgroup
// Pencils per workgroup for length L within `budget` complex elements:
// W*(L+1) + 2*(L/2+1) <= budget, W a power of two, W <= cap.
inline int pencils(int L, int cap, int budget) {
  int w = (budget - 2 * (L / 2 + 1)) / (L + 1);
  if (w < 1) w = 1;
  int p2 = 1;
  while (p2 * 2 <= w) p2 *= 2;
  return p2 < cap ? p2 : cap;
}

__device__ __forceinline__ int bitrev(int x, int logn) { return __brev(x) >> (32 - logn); }

__device__ __forceinline__ void bfly(double2& u, double2& t, const double2 w) {
  const double vr = t.x * w.x - t.y * w.y;
  const double vi = t.x * w.y + t.y * w.x;
  const double2 a = u;
  u = make_double2(a.x + vr, a.y + vi);
  t = make_double2(a.x - vr, a.y - vi);
}

// In-place radix-2 DIT over W pencils (stride ps complex) already in
// bit-reversed order.  tw[k] = (cos(2 pi k/L), sign*sin(2 pi k/L)), k <= L/2.
// Stages are applied in pairs from registers (each thread owns the four
// elements two consecutive stages combine), with exactly the butterflies and
// twiddles of the stage-by-stage algorithm, so the result is bit-identical.
__device__ void fft_lds(double2* buf, int L, int lg, int W, int ps, const double2* tw) {
  int half = 1, s = 0;
  const int q4 = L >> 2;
  for (; s + 1 < lg; s += 2, half <<= 2) {
    __syncthreads();
    const int stepA = L / (2 * half), stepB = L / (4 * half);
    for (int b = threadIdx.x; b < W * q4; b += FT) {
      const int p = b / q4;
      const int bb = b - p * q4;
      const int grp = bb / half;
      const int pos = bb - grp * half;
      const int i0 = p * ps + grp * 4 * half + pos;
      double2 x0 = buf[i0], x1 = buf[i0 + half], x2 = buf[i0 + 2 * half], x3 = buf[i0 + 3 * half];
      const double2 wa = tw[pos * stepA];
      bfly(x0, x1, wa);
      bfly(x2, x3, wa);
      bfly(x0, x2, tw[pos * stepB]);
      bfly(x1, x3, tw[(pos + half) * stepB]);
      buf[i0] = x0;
      buf[i0 + half] = x1;
      buf[i0 + 2 * half] = x2;
      buf[i0 + 3 * half] = x3;
    }
  }
  if (s < lg) {  // odd number of stages: the last one alone
    __syncthreads();
    const int step = L / (2 * half), h2 = L >> 1;
    for (int b = threadIdx.x; b < W * h2; b += FT) {
      const int p = b / h2;
      const int bb = b - p * h2;
      const int grp = bb / half;
      const int pos = bb - grp * half;
      const int i0 = p * ps + grp * 2 * half + pos;
      double2 u = buf[i0], t = buf[i0 + half];
      bfly(u, t, tw[pos * step]);
      buf[i0] = u;
      buf[i0 + half] = t;
    }
  }
  __syncthreads();
}

__device__ void load_tw(double2* tws, const double2* tw, int L) {
  for (int t = threadIdx.x; t <= L / 2; t += FT) tws[t] = tw[t];
}

struct PoisDev {
  Dev d;
  int nx, ny, nz, lx, ly, lz;  // interior sizes and log2
  int wx, wy, wz;              // pencils per workgroup for each pass
  double tot;                  // nx*ny*nz
};

// Pass X: per row (j,k) the residual r = (div(C grad P) - RHS) * cc with
// cc = -1/I_F2C(C) (poisson.cpp:393-403), forward FFT in x.
__global__ void __launch_bounds__(FT) k_pass_x_fwd(PoisDev q, const double* __restrict__ P,
                                                   const double* __restrict__ cx,
                                                   const double* __restrict__ cy,
                                                   const double* __restrict__ cz,
                                                   const double* __restrict__ RHS,
                                                   const double2* __restrict__ tw,
                                                   double2* __restrict__ out) {
  extern __shared__ double2 sm[];
  const Dev& d = q.d;
  const int L = q.nx, W = q.wx, ps = L + 1;
  double2* buf = sm;
  double2* tws = sm + W * ps;
  load_tw(tws, tw, L);
  const long rows = (long)q.ny * q.nz;
  const long row0 = (long)blockIdx.x * W;
  const long sx = 1, sy = d.Nx, sz = d.NxNy;
  for (int t = threadIdx.x; t < W * L; t += FT) {
    const int p = t / L, e = t - p * L;
    const long row = row0 + p;
    double r = 0.0;
    if (row < rows) {
      const int jj = (int)(row % q.ny), kk = (int)(row / q.ny);
      const long c = at(d, e + 1, jj + 1, kk + 1);
      const double Cx0 = cx[c], Cx1 = cx[c + sx];
      const double Cy0 = cy[c], Cy1 = cy[c + sy];
      const double Cz0 = cz[c], Cz1 = cz[c + sz];
      const double OneThird = 1. / 3.;
      const double ccf = OneThird * ((Cx0 + Cx1) * 0.5 + (Cy0 + Cy1) * 0.5 + (Cz0 + Cz1) * 0.5);
      const double cc = -1. / ccf;
      const double pc = P[c];
      const double D = (((P[c + sx] - pc) * d.idx) * Cx1 - ((pc - P[c - sx]) * d.idx) * Cx0) * d.idx +
                       (((P[c + sy] - pc) * d.idy) * Cy1 - ((pc - P[c - sy]) * d.idy) * Cy0) * d.idy +
                       (((P[c + sz] - pc) * d.idz) * Cz1 - ((pc - P[c - sz]) * d.idz) * Cz0) * d.idz;
      r = (D - RHS[c]) * cc;
    }
    buf[p * ps + bitrev(e, q.lx)] = make_double2(r, 0.0);
  }
  fft_lds(buf, L, q.lx, W, ps, tws);
  for (int t = threadIdx.x; t < W * L; t += FT) {
    const int p = t / L, e = t - p * L;
    const long row = row0 + p;
    if (row < rows) out[row * L + e] = buf[p * ps + e];
  }
}

// Strided pass along y (axis 1) or z (axis 2) over a tile of W consecutive x
// indices.  For axis 2 with `divide` set: forward FFT, divide by lambda except
// the zero mode (poisson.cpp:438-451), then inverse FFT (tw2) in LDS.
__global__ void __launch_bounds__(FT) k_pass_strided(PoisDev q, int axis, int divide,
                                                     const double2* __restrict__ tw,
                                                     const double2* __restrict__ tw2,
                                                     const double* __restrict__ lamx,
                                                     const double* __restrict__ lamy,
                                                     const double* __restrict__ lamz,
                                                     double2* __restrict__ data) {
  extern __shared__ double2 sm[];
  const int L = axis == 1 ? q.ny : q.nz;
  const int lg = axis == 1 ? q.ly : q.lz;
  const int W = axis == 1 ? q.wy : q.wz;
  const int ps = L + 1;
  double2* buf = sm;
  double2* tws = sm + W * ps;
  double2* tws2 = tws + (L / 2 + 1);
  load_tw(tws, tw, L);
  if (divide) load_tw(tws2, tw2, L);
  const int tiles_x = (q.nx + W - 1) / W;
  const int i0 = (blockIdx.x % tiles_x) * W;
  const int other = blockIdx.x / tiles_x;  // k for axis 1, j for axis 2
  const long stride = axis == 1 ? (long)q.nx : (long)q.nx * q.ny;
  const long base = axis == 1 ? (long)other * q.nx * q.ny : (long)other * q.nx;
  for (int t = threadIdx.x; t < W * L; t += FT) {
    const int e = t / W, p = t - e * W;
    if (i0 + p < q.nx) buf[p * ps + bitrev(e, lg)] = data[base + e * stride + i0 + p];
  }
  fft_lds(buf, L, lg, W, ps, tws);
  if (divide) {
    // lambda = (ax + ay) + az, host tables from the reference formula
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int e = t / W, p = t - e * W;
      const int ix = i0 + p, jy = other, kz = e;
      if (ix < q.nx && ix + jy + kz > 0) {
        const double lam = lamx[ix] + lamy[jy] + lamz[kz];
        double2 v = buf[p * ps + e];
        v.y /= lam;
        v.x /= lam;
        buf[p * ps + e] = v;
      }
    }
    __syncthreads();
    // bit-reverse permutation in place for the inverse transform
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int p = t / L, e = t - p * L;
      const int r = bitrev(e, lg);
      if (e < r) {
        double2 a = buf[p * ps + e];
        buf[p * ps + e] = buf[p * ps + r];
        buf[p * ps + r] = a;
      }
    }
    fft_lds(buf, L, lg, W, ps, tws2);
  }
  for (int t = threadIdx.x; t < W * L; t += FT) {
    const int e = t / W, p = t - e * W;
    if (i0 + p < q.nx) data[base + e * stride + i0 + p] = buf[p * ps + e];
  }
}

// Pass X': inverse FFT along x; dp = (-1)*Re/N (poisson.cpp:455-461);
// P += dp at the interior; per-block max|dp| (tensor0::max_abs).
__global__ void __launch_bounds__(FT) k_pass_x_inv(PoisDev q, const double2* __restrict__ in,
                                                   const double2* __restrict__ tw,
                                                   double* __restrict__ P,
                                                   double* __restrict__ bmax) {
  extern __shared__ double2 sm[];
  __shared__ double red[FT];
  const Dev& d = q.d;
  const int L = q.nx, W = q.wx, ps = L + 1;
  double2* buf = sm;
  double2* tws = sm + W * ps;
  load_tw(tws, tw, L);
  const long rows = (long)q.ny * q.nz;
  const long row0 = (long)blockIdx.x * W;
  for (int t = threadIdx.x; t < W * L; t += FT) {
    const int p = t / L, e = t - p * L;
    const long row = row0 + p;
    buf[p * ps + bitrev(e, q.lx)] = row < rows ? in[row * L + e] : make_double2(0.0, 0.0);
  }
  fft_lds(buf, L, q.lx, W, ps, tws);
  double m = 0.0;
  for (int t = threadIdx.x; t < W * L; t += FT) {
    const int p = t / L, e = t - p * L;
    const long row = row0 + p;
    if (row < rows) {
      const int jj = (int)(row % q.ny), kk = (int)(row / q.ny);
      const double dp = (-1) * buf[p * ps + e].x / q.tot;
      const long c = at(d, e + 1, jj + 1, kk + 1);
      P[c] = P[c] + dp;
      const double a = dp >= 0 ? dp : -dp;
      m = a > m ? a : m;
    }
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = FT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

__global__ void k_max_final(const double* __restrict__ b, int n, double* __restrict__ out) {
  __shared__ double red[FT];
  double m = 0.0;
  for (int t = threadIdx.x; t < n; t += FT) m = fmax(m, b[t]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = FT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace fr

using namespace fr;

struct fr_box_poisson {
  fr_geom g;
  PoisDev q;
  double2* data = nullptr;  // nx*ny*nz complex
  double2 *twf[3] = {}, *twi[3] = {};
  double *lam[3] = {};
  double* bmax = nullptr;
  double* dres = nullptr;
  double* hres = nullptr;
  int nbx = 0;
  int lds_max = 0;
};

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return (1 << l) == n ? l : -1;
}

extern "C" {

int fr_box_poisson_create(const fr_geom* g, fr_box_poisson** out) {
  if (!geom_ok(g) || !out || !g->is_box) return FR_EINVAL;
  const int n[3] = {g->Nx - 2, g->Ny - 2, g->Nz - 2};
  int lg[3];
  for (int a = 0; a < 3; a++) {
    lg[a] = ilog2(n[a]);
    if (lg[a] < 1 || n[a] > 1024) return FR_EINVAL;
  }
  fr_box_poisson* s = new fr_box_poisson();
  s->g = *g;
  s->q.d = make_dev(g);
  s->q.nx = n[0];
  s->q.ny = n[1];
  s->q.nz = n[2];
  s->q.lx = lg[0];
  s->q.ly = lg[1];
  s->q.lz = lg[2];
  s->q.tot = (double)((long)n[0] * n[1] * n[2]);
  const double PI_ = 3.141592653589793;
  const double Ls[3] = {g->Lx, g->Ly, g->Lz};
  const double ds[3] = {g->dx, g->dy, g->dz};
  FR_CHECK(hipMalloc(&s->data, sizeof(double2) * (size_t)n[0] * n[1] * n[2]));
  for (int a = 0; a < 3; a++) {
    // twiddles exactly as the oracle: a = 2.0*PI*k/n, (cos a, sign*sin a)
    std::vector<double2> f(n[a] / 2 + 1), iv(n[a] / 2 + 1);
    for (int k = 0; k <= n[a] / 2; k++) {
      double ang = 2.0 * PI_ * (double)k / (double)n[a];
      f[k] = make_double2(cos(ang), -1 * sin(ang));
      iv[k] = make_double2(cos(ang), 1 * sin(ang));
    }
    FR_CHECK(hipMalloc(&s->twf[a], sizeof(double2) * f.size()));
    FR_CHECK(hipMalloc(&s->twi[a], sizeof(double2) * iv.size()));
    FR_CHECK(hipMemcpy(s->twf[a], f.data(), sizeof(double2) * f.size(), hipMemcpyHostToDevice));
    FR_CHECK(hipMemcpy(s->twi[a], iv.data(), sizeof(double2) * iv.size(), hipMemcpyHostToDevice));
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
  // LDS budget: gfx950 has 160 KB per CU; use what the device grants a
  // workgroup, at most 160 KB.
  int dev = 0, maxsh = 65536;
  if (hipGetDevice(&dev) == hipSuccess)
    if (hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      maxsh = 65536;
  if (maxsh > 163840) maxsh = 163840;
  if (maxsh < 65536) maxsh = 65536;
  // Default budget 32 KB per workgroup (several workgroups per CU);
  // FR_FFT_LDS_BYTES overrides it for tuning, capped by the device limit.
  int want = 32768;
  if (const char* e = getenv("FR_FFT_LDS_BYTES")) want = atoi(e);
  if (want < 8192) want = 8192;
  if (want > maxsh) want = maxsh;
  const int budget = want / (int)sizeof(double2) - 160;  // room for the static reduction array
  s->q.wx = pencils(n[0], n[1] * n[2], budget);
  s->q.wy = pencils(n[1], n[0], budget);
  s->q.wz = pencils(n[2], n[0], budget);
  const size_t need = sizeof(double2) * (size_t)std::max(
      {s->q.wx * (n[0] + 1) + n[0] / 2 + 1, s->q.wy * (n[1] + 1) + 2 * (n[1] / 2 + 1),
       s->q.wz * (n[2] + 1) + 2 * (n[2] / 2 + 1)});
  if (need > 65536) {
    FR_CHECK(hipFuncSetAttribute((const void*)k_pass_x_fwd,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    FR_CHECK(hipFuncSetAttribute((const void*)k_pass_strided,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
    FR_CHECK(hipFuncSetAttribute((const void*)k_pass_x_inv,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
  }
  s->lds_max = (int)need;
  const int W = s->q.wx;
  s->nbx = (int)(((long)n[1] * n[2] + W - 1) / W);
  FR_CHECK(hipMalloc(&s->bmax, sizeof(double) * (s->nbx + 1)));
  FR_CHECK(hipMalloc(&s->dres, sizeof(double) * 2));
  FR_CHECK(hipHostMalloc(&s->hres, sizeof(double) * 2, hipHostMallocDefault));
  *out = s;
  return FR_OK;
}

int fr_box_poisson_destroy(fr_box_poisson* s) {
  if (!s) return FR_OK;
  hipFree(s->data);
  for (int a = 0; a < 3; a++) {
    hipFree(s->twf[a]);
    hipFree(s->twi[a]);
    hipFree(s->lam[a]);
  }
  hipFree(s->bmax);
  hipFree(s->dres);
  hipHostFree(s->hres);
  delete s;
  return FR_OK;
}

// One defect-correction iteration; leaves max|dp| in s->dres[0].
static int box_poisson_iter(fr_box_poisson* s, const double* cx, const double* cy,
                            const double* cz, double* P, const double* RHS,
                            hipStream_t st) {
  PoisDev& q = s->q;
  const size_t lds_x = sizeof(double2) * (q.wx * (q.nx + 1) + q.nx / 2 + 1);
  hipLaunchKernelGGL(k_pass_x_fwd, dim3(s->nbx), dim3(FT), lds_x, st, q, P, cx, cy, cz,
                     RHS, s->twf[0], s->data);
  {
    const int L = q.ny, W = q.wy;
    const size_t lds = sizeof(double2) * (W * (L + 1) + L / 2 + 1);
    const int tiles = (q.nx + W - 1) / W;
    hipLaunchKernelGGL(k_pass_strided, dim3(tiles * q.nz), dim3(FT), lds, st, q, 1, 0,
                       s->twf[1], s->twf[1], s->lam[0], s->lam[1], s->lam[2], s->data);
  }
  {
    const int L = q.nz, W = q.wz;
    const size_t lds = sizeof(double2) * (W * (L + 1) + 2 * (L / 2 + 1));
    const int tiles = (q.nx + W - 1) / W;
    hipLaunchKernelGGL(k_pass_strided, dim3(tiles * q.ny), dim3(FT), lds, st, q, 2, 1,
                       s->twf[2], s->twi[2], s->lam[0], s->lam[1], s->lam[2], s->data);
  }
  {
    const int L = q.ny, W = q.wy;
    const size_t lds = sizeof(double2) * (W * (L + 1) + L / 2 + 1);
    const int tiles = (q.nx + W - 1) / W;
    hipLaunchKernelGGL(k_pass_strided, dim3(tiles * q.nz), dim3(FT), lds, st, q, 1, 0,
                       s->twi[1], s->twi[1], s->lam[0], s->lam[1], s->lam[2], s->data);
  }
  hipLaunchKernelGGL(k_pass_x_inv, dim3(s->nbx), dim3(FT), lds_x, st, q, s->data, s->twi[0],
                     P, s->bmax);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(FT), 0, st, s->bmax, s->nbx, s->dres);
  FR_LAUNCH_CHECK();
  return ghost_fill(q.d, P, st);
}

int fr_box_poisson_solve(fr_box_poisson* s, const double* cx, const double* cy,
                         const double* cz, double* P, const double* RHS, int iterations,
                         double epsilon, int* num_iter, void* stream) {
  if (!s || !cx || !cy || !cz || !P || !RHS) return FR_EINVAL;
  hipStream_t st = S(stream);
  int i = 0;
  double error = 100;
  while ((i < iterations) && (error > epsilon)) {
    int rc = box_poisson_iter(s, cx, cy, cz, P, RHS, st);
    if (rc) return rc;
    FR_CHECK(hipMemcpyAsync(s->hres, s->dres, sizeof(double), hipMemcpyDeviceToHost, st));
    FR_CHECK(hipStreamSynchronize(st));
    error = s->hres[0];
    i++;
  }
  if (num_iter) *num_iter = i;
  return FR_OK;
}

}  // extern "C"
