// Duct pressure solve on gfx950: the defect-correction loop of
// poisson_IO::Solve (poisson_IO.cpp:81-114) around the constant-coefficient
// solve poisson_IO::CCP_FFT (poisson_IO.cpp:116-206): per x-slice a 2-D FFT
// over the periodic (y, z) directions, then for every (y, z) mode a Thomas
// solve along x with Neumann ends (poisson_IO.cpp:208-452), then the inverse
// 2-D FFT and P += Re/(Ny*Nz).
//
// The slice data and the Thomas coefficients are real, so conjugate modes stay
// conjugate and only the Ny/2+1 non-negative y frequencies are solved:
//   pass Y : residual (inline) -> real FFT along y    A[k][jh][i]  (i fastest)
//   pass Z : FFT along z                               T[i][jh][k]  (k fastest)
//   Thomas : one thread per mode (jh,k), serial in x, coalesced across modes
//   pass Z': inverse FFT along z                       A[k][jh][i]
//   pass Y': inverse real FFT along y, dp = x/(Ny*Nz), P += dp, max|dp|
// then the y,z periodic halo and the Neumann x-ghost copies of P.
// The CPU oracle runs the same transforms and the same Thomas arithmetic, so
// the two agree bit for bit.  The x pencils of a workgroup are consecutive i,
// so every global access of the y and z passes is an i-contiguous run.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "fft_lds.hpp"

namespace fr {

struct DuctDev {
  Dev d;
  int nxi, ny, nz;  // interior sizes
  int M, nh;        // ny/2, ny/2+1
  int lym, lz;      // log2(ny/2), log2(nz)
  int wy, wz;       // i-pencils per workgroup of the y and z passes (powers of two)
  int lwy, lwz;     // their log2
  long nm;          // modes nh*nz
  double tot;       // ny*nz
  double itot;      // 1/tot, exact (tot is a power of two): x*itot == x/tot
  double a;         // 1/dx^2 (a = c of poisson_IO.cpp:225-229)
};

// Pass Y: block (i-tile, k).  Residual rows r(i, j, k), j = 0..ny-1, packed as
// z_m = r_2m + i r_2m+1, an ny/2-point FFT, unpacked to X_jh, jh = 0..ny/2.
template <bool INV>
__global__ void __launch_bounds__(FT) k_duct_y_fwd(DuctDev q, const double* __restrict__ P,
                                                   const double* __restrict__ cx,
                                                   const double* __restrict__ cy,
                                                   const double* __restrict__ cz,
                                                   const double* __restrict__ RHS,
                                                   const double2* __restrict__ twM,
                                                   const double2* __restrict__ twN,
                                                   double2* __restrict__ A,
                                                   const int* __restrict__ gate) {
  extern __shared__ double2 sm[];
  if (*gate) return;  // the solve has converged: a queued iteration does nothing
  const Dev& d = q.d;
  const int M = q.M, W = q.wy, ps = M + 1;
  double2* buf = sm;
  double2* twm = sm + W * ps;
  double2* twn = twm + (M / 2 + 1);
  load_tw(twm, twM, M);
  load_tw(twn, twN, q.ny);
  const int lw = q.lwy;
  const int tiles = (q.nxi + W - 1) >> lw;
  const int k = blockIdx.x / tiles;
  const int i0 = ((int)blockIdx.x - k * tiles) << lw;
  for (int t = threadIdx.x; t < W * M; t += FT) {
    const int p = t & (W - 1), m = t >> lw;
    double r0 = 0.0, r1 = 0.0;
    if (i0 + p < q.nxi) {
      const long c = at(d, i0 + p + 1, 2 * m + 1, k + 1);
      r0 = residual_t<INV>(d, P, cx, cy, cz, RHS, c);
      r1 = residual_t<INV>(d, P, cx, cy, cz, RHS, c + d.Nx);
    }
    buf[p * ps + bitrev(m, q.lym)] = make_double2(r0, r1);
  }
  fft_lds(buf, M, q.lym, W, ps, twm);
  for (int t = threadIdx.x; t < W * (M + 1); t += FT) {
    const int p = t & (W - 1), jh = t >> lw;
    if (i0 + p >= q.nxi) continue;
    const double2 Av = buf[p * ps + (jh == M ? 0 : jh)];
    const double2 Bv = buf[p * ps + (jh == 0 ? 0 : M - jh)];
    const double Er = 0.5 * (Av.x + Bv.x), Ei = 0.5 * (Av.y - Bv.y);
    const double Dr = Av.x - Bv.x, Di = Av.y + Bv.y;
    const double Or = 0.5 * Di, Oi = -0.5 * Dr;
    const double2 w = twn[jh];
    const double Tr = Or * w.x - Oi * w.y, Ti = Or * w.y + Oi * w.x;
    A[((long)k * q.nh + jh) * q.nxi + i0 + p] = make_double2(Er + Tr, Ei + Ti);
  }
}

// Pass Z (forward, A -> T) and Z' (inverse, T -> A): block (i-tile, jh),
// W pencils of length nz.
template <bool FWD>
__global__ void __launch_bounds__(FT) k_duct_z(DuctDev q, const double2* __restrict__ tw,
                                               const double2* __restrict__ in,
                                               double2* __restrict__ out,
                                               const int* __restrict__ gate) {
  extern __shared__ double2 sm[];
  if (*gate) return;
  const int L = q.nz, W = q.wz, ps = L + 1;
  double2* buf = sm;
  double2* tws = sm + W * ps;
  load_tw(tws, tw, L);
  const int lw = q.lwz, lg = q.lz;
  const int tiles = (q.nxi + W - 1) >> lw;
  const int jh = blockIdx.x / tiles;
  const int i0 = ((int)blockIdx.x - jh * tiles) << lw;
  if (FWD) {
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int p = t & (W - 1), k = t >> lw;
      if (i0 + p < q.nxi) buf[p * ps + bitrev(k, q.lz)] = in[((long)k * q.nh + jh) * q.nxi + i0 + p];
    }
  } else {
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int k = t & (L - 1), p = t >> lg;
      if (i0 + p < q.nxi) buf[p * ps + bitrev(k, q.lz)] = in[((long)(i0 + p) * q.nh + jh) * L + k];
    }
  }
  fft_lds(buf, L, q.lz, W, ps, tws);
  if (FWD) {
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int k = t & (L - 1), p = t >> lg;
      if (i0 + p < q.nxi) out[((long)(i0 + p) * q.nh + jh) * L + k] = buf[p * ps + k];
    }
  } else {
    for (int t = threadIdx.x; t < W * L; t += FT) {
      const int p = t & (W - 1), k = t >> lw;
      if (i0 + p < q.nxi) out[((long)k * q.nh + jh) * q.nxi + i0 + p] = buf[p * ps + k];
    }
  }
}

// Thomas coefficients of mode m = jh*nz + k (poisson_IO.cpp:215-229):
// b0 = -(ly + lz) - a on the first and last rows, b = -(ly + lz) - 2a inside.
__device__ __forceinline__ void thomas_coef(const DuctDev& q, const double* lamy,
                                            const double* lamz, long m, double& b0,
                                            double& b) {
  const int jh = (int)(m / q.nz), k = (int)(m % q.nz);
  const double s = lamy[jh] + lamz[k];
  b0 = -(s)-q.a;
  b = -(s)-2 * q.a;
}

// c'_i of every mode, once per solver: c'_0 = c/b0, c'_i = c/(b - c'_{i-1} a)
// (poisson_IO.cpp:231-240); stored [i][m] for i < nxi-1.
__global__ void k_duct_cprime(DuctDev q, const double* __restrict__ lamy,
                              const double* __restrict__ lamz, double* __restrict__ cp) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= q.nm) return;
  double b0, b;
  thomas_coef(q, lamy, lamz, m, b0, b);
  const double a = q.a, c = q.a;
  double prev = c / b0;
  cp[m] = prev;
  for (int i = 1; i < q.nxi - 1; i++) {
    prev = c / (b - prev * a);
    cp[(long)i * q.nm + m] = prev;
  }
}

// Forward sweep and back substitution of every mode (poisson_IO.cpp:231-329
// for the real part, 353-452 for the imaginary part with the same
// coefficients), in place on T[i][m].  Mode (0,0) is pinned by a zero last d'.
// The sweeps are serial in i; the values and c' of TB consecutive rows are
// loaded together ahead of the dependent arithmetic.
constexpr int TB = 32;
// One thread per (mode, real or imaginary part): the two parts are
// independent systems with identical coefficients.
__global__ void __launch_bounds__(64) k_duct_thomas(DuctDev q, const double* __restrict__ lamy,
                                                    const double* __restrict__ lamz,
                                                    const double* __restrict__ cp,
                                                    double* __restrict__ T,
                                                    const int* __restrict__ gate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * q.nm || *gate) return;
  const long m = t >> 1;
  double b0, b;
  thomas_coef(q, lamy, lamz, m, b0, b);
  const double a = q.a;
  const long nm = q.nm;
  const int n = q.nxi;
  // T as doubles: row i, mode m, part t&1 at (i*nm + m)*2 + (t&1) = i*2nm + t
  const long rs = 2 * nm;
  // Register blocks of TB rows, the next block's loads issued before the
  // current block's sequential recurrence (the loads do not depend on it).
  auto load_f = [&](int i0, double* vv, double* cc) {
#pragma unroll
    for (int u = 0; u < TB; u++) {
      const int i = i0 + u;
      vv[u] = i < n ? T[(long)i * rs + t] : 0.0;
      cc[u] = (i >= 1 && i < n) ? cp[(long)(i - 1) * nm + m] : 0.0;
    }
  };
  double dd = 0.0;
  double vv[TB], cc[TB], nv[TB], nc[TB];
  load_f(0, vv, cc);
  for (int i0 = 0; i0 < n; i0 += TB) {
    if (i0 + TB < n) load_f(i0 + TB, nv, nc);
#pragma unroll
    for (int u = 0; u < TB; u++) {
      const int i = i0 + u;
      if (i < n) {
        const double v = vv[u];
        if (i == 0) {
          dd = v / b0;
        } else if (i < n - 1) {
          const double den = b - cc[u] * a;
          dd = (v - dd * a) / den;
        } else if (m == 0) {
          dd = 0.0;
        } else {
          const double den = b0 - cc[u] * a;
          dd = (v - dd * a) / den;
        }
        T[(long)i * rs + t] = dd;
      }
    }
#pragma unroll
    for (int u = 0; u < TB; u++) {
      vv[u] = nv[u];
      cc[u] = nc[u];
    }
  }
  // back substitution x_i = d'_i - c'_i x_{i+1}, from x_{n-1} = d'_{n-1}
  auto load_b = [&](int i1, double* vv, double* cc) {
#pragma unroll
    for (int u = 0; u < TB; u++) {
      const int i = i1 - u;
      vv[u] = i >= 0 ? T[(long)i * rs + t] : 0.0;
      cc[u] = i >= 0 ? cp[(long)i * nm + m] : 0.0;
    }
  };
  double x = dd;
  load_b(n - 2, vv, cc);
  for (int i1 = n - 2; i1 >= 0; i1 -= TB) {
    if (i1 - TB >= 0) load_b(i1 - TB, nv, nc);
#pragma unroll
    for (int u = 0; u < TB; u++) {
      const int i = i1 - u;
      if (i >= 0) {
        x = vv[u] - cc[u] * x;
        T[(long)i * rs + t] = x;
      }
    }
#pragma unroll
    for (int u = 0; u < TB; u++) {
      vv[u] = nv[u];
      cc[u] = nc[u];
    }
  }
}

// Pass Y': block (i-tile, k).  Z_jh = (X_jh + conj X_{M-jh}) + i w^jh (X_jh -
// conj X_{M-jh}), an ny/2-point inverse FFT, x_2m = Re z_m, x_2m+1 = Im z_m;
// dp = x/(Ny*Nz) (poisson_IO.cpp:203); P += dp; per-block max|dp|.
__global__ void __launch_bounds__(FT) k_duct_y_inv(DuctDev q, const double2* __restrict__ A,
                                                   const double2* __restrict__ twM,
                                                   const double2* __restrict__ twN,
                                                   double* __restrict__ P,
                                                   double* __restrict__ bmax,
                                                   const int* __restrict__ gate) {
  extern __shared__ double2 sm[];
  if (*gate) return;
  __shared__ double red[FT];
  const Dev& d = q.d;
  const int M = q.M, W = q.wy, ps = M + 1;
  double2* buf = sm;
  double2* twm = sm + W * ps;
  double2* twn = twm + (M / 2 + 1);
  load_tw(twm, twM, M);
  load_tw(twn, twN, q.ny);
  __syncthreads();
  const int lw = q.lwy;
  const int tiles = (q.nxi + W - 1) >> lw;
  const int k = blockIdx.x / tiles;
  const int i0 = ((int)blockIdx.x - k * tiles) << lw;
  const long row = (long)k * q.nh;
  for (int t = threadIdx.x; t < W * M; t += FT) {
    const int p = t & (W - 1), jh = t >> lw;
    double2 z = make_double2(0.0, 0.0);
    if (i0 + p < q.nxi) {
      const double2 Av = A[(row + jh) * q.nxi + i0 + p];
      const double2 Bv = A[(row + M - jh) * q.nxi + i0 + p];
      const double Sr = Av.x + Bv.x, Si = Av.y - Bv.y;
      const double Dr = Av.x - Bv.x, Di = Av.y + Bv.y;
      const double2 w = twn[jh];
      const double Tr = Dr * w.x - Di * w.y, Ti = Dr * w.y + Di * w.x;
      z = make_double2(Sr - Ti, Si + Tr);
    }
    buf[p * ps + bitrev(jh, q.lym)] = z;
  }
  fft_lds(buf, M, q.lym, W, ps, twm);
  double mx = 0.0;
  for (int t = threadIdx.x; t < W * M; t += FT) {
    const int p = t & (W - 1), e = t >> lw;
    if (i0 + p >= q.nxi) continue;
    const double2 z = buf[p * ps + e];
    const double dp0 = z.x * q.itot;
    const double dp1 = z.y * q.itot;
    const long c = at(d, i0 + p + 1, 2 * e + 1, k + 1);
    P[c] = P[c] + dp0;
    P[c + d.Nx] = P[c + d.Nx] + dp1;
    const double a0 = dp0 >= 0 ? dp0 : -dp0, a1 = dp1 >= 0 ? dp1 : -dp1;
    mx = a0 > mx ? a0 : mx;
    mx = a1 > mx ? a1 : mx;
  }
  block_max_store(mx, red, bmax + blockIdx.x);
}

// Neumann x ghosts of P on the inlet and outlet planes (poisson_IO.cpp:102-107):
// P(0,j,k) = P(1,j,k), P(Nx-1,j,k) = P(Nx-2,j,k) for every ghosted j, k.
__global__ void k_neumann_x(Dev d, double* __restrict__ P) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)d.Ny * d.Nz) return;
  const long base = t * d.Nx;
  P[base] = P[base + 1];
  P[base + d.Nx - 1] = P[base + d.Nx - 2];
}

}  // namespace fr

using namespace fr;

struct fr_duct_poisson {
  fr_geom g;
  DuctDev q;
  double2 *A = nullptr, *T = nullptr;
  double* cp = nullptr;
  double *lamy = nullptr, *lamz = nullptr;
  double2 *twfm = nullptr, *twim = nullptr;  // length ny/2
  double2 *twfy = nullptr, *twiy = nullptr;  // length ny
  double2 *twfz = nullptr, *twiz = nullptr;  // length nz
  double* bmax = nullptr;
  double* dres = nullptr;
  double* hres = nullptr;
  int nby = 0, nbz = 0;
  size_t lds_y = 0, lds_z = 0;
  // device convergence gate (done flag, iterations run), its host copy, and
  // the previous solve's iteration count (the next batch size)
  int* gate = nullptr;
  int* hgate = nullptr;
  int last_iters = 1;
  // optional phase timers (poisson_IO's FFT_TIME / SOLVER_TIME): four events
  // per iteration around the FFT passes and the Thomas sweeps
  bool timing = false;
  std::vector<hipEvent_t> tev;
  int tev_used = 0;
  double fft_ms = 0, solver_ms = 0;
};

static int duct_poisson_iter(fr_duct_poisson* s, const double* cx, const double* cy,
                             const double* cz, bool inv, double* P, const double* RHS,
                             double eps, hipStream_t st);

namespace fr {

// poisson_IO::Solve (poisson_IO.cpp:81-114): error = max|dp| of the iteration.
int duct_poisson_solve(fr_duct_poisson* s, const double* cx, const double* cy, const double* cz,
                       bool inv, double* P, const double* RHS, int iterations, double epsilon,
                       int* num_iter, hipStream_t st) {
  if (!s || !cx || !cy || !cz || !P || !RHS) return FR_EINVAL;
  // Iterations queued in batches behind the device gate, as in the box
  // solve: the same iterations run, one host synchronisation per batch.
  FR_CHECK(hipMemsetAsync(s->gate, 0, sizeof(int) * 2, st));
  int launched = 0, ran = 0, batch = s->last_iters;
  while (launched < iterations) {
    const int k = std::min({batch, iterations - launched, 32});
    for (int j = 0; j < k; j++) {
      int rc = duct_poisson_iter(s, cx, cy, cz, inv, P, RHS, epsilon, st);
      if (rc) return rc;
    }
    launched += k;
    FR_CHECK(hipMemcpyAsync(s->hgate, s->gate, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
    FR_CHECK(hipStreamSynchronize(st));
    ran = s->hgate[1];
    if (s->hgate[0]) break;
    batch = 1;
  }
  s->last_iters = ran > 0 ? ran : 1;
  if (num_iter) *num_iter = ran;
  if (s->timing) {  // the stream is synchronised: every recorded event is done
    for (int e = 0; e + 3 < s->tev_used; e += 4) {
      float a = 0, b = 0, c = 0;
      hipEventElapsedTime(&a, s->tev[e], s->tev[e + 1]);
      hipEventElapsedTime(&b, s->tev[e + 1], s->tev[e + 2]);
      hipEventElapsedTime(&c, s->tev[e + 2], s->tev[e + 3]);
      s->fft_ms += a + c;
      s->solver_ms += b;
    }
    s->tev_used = 0;
  }
  return FR_OK;
}

// Turns the phase timers on (creating 4 events per iteration up to
// `iterations`) and returns the summed milliseconds so far.
int duct_poisson_timing(fr_duct_poisson* s, int iterations, double* fft_ms, double* solver_ms) {
  if (!s) return FR_EINVAL;
  const size_t need = 4 * (size_t)std::max(iterations, 1);
  while (s->tev.size() < need) {
    hipEvent_t e;
    FR_CHECK(hipEventCreate(&e));
    s->tev.push_back(e);
  }
  s->timing = true;
  if (fft_ms) *fft_ms = s->fft_ms;
  if (solver_ms) *solver_ms = s->solver_ms;
  return FR_OK;
}

}  // namespace fr

extern "C" {

int fr_duct_poisson_create(const fr_geom* g, fr_duct_poisson** out) {
  if (!geom_ok(g) || !out || g->is_box) return FR_EINVAL;
  const int nxi = g->Nx - 2, ny = g->Ny - 2, nz = g->Nz - 2;
  const int ly = ilog2(ny), lz = ilog2(nz);
  if (nxi < 2 || ly < 2 || lz < 1 || ny > 1024 || nz > 1024) return FR_EINVAL;
  fr_duct_poisson* s = new fr_duct_poisson();
  s->g = *g;
  DuctDev& q = s->q;
  q.d = make_dev(g);
  q.nxi = nxi;
  q.ny = ny;
  q.nz = nz;
  q.M = ny / 2;
  q.nh = ny / 2 + 1;
  q.lym = ly - 1;
  q.lz = lz;
  q.nm = (long)q.nh * nz;
  q.tot = (double)(ny * nz);
  q.itot = 1.0 / q.tot;
  q.a = 1. / (g->dx * g->dx);
  const size_t nspec = (size_t)nxi * q.nm;
  FR_CHECK(hipMalloc(&s->A, sizeof(double2) * nspec));
  FR_CHECK(hipMalloc(&s->T, sizeof(double2) * nspec));
  FR_CHECK(hipMalloc(&s->cp, sizeof(double) * (size_t)(nxi - 1) * q.nm));
  int rc = upload_tw(q.M, -1, &s->twfm);
  if (!rc) rc = upload_tw(q.M, 1, &s->twim);
  if (!rc) rc = upload_tw(ny, -1, &s->twfy);
  if (!rc) rc = upload_tw(ny, 1, &s->twiy);
  if (!rc) rc = upload_tw(nz, -1, &s->twfz);
  if (!rc) rc = upload_tw(nz, 1, &s->twiz);
  if (rc) return rc;
  // (2/d^2)(1 - cos(kk d)) per y and z wavenumber, kk = ((j<n/2)?j:j-n)*2pi/L
  // (poisson_IO.cpp:215-227)
  {
    double TWO_PI(2 * 3.141592653589793);
    double Two_PI_Over_Ly = TWO_PI / g->Ly;
    double Two_PI_Over_Lz = TWO_PI / g->Lz;
    double dy = g->dy, dz = g->dz;
    double dy2 = 2. / (dy * dy), dz2 = 2. / (dz * dz);
    std::vector<double> ly_(q.nh), lz_(nz);
    for (int j = 0; j < q.nh; j++) {
      double jj = ((j < ny / 2) ? j : j - ny) * Two_PI_Over_Ly;
      ly_[j] = (dy2) * (1 - cos(jj * dy));
    }
    for (int k = 0; k < nz; k++) {
      double kk = ((k < nz / 2) ? k : k - nz) * Two_PI_Over_Lz;
      lz_[k] = (dz2) * (1 - cos(kk * dz));
    }
    FR_CHECK(hipMalloc(&s->lamy, sizeof(double) * q.nh));
    FR_CHECK(hipMalloc(&s->lamz, sizeof(double) * nz));
    FR_CHECK(hipMemcpy(s->lamy, ly_.data(), sizeof(double) * q.nh, hipMemcpyHostToDevice));
    FR_CHECK(hipMemcpy(s->lamz, lz_.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  }
  int want = 32768;
  if (const char* e = getenv("FR_FFT_LDS_BYTES")) want = atoi(e);
  if (want < 8192) want = 8192;
  if (want > 65536) want = 65536;
  const int budget = want / (int)sizeof(double2) - 160;
  q.wy = pencils(q.M, 64, budget, (q.M / 2 + 1) + (ny / 2 + 1));
  q.wz = pencils(nz, 64, budget, nz / 2 + 1);
  q.lwy = log2i(q.wy);
  q.lwz = log2i(q.wz);
  s->lds_y = sizeof(double2) * (q.wy * (q.M + 1) + (q.M / 2 + 1) + (ny / 2 + 1));
  s->lds_z = sizeof(double2) * (q.wz * (nz + 1) + (nz / 2 + 1));
  s->nby = ((nxi + q.wy - 1) / q.wy) * nz;
  s->nbz = ((nxi + q.wz - 1) / q.wz) * q.nh;
  FR_CHECK(hipMalloc(&s->bmax, sizeof(double) * (s->nby + 1)));
  FR_CHECK(hipMalloc(&s->dres, sizeof(double) * 2));
  FR_CHECK(hipHostMalloc(&s->hres, sizeof(double) * 2, hipHostMallocDefault));
  FR_CHECK(hipMalloc(&s->gate, sizeof(int) * 2));
  FR_CHECK(hipHostMalloc(&s->hgate, sizeof(int) * 2, hipHostMallocDefault));
  hipLaunchKernelGGL(k_duct_cprime, dim3(nblocks(q.nm, 256)), dim3(256), 0, 0, q, s->lamy,
                     s->lamz, s->cp);
  FR_LAUNCH_CHECK();
  FR_CHECK(hipDeviceSynchronize());
  *out = s;
  return FR_OK;
}

int fr_duct_poisson_destroy(fr_duct_poisson* s) {
  if (!s) return FR_OK;
  for (hipEvent_t e : s->tev) (void)hipEventDestroy(e);
  (void)hipFree(s->A);
  (void)hipFree(s->T);
  (void)hipFree(s->cp);
  (void)hipFree(s->lamy);
  (void)hipFree(s->lamz);
  (void)hipFree(s->twfm);
  (void)hipFree(s->twim);
  (void)hipFree(s->twfy);
  (void)hipFree(s->twiy);
  (void)hipFree(s->twfz);
  (void)hipFree(s->twiz);
  (void)hipFree(s->bmax);
  (void)hipFree(s->dres);
  (void)hipHostFree(s->hres);
  (void)hipFree(s->gate);
  (void)hipHostFree(s->hgate);
  delete s;
  return FR_OK;
}

// One defect-correction iteration; leaves max|dp| in s->dres[0].  With inv
// the coefficient arrays hold Rho_face and C = 1./Rho_face.
static int duct_poisson_iter(fr_duct_poisson* s, const double* cx, const double* cy,
                             const double* cz, bool inv, double* P, const double* RHS,
                             double eps, hipStream_t st) {
  DuctDev& q = s->q;
  const int* g = s->gate;
  hipEvent_t* ev = nullptr;
  if (s->timing && s->tev_used + 4 <= (int)s->tev.size()) {
    ev = &s->tev[s->tev_used];
    s->tev_used += 4;
    FR_CHECK(hipEventRecord(ev[0], st));
  }
  if (inv)
    hipLaunchKernelGGL(k_duct_y_fwd<true>, dim3(s->nby), dim3(FT), s->lds_y, st, q, P, cx, cy,
                       cz, RHS, s->twfm, s->twfy, s->A, g);
  else
    hipLaunchKernelGGL(k_duct_y_fwd<false>, dim3(s->nby), dim3(FT), s->lds_y, st, q, P, cx, cy,
                       cz, RHS, s->twfm, s->twfy, s->A, g);
  hipLaunchKernelGGL(k_duct_z<true>, dim3(s->nbz), dim3(FT), s->lds_z, st, q, s->twfz, s->A,
                     s->T, g);
  if (ev) FR_CHECK(hipEventRecord(ev[1], st));
  hipLaunchKernelGGL(k_duct_thomas, dim3(nblocks(2 * q.nm, 64)), dim3(64), 0, st, q, s->lamy,
                     s->lamz, s->cp, (double*)s->T, g);
  if (ev) FR_CHECK(hipEventRecord(ev[2], st));
  hipLaunchKernelGGL(k_duct_z<false>, dim3(s->nbz), dim3(FT), s->lds_z, st, q, s->twiz, s->T,
                     s->A, g);
  hipLaunchKernelGGL(k_duct_y_inv, dim3(s->nby), dim3(FT), s->lds_y, st, q, s->A, s->twim,
                     s->twiy, P, s->bmax, g);
  if (ev) FR_CHECK(hipEventRecord(ev[3], st));
  FR_LAUNCH_CHECK();
  // the ghost copies of P below (periodic y,z; Neumann x) are idempotent
  int rc = max_final(s->bmax, s->nby, s->dres, st, s->gate, eps);
  if (!rc) rc = ghost_fill(q.d, P, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_neumann_x, dim3(nblocks((long)q.d.Ny * q.d.Nz, 256)), dim3(256), 0, st,
                     q.d, P);
  FR_LAUNCH_CHECK();
  return FR_OK;
}

int fr_duct_poisson_solve(fr_duct_poisson* s, const double* cx, const double* cy,
                          const double* cz, double* P, const double* RHS, int iterations,
                          double epsilon, int* num_iter, void* stream) {
  return fr::duct_poisson_solve(s, cx, cy, cz, false, P, RHS, iterations, epsilon, num_iter,
                                S(stream));
}

}  // extern "C"
