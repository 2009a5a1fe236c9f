// Oracle pressure solves.  TEST INFRASTRUCTURE.
//   box : poisson.cpp:391-462 (defect correction around a 3-D FFT solve)
//   duct: poisson_IO.cpp:81-452 (same loop; 2-D FFT in y,z + Thomas in x)
#include "orc.h"

namespace orc {

void BoxPoisson::init(const Params* p, const Geom* gg) {
  par = p;
  g = gg;
  center_coeff.init(g);
  RHS_new.init(g);
  dummy.init(g);
}

// poisson.cpp:391-411
void BoxPoisson::Solve(const F1& C, F0& P, const F0& RHS) {
  center_coeff.Equal_I_F2C(C);
  center_coeff.Equal_Divide(-1., center_coeff);
  int i = 0;
  double error = 100;
  while ((i < par->Iteration) && (error > par->epsilon)) {
    dummy.Equal_Grad_C2F(P);
    dummy.mul(C);
    RHS_new.Equal_Div_F2C(dummy);
    RHS_new.sub(RHS);
    RHS_new.mul(center_coeff);
    CCP_FFT(RHS_new, dummy.x);
    P.add(dummy.x);
    P.Update_Ghosts();
    error = dummy.x.max_abs();
    i++;
  }
  num_iteration = i;
}

// poisson.cpp:413-462.  The residual is real and the divisor is real and
// even in every wavenumber, so the solve runs on the nx/2+1 non-negative x
// frequencies: real transform along x, complex along y and z, divide, complex
// inverse along z and y, real inverse along x.  This is the transform the
// reference's full complex FFT computes (its imaginary part is zero to
// round-off and is discarded at poisson.cpp:461); odd Nx falls back to the
// full complex transform.
void BoxPoisson::CCP_FFT(const F0& RHS, F0& P) {
  const int nx = g->Nx_tot, ny = g->Ny_tot, nz = g->Nz_tot, bs = g->bs;
  if (nx % 2 != 0 || nx < 4) {
    CCP_FFT_complex(RHS, P);
    return;
  }
  const int tot = g->size_tot();
  const int nh = nx / 2 + 1;
  const long sy = nh, sz = (long)nh * ny;
  double TWO_PI(2 * 3.141592653589793);
  double Two_PI_Over_Lx = TWO_PI / g->Lx;
  double Two_PI_Over_Ly = TWO_PI / g->Ly;
  double Two_PI_Over_Lz = TWO_PI / g->Lz;
  re.assign((size_t)nh * ny * nz, 0.0);
  im.assign((size_t)nh * ny * nz, 0.0);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++) {
    std::vector<double> row(nx);
    for (int j = 0; j < ny; j++) {
      for (int i = 0; i < nx; i++) row[i] = RHS(i + bs, j + bs, k + bs);
      rfft1d(row.data(), nx, &re[k * sz + j * sy], &im[k * sz + j * sy]);
    }
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nh; i++) fft1d(&re[k * sz + i], &im[k * sz + i], ny, sy, -1);
  double dx = g->dx, dy = g->dy, dz = g->dz;
  double dx2 = 2. / (dx * dx), dy2 = 2. / (dy * dy), dz2 = 2. / (dz * dz);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nh; i++) {
      double* r = &re[j * sy + i];
      double* m = &im[j * sy + i];
      fft1d(r, m, nz, sz, -1);
      for (int k = 0; k < nz; k++)
        if (i + j + k > 0) {
          double ii = ((i < nx / 2) ? i : i - nx) * Two_PI_Over_Lx;
          double jj = ((j < ny / 2) ? j : j - ny) * Two_PI_Over_Ly;
          double kk = ((k < nz / 2) ? k : k - nz) * Two_PI_Over_Lz;
          m[k * sz] /= ((dx2) * (1 - cos(ii * dx)) + (dy2) * (1 - cos(jj * dy)) +
                        (dz2) * (1 - cos(kk * dz)));
          r[k * sz] /= ((dx2) * (1 - cos(ii * dx)) + (dy2) * (1 - cos(jj * dy)) +
                        (dz2) * (1 - cos(kk * dz)));
        }
      fft1d(r, m, nz, sz, +1);
    }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int i = 0; i < nh; i++) fft1d(&re[k * sz + i], &im[k * sz + i], ny, sy, +1);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++) {
    std::vector<double> row(nx);
    for (int j = 0; j < ny; j++) {
      irfft1d(&re[k * sz + j * sy], &im[k * sz + j * sy], nx, row.data());
      for (int i = 0; i < nx; i++) P(i + bs, j + bs, k + bs) = (-1) * row[i] / (tot);
    }
  }
}

// poisson.cpp:413-462 with the full complex 3-D transform.
void BoxPoisson::CCP_FFT_complex(const F0& RHS, F0& P) {
  const int nx = g->Nx_tot, ny = g->Ny_tot, nz = g->Nz_tot, bs = g->bs;
  const int tot = g->size_tot();
  double TWO_PI(2 * 3.141592653589793);
  double Two_PI_Over_Lx = TWO_PI / g->Lx;
  double Two_PI_Over_Ly = TWO_PI / g->Ly;
  double Two_PI_Over_Lz = TWO_PI / g->Lz;
  re.assign((size_t)tot, 0.0);
  im.assign((size_t)tot, 0.0);
  size_t count = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) re[count++] = RHS(i + bs, j + bs, k + bs);
  fft3d(re, im, nx, ny, nz, -1);
  double dx = g->dx, dy = g->dy, dz = g->dz;
  double dx2 = 2. / (dx * dx), dy2 = 2. / (dy * dy), dz2 = 2. / (dz * dz);
  count = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (i + j + k > 0) {
          double ii = ((i < nx / 2) ? i : i - nx) * Two_PI_Over_Lx;
          double jj = ((j < ny / 2) ? j : j - ny) * Two_PI_Over_Ly;
          double kk = ((k < nz / 2) ? k : k - nz) * Two_PI_Over_Lz;
          im[count] /= ((dx2) * (1 - cos(ii * dx)) + (dy2) * (1 - cos(jj * dy)) +
                        (dz2) * (1 - cos(kk * dz)));
          re[count] /= ((dx2) * (1 - cos(ii * dx)) + (dy2) * (1 - cos(jj * dy)) +
                        (dz2) * (1 - cos(kk * dz)));
        }
        count++;
      }
  fft3d(re, im, nx, ny, nz, +1);
  count = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) P(i + bs, j + bs, k + bs) = (-1) * re[count++] / (tot);
}

void DuctPoisson::init(const Params* p, const Geom* gg) {
  par = p;
  g = gg;
  center_coeff.init(g);
  RHS_new.init(g);
  fft_real.init(g);
  fft_imag.init(g);
  dummy.init(g);
  thomas_c.init(g);
  thomas_d.init(g);
}

// poisson_IO.cpp:81-114.  Note the order: err is taken before P += dp, and the
// inlet/outlet x-ghosts of P are Neumann copies.
void DuctPoisson::Solve(const F1& C, F0& P, const F0& RHS) {
  center_coeff.Equal_I_F2C(C);
  center_coeff.Equal_Divide(-1., center_coeff);
  int i = 0;
  num_iteration2 = 0;
  double error = 100;
  const int Nx = g->Nx, Ny = g->Ny, Nz = g->Nz;
  while ((i < par->Iteration) && (error > par->epsilon)) {
    dummy.Equal_Grad_C2F(P);
    dummy.mul(C);
    RHS_new.Equal_Div_F2C(dummy);
    RHS_new.sub(RHS);
    RHS_new.mul(center_coeff);
    CCP_FFT(RHS_new, dummy.x);
    error = dummy.x.max_abs();
    P.add(dummy.x);
    P.Update_Ghosts();
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) P(0, j, k) = P(1, j, k);
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) P(Nx - 1, j, k) = P(Nx - 2, j, k);
    i++;
    // num_iteration2_ accumulates an uninitialised per-solve count in the
    // reference (poisson_IO.cpp:110); one direct solve per iteration here.
    num_iteration2 += 1;
  }
  num_iteration1 = i;
  num_iteration2 /= i;
}

// poisson_IO.cpp:116-206: per interior x-slice a 2-D FFT over (y,z), a
// Thomas solve along x per (y,z) mode, the inverse 2-D FFT, P = Re/(Ny*Nz).
// The slice data and the Thomas coefficients are real, so conjugate modes
// stay conjugate: the solve runs on the Ny/2+1 non-negative y frequencies
// (real transform along y, complex along z; inverse: complex along z, real
// along y), which is the reference's full complex solve with its zero
// imaginary part dropped.  Odd Ny falls back to the full complex transform.
void DuctPoisson::CCP_FFT(const F0& RHS, F0& P) {
  const int Nx = g->Nx, bs = g->bs, ny = g->Ny_tot, nz = g->Nz_tot;
  if (ny % 2 != 0 || ny < 4) {
    CCP_FFT_complex(RHS, P);
    return;
  }
  const int nxi = Nx - 2 * bs, nh = ny / 2 + 1;
  const long nmode = (long)nh * nz;
  // spectrum [i][jh][k], k fastest
  std::vector<double> sr((size_t)nxi * nmode), si((size_t)nxi * nmode);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nxi; i++) {
    std::vector<double> row(ny);
    double* r = &sr[(size_t)i * nmode];
    double* m = &si[(size_t)i * nmode];
    std::vector<double> xr(nh), xi(nh);
    for (int k = 0; k < nz; k++) {
      for (int j = 0; j < ny; j++) row[j] = RHS(i + bs, j + bs, k + bs);
      rfft1d(row.data(), ny, xr.data(), xi.data());
      for (int jh = 0; jh < nh; jh++) {
        r[(long)jh * nz + k] = xr[jh];
        m[(long)jh * nz + k] = xi[jh];
      }
    }
    for (int jh = 0; jh < nh; jh++) fft1d(r + (long)jh * nz, m + (long)jh * nz, nz, 1, -1);
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; k++)
    for (int jh = 0; jh < nh; jh++) {
      thomas_line(sr.data(), nmode, (long)jh * nz + k, jh, k);
      thomas_line(si.data(), nmode, (long)jh * nz + k, jh, k);
    }
  const int tot = ny * nz;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nxi; i++) {
    std::vector<double> row(ny), xr(nh), xi(nh);
    double* r = &sr[(size_t)i * nmode];
    double* m = &si[(size_t)i * nmode];
    for (int jh = 0; jh < nh; jh++) fft1d(r + (long)jh * nz, m + (long)jh * nz, nz, 1, +1);
    for (int k = 0; k < nz; k++) {
      for (int jh = 0; jh < nh; jh++) {
        xr[jh] = r[(long)jh * nz + k];
        xi[jh] = m[(long)jh * nz + k];
      }
      irfft1d(xr.data(), xi.data(), ny, row.data());
      for (int j = 0; j < ny; j++) P(i + bs, j + bs, k + bs) = row[j] / (tot);
    }
  }
}

// One Thomas line (poisson_IO.cpp:208-329) over the nxi interior x-points of
// mode (j, k), values v[i*stride + off]: first and last rows Neumann (b0),
// mode (0,0) pinned by a zero last d'.
void DuctPoisson::thomas_line(double* v, long stride, long off, int j, int k) const {
  const int bs = g->bs, Nyt = g->Ny_tot, Nzt = g->Nz_tot;
  const int n = g->Nx - 2 * bs;
  double TWO_PI(2 * 3.141592653589793);
  double Two_PI_Over_Ly = TWO_PI / g->Ly;
  double Two_PI_Over_Lz = TWO_PI / g->Lz;
  double jj = ((j < Nyt / 2) ? j : j - Nyt) * Two_PI_Over_Ly;
  double kk = ((k < Nzt / 2) ? k : k - Nzt) * Two_PI_Over_Lz;
  double inv_dx2 = 1. / (g->dx * g->dx);
  double dy = g->dy, dz = g->dz;
  double dy2 = 2. / (dy * dy), dz2 = 2. / (dz * dz);
  double a = inv_dx2;
  double b0 = -((dy2) * (1 - cos(jj * dy)) + (dz2) * (1 - cos(kk * dz))) - inv_dx2;
  double b = -((dy2) * (1 - cos(jj * dy)) + (dz2) * (1 - cos(kk * dz))) - 2 * inv_dx2;
  double c = inv_dx2;
  std::vector<double> cp(n), dp(n);
  auto V = [&](int i) -> double& { return v[(long)i * stride + off]; };
  cp[0] = c / b0;
  dp[0] = V(0) / b0;
  for (int i = 1; i < n - 1; i++) {
    cp[i] = c / (b - cp[i - 1] * a);
    dp[i] = (V(i) - dp[i - 1] * a) / (b - cp[i - 1] * a);
  }
  if (j + k == 0)
    dp[n - 1] = 0;
  else
    dp[n - 1] = (V(n - 1) - dp[n - 2] * a) / (b0 - cp[n - 2] * a);
  V(n - 1) = dp[n - 1];
  for (int i = n - 2; i >= 0; i--) V(i) = dp[i] - cp[i] * V(i + 1);
}

// poisson_IO.cpp:116-206 with the full complex 2-D transform (odd Ny).
void DuctPoisson::CCP_FFT_complex(const F0& RHS, F0& P) {
  const int Nx = g->Nx, bs = g->bs, ny = g->Ny_tot, nz = g->Nz_tot;
  std::vector<double> re((size_t)ny * nz), im((size_t)ny * nz);
  for (int i = bs; i < Nx - bs; i++) {
    size_t c = 0;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        im[c] = 0;
        re[c++] = RHS(i, j + bs, k + bs);
      }
    fft2d(re.data(), im.data(), ny, nz, -1);
    c = 0;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        fft_real(i, j + bs, k + bs) = re[c];
        fft_imag(i, j + bs, k + bs) = im[c];
        c++;
      }
  }
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++) {
      thomas(fft_real, thomas_c, thomas_d, 0, j, k);
      thomas(fft_imag, thomas_c, thomas_d, 1, j, k);
    }
  int tot = ny * nz;
  for (int i = bs; i < Nx - bs; i++) {
    size_t c = 0;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        im[c] = fft_imag(i, j + bs, k + bs);
        re[c] = fft_real(i, j + bs, k + bs);
        c++;
      }
    fft2d(re.data(), im.data(), ny, nz, +1);
    c = 0;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) P(i, j + bs, k + bs) = re[c++] / (tot);
  }
}

// poisson_IO.cpp:208-329 (real part; the imaginary part at 331-452 is the
// same with thomas_c.y/thomas_d.y), for a rank that is both inlet and outlet:
// forward sweep with Neumann first/last rows (b0), mode (0,0) pinned by a zero
// last d', then back substitution.
void DuctPoisson::thomas(F0& val, F1& tc, F1& td, int comp, int j, int k) {
  F0& c_ = comp == 0 ? tc.x : tc.y;
  F0& d_ = comp == 0 ? td.x : td.y;
  const int Nx = g->Nx, bs = g->bs, Nyt = g->Ny_tot, Nzt = g->Nz_tot;
  double TWO_PI(2 * 3.141592653589793);
  double Two_PI_Over_Ly = TWO_PI / g->Ly;
  double Two_PI_Over_Lz = TWO_PI / g->Lz;
  double jj = ((j < Nyt / 2) ? j : j - Nyt) * Two_PI_Over_Ly;
  double kk = ((k < Nzt / 2) ? k : k - Nzt) * Two_PI_Over_Lz;
  int jl = j + bs, kl = k + bs;
  double inv_dx2 = 1. / (g->dx * g->dx);
  double dy = g->dy, dz = g->dz;
  double dy2 = 2. / (dy * dy), dz2 = 2. / (dz * dz);
  double a = inv_dx2;
  double b0 = -((dy2) * (1 - cos(jj * dy)) + (dz2) * (1 - cos(kk * dz))) - inv_dx2;
  double b = -((dy2) * (1 - cos(jj * dy)) + (dz2) * (1 - cos(kk * dz))) - 2 * inv_dx2;
  double c = inv_dx2;
  c_(bs, jl, kl) = c / b0;
  d_(bs, jl, kl) = val(bs, jl, kl) / b0;
  for (int i = bs + 1; i < Nx - bs - 1; i++) {
    c_(i, jl, kl) = c / (b - c_(i - 1, jl, kl) * a);
    d_(i, jl, kl) = (val(i, jl, kl) - d_(i - 1, jl, kl) * a) / (b - c_(i - 1, jl, kl) * a);
  }
  if (j + k == 0)
    d_(Nx - bs - 1, jl, kl) = 0;
  else
    d_(Nx - bs - 1, jl, kl) = (val(Nx - bs - 1, jl, kl) - d_(Nx - bs - 2, jl, kl) * a) /
                              (b0 - c_(Nx - bs - 2, jl, kl) * a);
  val(Nx - bs - 1, jl, kl) = d_(Nx - bs - 1, jl, kl);
  for (int i = Nx - bs - 2; i >= bs; i--)
    val(i, jl, kl) = d_(i, jl, kl) - c_(i, jl, kl) * val(i + 1, jl, kl);
}

}  // namespace orc
