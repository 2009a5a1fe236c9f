// Oracle FFT: unnormalised complex DFT with sign convention
//   X_k = sum_n x_n exp(sign * 2 pi i n k / N).
// TEST INFRASTRUCTURE.  Stands in for the Sandia fft_3d/fft_2d package over
// FFTW 2.1.5 that the reference links (poisson.cpp:87,431,453;
// poisson_IO.cpp:41,168,196); that library is absent, so this transform is
// pinned by the DFT definition (brute force and numpy fixtures in tests/).
#include <vector>

#include "orc.h"

namespace orc {

static void twiddles(int n, int sign, std::vector<double>& c,
                     std::vector<double>& s) {
  c.resize(n / 2 + 1);
  s.resize(n / 2 + 1);
  for (int k = 0; k <= n / 2; k++) {
    double a = 2.0 * PI * (double)k / (double)n;
    c[k] = cos(a);
    s[k] = sign * sin(a);
  }
}

// In-place transform of n elements spaced by `stride`.
void fft1d(double* re, double* im, int n, long stride, int sign) {
  if (n <= 1) return;
  std::vector<double> xr(n), xi(n);
  for (int i = 0; i < n; i++) {
    xr[i] = re[i * stride];
    xi[i] = im[i * stride];
  }
  bool pow2 = (n & (n - 1)) == 0;
  if (pow2) {
    // bit reversal
    for (int i = 1, j = 0; i < n; i++) {
      int bit = n >> 1;
      for (; j & bit; bit >>= 1) j ^= bit;
      j ^= bit;
      if (i < j) {
        std::swap(xr[i], xr[j]);
        std::swap(xi[i], xi[j]);
      }
    }
    std::vector<double> c, s;
    twiddles(n, sign, c, s);
    for (int len = 2; len <= n; len <<= 1) {
      int half = len >> 1, step = n / len;
      for (int b = 0; b < n; b += len)
        for (int m = 0; m < half; m++) {
          double wr = c[m * step], wi = s[m * step];
          double ur = xr[b + m], ui = xi[b + m];
          double tr = xr[b + m + half], ti = xi[b + m + half];
          double vr = tr * wr - ti * wi, vi = tr * wi + ti * wr;
          xr[b + m] = ur + vr;
          xi[b + m] = ui + vi;
          xr[b + m + half] = ur - vr;
          xi[b + m + half] = ui - vi;
        }
    }
  } else {
    std::vector<double> yr(n), yi(n), c(n), s(n);
    for (int k = 0; k < n; k++) {
      double a = 2.0 * PI * (double)k / (double)n;
      c[k] = cos(a);
      s[k] = sign * sin(a);
    }
    for (int k = 0; k < n; k++) {
      double ar = 0, ai = 0;
      for (int m = 0; m < n; m++) {
        int t = (int)(((long)m * k) % n);
        ar += xr[m] * c[t] - xi[m] * s[t];
        ai += xr[m] * s[t] + xi[m] * c[t];
      }
      yr[k] = ar;
      yi[k] = ai;
    }
    xr.swap(yr);
    xi.swap(yi);
  }
  for (int i = 0; i < n; i++) {
    re[i * stride] = xr[i];
    im[i * stride] = xi[i];
  }
}

// 3-D transform of an [nz][ny][nx] array (x fastest).  The forward transform
// (sign -1) runs the x, y, z passes; the inverse (sign +1) runs z, y, x.  The
// pass order does not change the transform; this one is chosen so the device
// solver, which fuses the forward and inverse z passes, matches bit for bit.
void fft3d(std::vector<double>& re, std::vector<double>& im, int nx, int ny,
           int nz, int sign) {
  const long sx = 1, sy = nx, sz = (long)nx * ny;
  auto px = [&]() {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        fft1d(&re[k * sz + j * sy], &im[k * sz + j * sy], nx, sx, sign);
  };
  auto py = [&]() {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++)
        fft1d(&re[k * sz + i], &im[k * sz + i], ny, sy, sign);
  };
  auto pz = [&]() {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        fft1d(&re[j * sy + i], &im[j * sy + i], nz, sz, sign);
  };
  if (sign < 0) {
    px();
    py();
    pz();
  } else {
    pz();
    py();
    px();
  }
}

// Real-input transform by the packing method (e.g. Press et al., Numerical
// Recipes 12.3): with z_m = x_2m + i x_2m+1 and Z its M-point DFT (M = n/2),
//   E_k = (Z_k + conj Z_{M-k})/2,  O_k = (Z_k - conj Z_{M-k})/(2i),
//   X_k = E_k + W^k O_k,  W = exp(-2 pi i/n),  k = 0..M  (Z_M = Z_0).
void rfft1d(const double* x, int n, double* Xre, double* Xim) {
  const int M = n / 2;
  std::vector<double> zr(M), zi(M), c, s;
  for (int m = 0; m < M; m++) {
    zr[m] = x[2 * m];
    zi[m] = x[2 * m + 1];
  }
  fft1d(zr.data(), zi.data(), M, 1, -1);
  twiddles(n, -1, c, s);
  for (int k = 0; k <= M; k++) {
    const int a = (k == M) ? 0 : k, b = (k == 0) ? 0 : M - k;
    const double Ar = zr[a], Ai = zi[a], Br = zr[b], Bi = zi[b];
    const double Er = 0.5 * (Ar + Br), Ei = 0.5 * (Ai - Bi);
    const double Dr = Ar - Br, Di = Ai + Bi;
    const double Or = 0.5 * Di, Oi = -0.5 * Dr;
    const double Tr = Or * c[k] - Oi * s[k], Ti = Or * s[k] + Oi * c[k];
    Xre[k] = Er + Tr;
    Xim[k] = Ei + Ti;
  }
}

// Inverse: Z_k = (X_k + conj X_{M-k}) + i V^k (X_k - conj X_{M-k}),
// V = exp(+2 pi i/n), k < M; z = unnormalised M-point inverse DFT of Z;
// x_2m = Re z_m, x_2m+1 = Im z_m.  Returns n times the input of rfft1d.
void irfft1d(const double* Xre, const double* Xim, int n, double* x) {
  const int M = n / 2;
  std::vector<double> zr(M), zi(M), c, s;
  twiddles(n, +1, c, s);
  for (int k = 0; k < M; k++) {
    const double Ar = Xre[k], Ai = Xim[k], Br = Xre[M - k], Bi = Xim[M - k];
    const double Sr = Ar + Br, Si = Ai - Bi;
    const double Dr = Ar - Br, Di = Ai + Bi;
    const double Tr = Dr * c[k] - Di * s[k], Ti = Dr * s[k] + Di * c[k];
    zr[k] = Sr - Ti;
    zi[k] = Si + Tr;
  }
  fft1d(zr.data(), zi.data(), M, 1, +1);
  for (int m = 0; m < M; m++) {
    x[2 * m] = zr[m];
    x[2 * m + 1] = zi[m];
  }
}

// 2-D transform of an [nz][ny] plane (y fastest).
void fft2d(double* re, double* im, int ny, int nz, int sign) {
  for (int k = 0; k < nz; k++) fft1d(re + (long)k * ny, im + (long)k * ny, ny, 1, sign);
  for (int j = 0; j < ny; j++) fft1d(re + j, im + j, nz, ny, sign);
}

}  // namespace orc
