// CPU oracle of the learned preconditioner's network (TEST INFRASTRUCTURE):
// the 3-level 3-D U-Net of the HIP build, forward and backward, in FP64.
// No reference counterpart (the reference has no learned part; its
// variable-coefficient solve is Hypre only, poisson.cpp:115-388).
//
// Layout and parameter order are those of include/fr_amd.h: activations
// channels-last [z][y][x][c] over the interior grid, conv weights
// W[tap][ci][co] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding,
// 2x2x2 average pooling, 2x2x2 stride-2 transposed convolutions with
// W[o][ci][co], o = 4 oz + 2 oy + ox, ReLU after every 3x3x3 convolution.
//
// bf16 = 1 emulates the GPU's storage precision: weights (not biases) and
// every stored activation are rounded to BF16 (round to nearest even), the
// backward operand dL/d(pre-activation) is rounded to BF16 where the GPU
// feeds it to the matrix cores; all arithmetic stays FP64.  bf16 = 0 is the
// exact FP64 network, pinned by torch's conv3d and by finite differences.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double bfr(double v, bool on) {
  if (!on) return v;
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return v;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Dims {
  int nx, ny, nz;
  long n() const { return (long)nx * ny * nz; }
};

const int kConv[10][2] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                          {64, 64}, {64, 32}, {32, 32}, {32, 16}, {16, 16}};
enum { C1A, C1B, C2A, C2B, C3A, C3B, D2A, D2B, D1A, D1B };

struct Offsets {
  long cw[10], cb[10], uw[2], ub[2], ow, ob, n;
  Offsets() {
    long o = 0;
    for (int l = 0; l < 10; l++) {
      cw[l] = o;
      o += 27L * kConv[l][0] * kConv[l][1];
      cb[l] = o;
      o += kConv[l][1];
    }
    const int u[2][2] = {{64, 32}, {32, 16}};
    for (int k = 0; k < 2; k++) {
      uw[k] = o;
      o += 8L * u[k][0] * u[k][1];
      ub[k] = o;
      o += u[k][1];
    }
    ow = o;
    o += 16;
    ob = o;
    o += 1;
    n = o;
  }
};

using Vec = std::vector<double>;

// y[p][co] = act(b[co] + sum_tap sum_ci x[p + off(tap)][ci] W[tap][ci][co])
void conv3(Dims d, int ci, int co, const Vec& x, const double* W, const double* b, bool relu,
           bool bf, Vec& y) {
  y.assign(d.n() * co, 0.0);
  std::vector<double> acc(co);
  for (int z = 0; z < d.nz; z++)
    for (int yy = 0; yy < d.ny; yy++)
      for (int xx = 0; xx < d.nx; xx++) {
        for (int c = 0; c < co; c++) acc[c] = 0.0;
        for (int t = 0; t < 27; t++) {
          const int sz = z + t / 9 - 1, sy = yy + (t / 3) % 3 - 1, sx = xx + t % 3 - 1;
          if (sz < 0 || sz >= d.nz || sy < 0 || sy >= d.ny || sx < 0 || sx >= d.nx) continue;
          const double* xp = &x[(((long)sz * d.ny + sy) * d.nx + sx) * ci];
          const double* wt = W + (long)t * ci * co;
          for (int a = 0; a < ci; a++) {
            const double v = xp[a];
            if (v == 0.0) continue;
            for (int c = 0; c < co; c++) acc[c] += v * wt[a * co + c];
          }
        }
        double* yp = &y[(((long)z * d.ny + yy) * d.nx + xx) * co];
        for (int c = 0; c < co; c++) {
          double v = acc[c] + (b ? b[c] : 0.0);
          if (relu && v < 0) v = 0;
          yp[c] = bfr(v, bf);
        }
      }
}

// dx[q][ci] = sum_tap sum_co dp[q - off(tap)][co] W[tap][ci][co];
// dW[tap][ci][co] += sum_p x[p + off(tap)][ci] dp[p][co]
void conv3_bwd(Dims d, int ci, int co, const Vec& x, const double* W, const Vec& dp, Vec* dx,
               double* dW) {
  if (dx) dx->assign(d.n() * ci, 0.0);
  for (int z = 0; z < d.nz; z++)
    for (int yy = 0; yy < d.ny; yy++)
      for (int xx = 0; xx < d.nx; xx++) {
        const double* g = &dp[(((long)z * d.ny + yy) * d.nx + xx) * co];
        for (int t = 0; t < 27; t++) {
          const int sz = z + t / 9 - 1, sy = yy + (t / 3) % 3 - 1, sx = xx + t % 3 - 1;
          if (sz < 0 || sz >= d.nz || sy < 0 || sy >= d.ny || sx < 0 || sx >= d.nx) continue;
          const long q = ((long)sz * d.ny + sy) * d.nx + sx;
          const double* xp = &x[q * ci];
          const double* wt = W + (long)t * ci * co;
          double* dwt = dW + (long)t * ci * co;
          for (int a = 0; a < ci; a++) {
            double s = 0.0;
            for (int c = 0; c < co; c++) {
              dwt[a * co + c] += xp[a] * g[c];
              s += wt[a * co + c] * g[c];
            }
            if (dx) (*dx)[q * ci + a] += s;
          }
        }
      }
}

void pool(Dims dc, int C, const Vec& in, bool bf, Vec& out) {
  const int fx = 2 * dc.nx, fy = 2 * dc.ny;
  out.assign(dc.n() * C, 0.0);
  for (long q = 0; q < dc.n(); q++) {
    const int x = q % dc.nx, y = (q / dc.nx) % dc.ny, z = q / ((long)dc.nx * dc.ny);
    for (int c = 0; c < C; c++) {
      double s = 0;
      for (int o = 0; o < 8; o++) {
        const long p = ((long)(2 * z + (o >> 2)) * fy + (2 * y + ((o >> 1) & 1))) * fx + 2 * x +
                       (o & 1);
        s += in[p * C + c];
      }
      out[q * C + c] = bfr(s * 0.125, bf);
    }
  }
}

long child(Dims dc, long q, int o) {
  const int x = q % dc.nx, y = (q / dc.nx) % dc.ny, z = q / ((long)dc.nx * dc.ny);
  return ((long)(2 * z + (o >> 2)) * (2 * dc.ny) + (2 * y + ((o >> 1) & 1))) * (2 * dc.nx) +
         2 * x + (o & 1);
}

void upconv(Dims dc, int ci, int co, const Vec& in, const double* W, const double* b, bool bf,
            Vec& out) {
  out.assign(8 * dc.n() * co, 0.0);
  for (long q = 0; q < dc.n(); q++)
    for (int o = 0; o < 8; o++) {
      const long p = child(dc, q, o);
      for (int c = 0; c < co; c++) {
        double s = 0;
        for (int a = 0; a < ci; a++) s += in[q * ci + a] * W[((long)o * ci + a) * co + c];
        out[p * co + c] = bfr(s + b[c], bf);
      }
    }
}

struct Net {
  Offsets O;
  Dims L1, L2, L3;
  bool bf;
  Vec w;  // effective weights (BF16-rounded when bf, biases exact)
  Vec x0, e1a, e1, p1, e2a, e2, p2, b1, b2, u2, cat2, d2a, d2, u1, cat1, d1a, d1;
  Vec y;

  Net(int nx, int ny, int nz, const float* params, bool bf16) : bf(bf16) {
    L1 = {nx, ny, nz};
    L2 = {nx / 2, ny / 2, nz / 2};
    L3 = {nx / 4, ny / 4, nz / 4};
    w.resize(O.n);
    for (long i = 0; i < O.n; i++) w[i] = params[i];
    auto roundw = [&](long off, long cnt) {
      for (long i = 0; i < cnt; i++) w[off + i] = bfr(w[off + i], bf);
    };
    for (int l = 0; l < 10; l++) roundw(O.cw[l], 27L * kConv[l][0] * kConv[l][1]);
    roundw(O.uw[0], 8L * 64 * 32);
    roundw(O.uw[1], 8L * 32 * 16);
    roundw(O.ow, 16);
  }

  static void concat(long n, int ca, const Vec& a, int cb, const Vec& b, Vec& out) {
    out.assign(n * (ca + cb), 0.0);
    for (long p = 0; p < n; p++) {
      for (int c = 0; c < ca; c++) out[p * (ca + cb) + c] = a[p * ca + c];
      for (int c = 0; c < cb; c++) out[p * (ca + cb) + ca + c] = b[p * cb + c];
    }
  }

  void forward(const double* x3) {
    x0.assign(L1.n() * 16, 0.0);
    for (long p = 0; p < L1.n(); p++)
      for (int c = 0; c < 3; c++) x0[p * 16 + c] = bfr(x3[p * 3 + c], bf);
    auto C = [&](int l, Dims d, const Vec& in, Vec& out) {
      conv3(d, kConv[l][0], kConv[l][1], in, &w[O.cw[l]], &w[O.cb[l]], true, bf, out);
    };
    C(C1A, L1, x0, e1a);
    C(C1B, L1, e1a, e1);
    pool(L2, 16, e1, bf, p1);
    C(C2A, L2, p1, e2a);
    C(C2B, L2, e2a, e2);
    pool(L3, 32, e2, bf, p2);
    C(C3A, L3, p2, b1);
    C(C3B, L3, b1, b2);
    upconv(L3, 64, 32, b2, &w[O.uw[0]], &w[O.ub[0]], bf, u2);
    concat(L2.n(), 32, u2, 32, e2, cat2);
    C(D2A, L2, cat2, d2a);
    C(D2B, L2, d2a, d2);
    upconv(L2, 32, 16, d2, &w[O.uw[1]], &w[O.ub[1]], bf, u1);
    concat(L1.n(), 16, u1, 16, e1, cat1);
    C(D1A, L1, cat1, d1a);
    C(D1B, L1, d1a, d1);
    y.assign(L1.n(), 0.0);
    for (long p = 0; p < L1.n(); p++) {
      double s = 0;
      for (int c = 0; c < 16; c++) s += d1[p * 16 + c] * w[O.ow + c];
      y[p] = s + w[O.ob];
    }
  }

  // one 3x3x3 ReLU layer backward: g = dL/d(out) -> grads, returns dL/d(in)
  void layer_bwd(int l, Dims d, const Vec& in, const Vec& out, const Vec& g, Vec* gin,
                 double* G) {
    const int ci = kConv[l][0], co = kConv[l][1];
    Vec dp(g.size()), dpr(g.size());
    for (size_t i = 0; i < g.size(); i++) {
      dp[i] = out[i] > 0 ? g[i] : 0.0;
      dpr[i] = bfr(dp[i], bf);
    }
    for (long p = 0; p < d.n(); p++)
      for (int c = 0; c < co; c++) G[O.cb[l] + c] += dp[p * co + c];
    conv3_bwd(d, ci, co, in, &w[O.cw[l]], dpr, gin, &G[O.cw[l]]);
  }

  void upconv_bwd(int k, Dims dc, int ci, int co, const Vec& in, const Vec& g, Vec& gin,
                  double* G) {
    gin.assign(dc.n() * ci, 0.0);
    const double* W = &w[O.uw[k]];
    for (long q = 0; q < dc.n(); q++)
      for (int o = 0; o < 8; o++) {
        const long p = child(dc, q, o);
        for (int c = 0; c < co; c++) {
          const double gv = g[p * co + c];
          G[O.ub[k] + c] += gv;
          for (int a = 0; a < ci; a++) {
            gin[q * ci + a] += W[((long)o * ci + a) * co + c] * gv;
            G[O.uw[k] + ((long)o * ci + a) * co + c] += in[q * ci + a] * gv;
          }
        }
      }
  }

  static void split(long n, int ca, int cb, const Vec& g, Vec& a, Vec& b) {
    a.assign(n * ca, 0.0);
    b.assign(n * cb, 0.0);
    for (long p = 0; p < n; p++) {
      for (int c = 0; c < ca; c++) a[p * ca + c] = g[p * (ca + cb) + c];
      for (int c = 0; c < cb; c++) b[p * cb + c] = g[p * (ca + cb) + ca + c];
    }
  }

  static void pool_bwd(Dims dc, int C, const Vec& g, Vec& gin) {  // gin += P^T g
    for (long q = 0; q < dc.n(); q++)
      for (int o = 0; o < 8; o++) {
        const long p = child(dc, q, o);
        for (int c = 0; c < C; c++) gin[p * C + c] += g[q * C + c] * 0.125;
      }
  }

  void backward(const double* gy, double* G) {
    for (long i = 0; i < O.n; i++) G[i] = 0.0;
    Vec g_d1(L1.n() * 16);
    for (long p = 0; p < L1.n(); p++) {
      G[O.ob] += gy[p];
      for (int c = 0; c < 16; c++) {
        g_d1[p * 16 + c] = w[O.ow + c] * gy[p];
        G[O.ow + c] += d1[p * 16 + c] * gy[p];
      }
    }
    Vec g_d1a, g_cat1, g_u1, g_e1, g_d2, g_d2a, g_cat2, g_u2, g_e2, g_b2, g_b1, g_p2, g_e2a, g_p1,
        g_e1a;
    layer_bwd(D1B, L1, d1a, d1, g_d1, &g_d1a, G);
    layer_bwd(D1A, L1, cat1, d1a, g_d1a, &g_cat1, G);
    split(L1.n(), 16, 16, g_cat1, g_u1, g_e1);
    upconv_bwd(1, L2, 32, 16, d2, g_u1, g_d2, G);
    layer_bwd(D2B, L2, d2a, d2, g_d2, &g_d2a, G);
    layer_bwd(D2A, L2, cat2, d2a, g_d2a, &g_cat2, G);
    split(L2.n(), 32, 32, g_cat2, g_u2, g_e2);
    upconv_bwd(0, L3, 64, 32, b2, g_u2, g_b2, G);
    layer_bwd(C3B, L3, b1, b2, g_b2, &g_b1, G);
    layer_bwd(C3A, L3, p2, b1, g_b1, &g_p2, G);
    pool_bwd(L3, 32, g_p2, g_e2);
    layer_bwd(C2B, L2, e2a, e2, g_e2, &g_e2a, G);
    layer_bwd(C2A, L2, p1, e2a, g_e2a, &g_p1, G);
    pool_bwd(L2, 16, g_p1, g_e1);
    layer_bwd(C1B, L1, e1a, e1, g_e1, &g_e1a, G);
    layer_bwd(C1A, L1, x0, e1a, g_e1a, nullptr, G);
  }
};

}  // namespace

extern "C" {

long orc_unet_num_params() { return Offsets().n; }

// y = net(x); x is [n][3], y is [n] (n = nx*ny*nz, x fastest).
int orc_unet_forward(int nx, int ny, int nz, const float* params, int bf16, const double* x,
                     double* y) {
  if (nx % 4 || ny % 4 || nz % 4) return 1;
  Net net(nx, ny, nz, params, bf16 != 0);
  net.forward(x);
  std::memcpy(y, net.y.data(), sizeof(double) * net.y.size());
  return 0;
}

// Forward, then the parameter gradient of sum_p gy[p] * y[p] into grads.
int orc_unet_grad(int nx, int ny, int nz, const float* params, int bf16, const double* x,
                  const double* gy, double* y, double* grads) {
  if (nx % 4 || ny % 4 || nz % 4) return 1;
  Net net(nx, ny, nz, params, bf16 != 0);
  net.forward(x);
  if (y) std::memcpy(y, net.y.data(), sizeof(double) * net.y.size());
  net.backward(gy, grads);
  return 0;
}

// One 3x3x3 convolution (no activation) with FP64 weights W[27][ci][co]:
// the unit that torch's conv3d pins.
int orc_conv3(int nx, int ny, int nz, int ci, int co, const double* W, const double* b,
              const double* x, double* y) {
  Dims d{nx, ny, nz};
  Vec xv(x, x + d.n() * ci), yv;
  conv3(d, ci, co, xv, W, b, false, false, yv);
  std::memcpy(y, yv.data(), sizeof(double) * yv.size());
  return 0;
}

}  // extern "C"
