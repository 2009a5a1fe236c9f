// Oracle fields: geometry, tensor0/tensor1 arithmetic, stencils with the
// reference write ranges, one-rank halos and reductions.  TEST INFRASTRUCTURE.
#include <cstdio>
#include <fstream>
#include <sstream>

#include "orc.h"

namespace orc {

// params.cpp:6-64: drop lines whose first char is NUL or '#', then read the
// fields in a fixed order; params.cpp:50-64 derived values.
bool Params::parse(const char* path) {
  std::ifstream in(path);
  if (!in) return false;
  std::ostringstream trimmed;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    trimmed << line << "\n";
  }
  std::istringstream s(trimmed.str());
  s >> IsBox;
  s >> Nx_tot >> Ny_tot >> Nz_tot;
  s >> Lx >> Ly >> Lz;
  s >> bs;
  s >> Np_track;
  s >> T_final >> dt >> threshold;
  s >> Statistics >> Stat_print;
  s >> Data_freq_fast >> Data_freq_slow;
  s >> Data_dir >> Stat_dir;
  s >> Initial;
  s >> PreCond >> Solver;
  s >> cooling;
  s >> Iteration1 >> Iteration2 >> epsilon;
  s >> TWC >> ParticleGravity;
  s >> A1 >> A2;
  s >> Rho0;
  s >> U0;
  s >> T0;
  s >> Mu0 >> k;
  s >> gx1 >> gy1 >> gz1 >> gx2 >> gy2 >> gz2;
  s >> Cp >> Cv;
  s >> np0;
  s >> Dp >> Rhop;
  s >> Cvp;
  s >> epsilonp;
  s >> Nu;
  s >> I01 >> I02;
  if (s.fail()) return false;
  // extension fields appended after the reference's last one (the
  // reference's parser stops before them): per-particle response times
  // uniform on [Tp_lo, Tp_hi]; absent or 0 0 = the single Tp below
  double lo = 0, hi = 0;
  if (s >> lo >> hi) {
    Tp_lo = lo;
    Tp_hi = hi;
  }
  R = Cp - Cv;
  P0 = Rho0 * T0 * R;
  Tp = Rhop * Dp * Dp / (18 * Mu0);
  mp = Rhop * PI * Dp * Dp * Dp / 6.;
  W = Ly;
  N0 = (int)(np0 * Vol());
  Iteration = Iteration1;
  A = A1;
  gx = gx1;
  gy = gy1;
  gz = gz1;
  I0 = I01;
  return true;
}

void Params::update(double time) {
  if (time > threshold) {
    Iteration = Iteration2;
    A = A2;
    gx = gx2;
    gy = gy2;
    gz = gz2;
    I0 = I02;
  }
}

// gridsize.cpp:22-28, 83-88 with one rank (il = jl = kl = 0).
void Geom::setup(int nxt, int nyt, int nzt, int bs_, double lx, double ly,
                 double lz, bool is_box) {
  Nx_tot = nxt;
  Ny_tot = nyt;
  Nz_tot = nzt;
  bs = bs_;
  Lx = lx;
  Ly = ly;
  Lz = lz;
  box = is_box;
  Nx = nxt + 2 * bs;
  Ny = nyt + 2 * bs;
  Nz = nzt + 2 * bs;
  dx = Lx / Nx_tot;
  dy = Ly / Ny_tot;
  dz = Lz / Nz_tot;
  il = jl = kl = 0;
  int ih = il + Nx - 2 * bs - 1, jh = jl + Ny - 2 * bs - 1,
      kh = kl + Nz - 2 * bs - 1;
  xl = il * dx;
  xh = (ih + 1) * dx;
  yl = jl * dy;
  yh = (jh + 1) * dy;
  zl = kl * dz;
  zh = (kh + 1) * dz;
}

// tensor0.cpp:12-31
void F0::init(const Geom* gg) {
  g = gg;
  v.assign(g->size(), 0.0);
  red.assign(g->Nx, 0.0);
  inv_dx = 1. / g->dx;
  inv_dy = 1. / g->dy;
  inv_dz = 1. / g->dz;
  inv_2dx = 1. / (2 * g->dx);
  inv_2dy = 1. / (2 * g->dy);
  inv_2dz = 1. / (2 * g->dz);
  inv_dx2 = 1. / (g->dx * g->dx);
  inv_dy2 = 1. / (g->dy * g->dy);
  inv_dz2 = 1. / (g->dz * g->dz);
  Nx = g->Nx;
  Ny = g->Ny;
  Nz = g->Nz;
  NxNy = Nx * Ny;
  bs = g->bs;
}

// ---- halos on one rank -----------------------------------------------------
// communicator.cpp:177-278 (box): x phase over interior j,k; y phase over all
// i and interior k; z phase over all i,j.  Each phase copies the far interior
// layer into the near ghost.  communicator.cpp:280-396 (duct): the x phase is
// skipped on a rank that is both inlet and outlet.
void F0::Update_Ghosts() {
  F0& T = *this;
  if (g->box) {
    for (int k = bs; k < Nz - bs; k++)
      for (int j = bs; j < Ny - bs; j++)
        for (int b = 0; b < bs; b++) {
          T(Nx - bs + b, j, k) = T(bs + b, j, k);
          T(b, j, k) = T(Nx - 2 * bs + b, j, k);
        }
  }
  for (int k = bs; k < Nz - bs; k++)
    for (int b = 0; b < bs; b++)
      for (int i = 0; i < Nx; i++) {
        T(i, Ny - bs + b, k) = T(i, bs + b, k);
        T(i, b, k) = T(i, Ny - 2 * bs + b, k);
      }
  for (int b = 0; b < bs; b++)
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        T(i, j, Nz - bs + b) = T(i, j, bs + b);
        T(i, j, b) = T(i, j, Nz - 2 * bs + b);
      }
}

// communicator.cpp:422-511 (box) / 520-626 (duct): ghost deposits are added
// into the first/last interior layers, z phase (all i,j), then y phase (all i,
// interior k), then x phase (interior j,k; skipped for the duct on one rank).
// Both send buffers are packed before either receive is added.
void F0::Update_Ghosts_CUM() {
  F0& T = *this;
  std::vector<double> s0, s1;
  // z
  s0.clear();
  s1.clear();
  for (int i = 0; i < Nx; i++)
    for (int j = 0; j < Ny; j++)
      for (int k = Nz - bs; k < Nz; k++) s0.push_back(T(i, j, k));
  for (int i = 0; i < Nx; i++)
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < bs; k++) s1.push_back(T(i, j, k));
  {
    size_t c = 0;
    for (int i = 0; i < Nx; i++)
      for (int j = 0; j < Ny; j++)
        for (int k = Nz - 2 * bs; k < Nz - bs; k++) T(i, j, k) += s1[c++];
    c = 0;
    for (int i = 0; i < Nx; i++)
      for (int j = 0; j < Ny; j++)
        for (int k = bs; k < 2 * bs; k++) T(i, j, k) += s0[c++];
  }
  // y
  s0.clear();
  s1.clear();
  for (int i = 0; i < Nx; i++)
    for (int j = Ny - bs; j < Ny; j++)
      for (int k = bs; k < Nz - bs; k++) s0.push_back(T(i, j, k));
  for (int i = 0; i < Nx; i++)
    for (int j = 0; j < bs; j++)
      for (int k = bs; k < Nz - bs; k++) s1.push_back(T(i, j, k));
  {
    size_t c = 0;
    for (int i = 0; i < Nx; i++)
      for (int j = Ny - 2 * bs; j < Ny - bs; j++)
        for (int k = bs; k < Nz - bs; k++) T(i, j, k) += s1[c++];
    c = 0;
    for (int i = 0; i < Nx; i++)
      for (int j = bs; j < 2 * bs; j++)
        for (int k = bs; k < Nz - bs; k++) T(i, j, k) += s0[c++];
  }
  if (!g->box) return;
  // x
  s0.clear();
  s1.clear();
  for (int i = Nx - bs; i < Nx; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++) s0.push_back(T(i, j, k));
  for (int i = 0; i < bs; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++) s1.push_back(T(i, j, k));
  {
    size_t c = 0;
    for (int i = Nx - 2 * bs; i < Nx - bs; i++)
      for (int j = bs; j < Ny - bs; j++)
        for (int k = bs; k < Nz - bs; k++) T(i, j, k) += s1[c++];
    c = 0;
    for (int i = bs; i < 2 * bs; i++)
      for (int j = bs; j < Ny - bs; j++)
        for (int k = bs; k < Nz - bs; k++) T(i, j, k) += s0[c++];
  }
}

// ---- whole-array arithmetic (tensor0.cpp:93-241) ---------------------------
#define ORC_EACH(expr)                                      \
  {                                                         \
    const size_t n_ = v.size();                             \
    _Pragma("omp parallel for schedule(static)") for (size_t n = 0; n < n_; n++) expr; \
  }
void F0::set(double a) ORC_EACH(v[n] = a)
void F0::add(double a) ORC_EACH(v[n] += a)
void F0::sub(double a) ORC_EACH(v[n] -= a)
void F0::mul(double a) ORC_EACH(v[n] *= a)
void F0::div(double a) ORC_EACH(v[n] /= a)
void F0::copy(const F0& a) ORC_EACH(v[n] = a.v[n])
void F0::add(const F0& a) ORC_EACH(v[n] += a.v[n])
void F0::sub(const F0& a) ORC_EACH(v[n] -= a.v[n])
void F0::mul(const F0& a) ORC_EACH(v[n] *= a.v[n])
void F0::div(const F0& a) ORC_EACH(v[n] /= a.v[n])
void F0::PlusEqual_Mult(const F0& a, const F0& b) ORC_EACH(v[n] += a.v[n] * b.v[n])
void F0::PlusEqual_Mult(const F0& a, double b) ORC_EACH(v[n] += a.v[n] * b)
void F0::PlusEqual_Mult(double a, const F0& b) ORC_EACH(v[n] += a * b.v[n])
void F0::Equal_Divide(const F0& a, const F0& b) ORC_EACH(v[n] = a.v[n] / b.v[n])
void F0::Equal_Divide(const F0& a, double b) ORC_EACH(v[n] = a.v[n] / b)
void F0::Equal_Divide(double a, const F0& b) ORC_EACH(v[n] = a / b.v[n])
void F0::Equal_Mult(const F0& a, const F0& b) ORC_EACH(v[n] = a.v[n] * b.v[n])
void F0::Equal_Mult(const F0& a, double b) ORC_EACH(v[n] = a.v[n] * b)
void F0::Equal_Mult(double a, const F0& b) ORC_EACH(v[n] = a * b.v[n])
void F0::Equal_LinComb(double a, const F0& T, double b, const F0& R)
    ORC_EACH(v[n] = a * T.v[n] + b * R.v[n])
void F0::Equal_LinComb(const F0& A, const F0& T, double b, const F0& R)
    ORC_EACH(v[n] = A.v[n] * T.v[n] + b * R.v[n])
void F0::Equal_LinComb(const F0& A, const F0& T, const F0& B, const F0& R)
    ORC_EACH(v[n] = A.v[n] * T.v[n] + B.v[n] * R.v[n])
void F0::Equal_LinComb(double a, const F0& T, const F0& B, const F0& R)
    ORC_EACH(v[n] = a * T.v[n] + B.v[n] * R.v[n])
void F0::Equal_LinComb(double a, const F0& T, double b, const F0& R, double c,
                       const F0& S)
    ORC_EACH(v[n] = a * T.v[n] + b * R.v[n] + c * S.v[n])
#undef ORC_EACH

// ---- stencils --------------------------------------------------------------
// Loops run k outermost for speed; each write depends only on its inputs, so
// the loop order does not change any value.
#define ORC_RANGE(i0, i1, j0, j1, k0, k1)                          \
  _Pragma("omp parallel for schedule(static)") for (int k = (k0); k < (k1); k++) \
    for (int j = (j0); j < (j1); j++)                               \
      for (int i = (i0); i < (i1); i++)

void F0::Equal_Div_F2C(const F1& T) {  // tensor0.cpp:303-310
  ORC_RANGE(0, Nx - bs, 0, Ny - bs, 0, Nz - bs)
  (*this)(i, j, k) = T.x.ddx_F2C(i, j, k) + T.y.ddy_F2C(i, j, k) + T.z.ddz_F2C(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Del2(const F0& T) {  // tensor0.cpp:312-319
  ORC_RANGE(bs, Nx - bs, bs, Ny - bs, bs, Nz - bs)
  (*this)(i, j, k) = T.d2dx2(i, j, k) + T.d2dy2(i, j, k) + T.d2dz2(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Ix_C2F(const F0& T) {  // tensor0.cpp:353-360
  ORC_RANGE(bs, Nx, 0, Ny, 0, Nz)(*this)(i, j, k) = T.Ix_C2F(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Iy_C2F(const F0& T) {  // tensor0.cpp:362-369
  ORC_RANGE(0, Nx, bs, Ny, 0, Nz)(*this)(i, j, k) = T.Iy_C2F(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Iz_C2F(const F0& T) {  // tensor0.cpp:371-378
  ORC_RANGE(0, Nx, 0, Ny, bs, Nz)(*this)(i, j, k) = T.Iz_C2F(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Ix_F2C(const F0& T) {  // tensor0.cpp:381-388
  ORC_RANGE(0, Nx - bs, 0, Ny, 0, Nz)(*this)(i, j, k) = T.Ix_F2C(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Iy_F2C(const F0& T) {  // tensor0.cpp:390-397
  ORC_RANGE(0, Nx, 0, Ny - bs, 0, Nz)(*this)(i, j, k) = T.Iy_F2C(i, j, k);
  Update_Ghosts();
}
void F0::Equal_Iz_F2C(const F0& T) {  // tensor0.cpp:399-406
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz - bs)(*this)(i, j, k) = T.Iz_F2C(i, j, k);
  Update_Ghosts();
}
void F0::Equal_I_F2C(const F1& T) {  // tensor0.cpp:408-416
  double OneThird = 1. / 3.;
  ORC_RANGE(0, Nx - bs, 0, Ny - bs, 0, Nz - bs)
  (*this)(i, j, k) =
      OneThird * (T.x.Ix_F2C(i, j, k) + T.y.Iy_F2C(i, j, k) + T.z.Iz_F2C(i, j, k));
  Update_Ghosts();
}
void F1::Equal_Grad_C2F(const F0& T) {  // tensor1.cpp:242-260
  const int Nx = T.Nx, Ny = T.Ny, Nz = T.Nz, bs = T.bs;
  ORC_RANGE(bs, Nx, 0, Ny, 0, Nz) x(i, j, k) = T.ddx_C2F(i, j, k);
  ORC_RANGE(0, Nx, bs, Ny, 0, Nz) y(i, j, k) = T.ddy_C2F(i, j, k);
  ORC_RANGE(0, Nx, 0, Ny, bs, Nz) z(i, j, k) = T.ddz_C2F(i, j, k);
  Update_Ghosts();
}

// ---- reductions (tensor0.cpp:418-516): i outer, j, k inner -----------------
double F0::sum() const {
  double s = 0;
  for (int i = bs; i < Nx - bs; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++) s += (*this)(i, j, k);
  return s;
}
double F0::mean() const { return sum() / g->size_tot(); }
void F0::make_mean_zero() {
  double m = mean();
  sub(m);
}
double F0::max() const {
  double m = (*this)(bs, bs, bs);
  for (int i = bs; i < Nx - bs; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++)
        if ((*this)(i, j, k) > m) m = (*this)(i, j, k);
  return m;
}
double F0::min() const {
  double m = (*this)(bs, bs, bs);
  for (int i = bs; i < Nx - bs; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++)
        if ((*this)(i, j, k) < m) m = (*this)(i, j, k);
  return m;
}
static inline double ABS(double a) { return (a >= 0) ? a : -a; }
double F0::max_abs() const {
  double m = ABS((*this)(bs, bs, bs));
  for (int i = bs; i < Nx - bs; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++)
        if (ABS((*this)(i, j, k)) > m) m = ABS((*this)(i, j, k));
  return m;
}
double F0::sum_squares() const {
  double s = 0;
  for (int i = bs; i < Nx - bs; i++)
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++) s += (*this)(i, j, k) * (*this)(i, j, k);
  return s;
}
// communicator.cpp:635-646: y-z plane sums over interior j,k for every i,
// divided by Ny_tot*Nz_tot (Allreduce over one rank is identity).
void F0::reduceX() {
  for (int i = 0; i < Nx; i++) {
    double s = 0;
    for (int j = bs; j < Ny - bs; j++)
      for (int k = bs; k < Nz - bs; k++) s += (*this)(i, j, k);
    s /= (g->Ny_tot * g->Nz_tot);
    red[i] = s;
  }
}
double F1::max() const {  // tensor1.cpp:318-326
  double mx = x.max(), my = y.max(), mz = z.max();
  mx = (my > mx) ? my : mx;
  mx = (mz > mx) ? mz : mx;
  return mx;
}

// ---- post-process (tensor0.cpp:584-625; tensor1.cpp:415-426) --------------
void F0::Equal_SijSij(const F1& U) {
  ORC_RANGE(bs, Nx - bs, bs, Ny - bs, bs, Nz - bs) {
    double& r = (*this)(i, j, k);
    r = pow(U.x.ddx_F2C(i, j, k), 2) + pow(U.y.ddy_F2C(i, j, k), 2) +
        pow(U.z.ddz_F2C(i, j, k), 2);
    r += 0.125 * pow(U.x.ddy_C2C(i, j, k) + U.x.ddy_C2C(i + 1, j, k) +
                         U.y.ddx_C2C(i, j, k) + U.y.ddx_C2C(i, j + 1, k),
                     2);
    r += 0.125 * pow(U.x.ddz_C2C(i, j, k) + U.x.ddz_C2C(i + 1, j, k) +
                         U.z.ddx_C2C(i, j, k) + U.z.ddx_C2C(i, j, k + 1),
                     2);
    r += 0.125 * pow(U.y.ddz_C2C(i, j, k) + U.y.ddz_C2C(i, j + 1, k) +
                         U.z.ddy_C2C(i, j, k) + U.z.ddy_C2C(i, j, k + 1),
                     2);
  }
}
void F0::Equal_nu(double mu, const F0& Rho) {
  Equal_Divide(mu, Rho);
  reduceX();
}
void F0::Equal_dissipation(const F0& nu, const F0& sijsij) {
  const size_t n_ = v.size();
  for (size_t n = 0; n < n_; n++) v[n] = 2 * nu.v[n] * sijsij.v[n];
  reduceX();
}
void F1::Equal_TKE(const F1& U, const F0& R) {
  const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    x(i, j, k) = 0.5 * R(i, j, k) * (U.x(i, j, k) - U.x.red[i]) * (U.x(i, j, k) - U.x.red[i]);
    y(i, j, k) = 0.5 * R(i, j, k) * (U.y(i, j, k) - U.y.red[i]) * (U.y(i, j, k) - U.y.red[i]);
    z(i, j, k) = 0.5 * R(i, j, k) * (U.z(i, j, k) - U.z.red[i]) * (U.z(i, j, k) - U.z.red[i]);
  }
  reduceX();
}
// tensor0.cpp:597-607: adds to the prior content (the reference never
// clears it; grid::Statistics passes a field holding RU.y/Rho)
void F0::Equal_WijWij(const F1& U) {
  ORC_RANGE(bs, Nx - bs, bs, Ny - bs, bs, Nz - bs) {
    double& r = (*this)(i, j, k);
    r += 0.125 * pow(U.x.ddy_C2C(i, j, k) + U.x.ddy_C2C(i + 1, j, k) - U.y.ddx_C2C(i, j, k) -
                         U.y.ddx_C2C(i, j + 1, k),
                     2);
    r += 0.125 * pow(U.x.ddz_C2C(i, j, k) + U.x.ddz_C2C(i + 1, j, k) - U.z.ddx_C2C(i, j, k) -
                         U.z.ddx_C2C(i, j, k + 1),
                     2);
    r += 0.125 * pow(U.y.ddz_C2C(i, j, k) + U.y.ddz_C2C(i, j + 1, k) - U.z.ddy_C2C(i, j, k) -
                         U.z.ddy_C2C(i, j, k + 1),
                     2);
  }
}
void F0::Equal_eta(const F0& nu, const F0& eps) {
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    (*this)(i, j, k) = pow(nu(i, j, k) * nu(i, j, k) * nu(i, j, k) / eps(i, j, k), 0.25);
  }
  reduceX();
}
void F0::Equal_ST(double Tp, const F0& sijsij) {
  const double Tp2 = 2 * Tp * Tp;
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) { (*this)(i, j, k) = sqrt(Tp2 * sijsij(i, j, k)); }
  reduceX();
}
void F0::Equal_Reduced_sd(F0& T) {
  T.reduceX();
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    (*this)(i, j, k) = (T(i, j, k) - T.red[i]) * (T(i, j, k) - T.red[i]);
  }
  reduceX();
  for (int i = 0; i < Nx; i++) red[i] = sqrt(red[i]);
}
double F1::min() const {
  double mx = x.min(), my = y.min(), mz = z.min();
  mx = (my < mx) ? my : mx;
  mx = (mz < mx) ? mz : mx;
  return mx;
}
double F1::max_cfl(double dt) const {
  double mx = x.max() * dt / x.g->dx;
  double my = y.max() * dt / x.g->dy;
  double mz = z.max() * dt / x.g->dz;
  mx = (my > mx) ? my : mx;
  mx = (mz > mx) ? mz : mx;
  return mx;
}
void F1::Equal_TKE2(const F1& RU, const F1& U, const F0& R) {
  const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    x(i, j, k) = 0.5 * R(i, j, k) * (U.x(i, j, k) - RU.x.red[i] / R.red[i]) *
                 (U.x(i, j, k) - RU.x.red[i] / R.red[i]);
    y(i, j, k) = 0.5 * R(i, j, k) * (U.y(i, j, k) - RU.y.red[i] / R.red[i]) *
                 (U.y(i, j, k) - RU.y.red[i] / R.red[i]);
    z(i, j, k) = 0.5 * R(i, j, k) * (U.z(i, j, k) - RU.z.red[i] / R.red[i]) *
                 (U.z(i, j, k) - RU.z.red[i] / R.red[i]);
  }
  reduceX();
}
void F1::Equal_UiUi2(const F1& RU, const F1& U, const F0& R) {
  const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    x(i, j, k) = (U.x(i, j, k) - RU.x.red[i] / R.red[i]) * (U.x(i, j, k) - RU.x.red[i] / R.red[i]);
    y(i, j, k) = (U.y(i, j, k) - RU.y.red[i] / R.red[i]) * (U.y(i, j, k) - RU.y.red[i] / R.red[i]);
    z(i, j, k) = (U.z(i, j, k) - RU.z.red[i] / R.red[i]) * (U.z(i, j, k) - RU.z.red[i] / R.red[i]);
  }
  reduceX();
}
void F1::Equal_Cor_Rho_U(const F1& U, const F0& R) {
  const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    x(i, j, k) = (U.x(i, j, k) - U.x.red[i]) * (R(i, j, k) - R.red[i]);
    y(i, j, k) = (U.y(i, j, k) - U.y.red[i]) * (R(i, j, k) - R.red[i]);
    z(i, j, k) = (U.z(i, j, k) - U.z.red[i]) * (R(i, j, k) - R.red[i]);
  }
  reduceX();
}
void F1::Equal_UiUi(const F1& U, const F0& R) {
  const int Nx = R.Nx, Ny = R.Ny, Nz = R.Nz;
  ORC_RANGE(0, Nx, 0, Ny, 0, Nz) {
    x(i, j, k) = (U.x(i, j, k) - U.x.red[i]) * (U.x(i, j, k) - U.x.red[i]);
    y(i, j, k) = (U.y(i, j, k) - U.y.red[i]) * (U.y(i, j, k) - U.y.red[i]);
    z(i, j, k) = (U.z(i, j, k) - U.z.red[i]) * (U.z(i, j, k) - U.z.red[i]);
  }
  reduceX();
}
#undef ORC_RANGE

}  // namespace orc
