// CPU oracle: a serial restatement of the reference solver's time-step path
// (fields, stencils, halos, reductions, box/duct pressure solves, particles)
// for ONE MPI rank (1x1x1 process grid: every neighbour is self).
//
// TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP path and the CPU
// baseline timed beside it; nothing in the product links it.  Built with
// -O3 -ffp-contract=off so that a*b+c is never fused (the reference's GCC -O3
// x86-64 build emits no FMA).
//
// Every function cites the reference file:line it restates.
#pragma once
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

const double PI = 3.141592653589793;  // params.h:5

// ---------------------------------------------------------------- params
// params.cpp:6-64 (fixed-order, '#'-commented file) and params.cpp:66-77.
struct Params {
  int IsBox = 1;
  int Nx_tot = 0, Ny_tot = 0, Nz_tot = 0;
  double Lx = 0, Ly = 0, Lz = 0;
  int bs = 1;
  int Np_track = 0;
  double T_final = 0, dt = 0, threshold = 0;
  int Statistics = 0, Stat_print = 0;
  int Data_freq_fast = 1, Data_freq_slow = 1;
  std::string Data_dir, Stat_dir;
  int Initial = 0;
  int PreCond = 0, Solver = 0;
  int cooling = 0;
  int Iteration = 30, Iteration1 = 30, Iteration2 = 30;
  double epsilon = 1e-7;
  int TWC = 1, ParticleGravity = 0;
  double A = 0, A1 = 0, A2 = 0;
  double Rho0 = 1, U0 = 0, T0 = 1;
  double Mu0 = 0, k = 0;
  double gx = 0, gy = 0, gz = 0, gx1 = 0, gy1 = 0, gz1 = 0, gx2 = 0, gy2 = 0, gz2 = 0;
  double Cp = 0, Cv = 0, R = 0;
  double np0 = 0, Dp = 0, Rhop = 0, Cvp = 0, epsilonp = 0, Nu = 0;
  double I0 = 0, I01 = 0, I02 = 0;
  double Tp = 0, mp = 0, P0 = 0, W = 0;
  int N0 = 0;
  // extension (no reference counterpart): per-particle response times
  // uniform on [Tp_lo, Tp_hi], a function of the particle's flag (TpOfFlag);
  // both 0 keeps the single Tp of params.cpp:53
  double Tp_lo = 0, Tp_hi = 0;
  bool per_particle_Tp() const { return Tp_lo != 0 || Tp_hi != 0; }
  double Vol() const { return Lx * Ly * Lz; }
  bool parse(const char* path);  // params.cpp:6-64
  void update(double time);      // params.cpp:66-77
};

// Response time of the particle with flag f under the per-particle
// extension: u = the top 24 bits of the murmur3 32-bit finaliser of f, as a
// fraction in [0, 1); Tp = lo + span * u.  Flags travel with particles, so a
// particle keeps its Tp through migration, storage and restart.
inline double TpOfFlag(int f, double lo, double span) {
  uint32_t h = (uint32_t)f;
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return lo + span * ((double)(h >> 8) * (1.0 / 16777216.0));
}

// ---------------------------------------------------------------- geometry
// gridsize.cpp:4-88 on one rank: Nx = Nx_tot + 2 bs, dx = L/N, il = 0.
struct Geom {
  int Nx = 0, Ny = 0, Nz = 0, bs = 1;
  int Nx_tot = 0, Ny_tot = 0, Nz_tot = 0;
  double Lx = 0, Ly = 0, Lz = 0, dx = 0, dy = 0, dz = 0;
  bool box = true;
  int il = 0, jl = 0, kl = 0;
  double xl = 0, xh = 0, yl = 0, yh = 0, zl = 0, zh = 0;
  void setup(int nxt, int nyt, int nzt, int bs_, double lx, double ly,
             double lz, bool is_box);
  size_t size() const { return (size_t)Nx * Ny * Nz; }
  int size_tot() const { return Nx_tot * Ny_tot * Nz_tot; }
  double Vcell() const { return dx * dy * dz; }
};

// ---------------------------------------------------------------- tensor0
// tensor0.h:6-132: ghosted double[Nz][Ny][Nx], i fastest.
struct F0 {
  const Geom* g = nullptr;
  std::vector<double> v;
  std::vector<double> red;  // ptr_reduce_ (tensor0.h:17)
  double inv_dx, inv_dy, inv_dz, inv_2dx, inv_2dy, inv_2dz, inv_dx2, inv_dy2,
      inv_dz2;
  int Nx, Ny, Nz, NxNy, bs;
  void init(const Geom* gg);  // tensor0.cpp:12-31
  inline double& operator()(int i, int j, int k) {
    return v[(size_t)k * NxNy + (size_t)j * Nx + i];
  }
  inline double operator()(int i, int j, int k) const {
    return v[(size_t)k * NxNy + (size_t)j * Nx + i];
  }
  size_t size() const { return v.size(); }
  // halos (tensor0.cpp:83-91 -> communicator.cpp:177-396, 422-626)
  void Update_Ghosts();
  void Update_Ghosts_CUM();
  // whole-array arithmetic (tensor0.cpp:93-241)
  void set(double a);
  void add(double a);
  void sub(double a);
  void mul(double a);
  void div(double a);
  void copy(const F0& a);
  void add(const F0& a);
  void sub(const F0& a);
  void mul(const F0& a);
  void div(const F0& a);
  void PlusEqual_Mult(const F0& a, const F0& b);
  void PlusEqual_Mult(const F0& a, double b);
  void PlusEqual_Mult(double a, const F0& b);
  void Equal_Divide(const F0& a, const F0& b);
  void Equal_Divide(const F0& a, double b);
  void Equal_Divide(double a, const F0& b);
  void Equal_Mult(const F0& a, const F0& b);
  void Equal_Mult(const F0& a, double b);
  void Equal_Mult(double a, const F0& b);
  void Equal_LinComb(double a, const F0& T, double b, const F0& R);
  void Equal_LinComb(const F0& A, const F0& T, double b, const F0& R);
  void Equal_LinComb(const F0& A, const F0& T, const F0& B, const F0& R);
  void Equal_LinComb(double a, const F0& T, const F0& B, const F0& R);
  void Equal_LinComb(double a, const F0& T, double b, const F0& R, double c,
                     const F0& S);
  // point differences (tensor0.cpp:243-301) and interpolations (322-350)
  inline double ddx_F2C(int i, int j, int k) const {
    return ((*this)(i + 1, j, k) - (*this)(i, j, k)) * inv_dx;
  }
  inline double ddy_F2C(int i, int j, int k) const {
    return ((*this)(i, j + 1, k) - (*this)(i, j, k)) * inv_dy;
  }
  inline double ddz_F2C(int i, int j, int k) const {
    return ((*this)(i, j, k + 1) - (*this)(i, j, k)) * inv_dz;
  }
  inline double ddx_C2C(int i, int j, int k) const {
    return ((*this)(i + 1, j, k) - (*this)(i - 1, j, k)) * inv_2dx;
  }
  inline double ddy_C2C(int i, int j, int k) const {
    return ((*this)(i, j + 1, k) - (*this)(i, j - 1, k)) * inv_2dy;
  }
  inline double ddz_C2C(int i, int j, int k) const {
    return ((*this)(i, j, k + 1) - (*this)(i, j, k - 1)) * inv_2dz;
  }
  inline double ddx_C2F(int i, int j, int k) const {
    return ((*this)(i, j, k) - (*this)(i - 1, j, k)) * inv_dx;
  }
  inline double ddy_C2F(int i, int j, int k) const {
    return ((*this)(i, j, k) - (*this)(i, j - 1, k)) * inv_dy;
  }
  inline double ddz_C2F(int i, int j, int k) const {
    return ((*this)(i, j, k) - (*this)(i, j, k - 1)) * inv_dz;
  }
  inline double d2dx2(int i, int j, int k) const {
    return ((*this)(i + 1, j, k) - 2 * (*this)(i, j, k) + (*this)(i - 1, j, k)) *
           inv_dx2;
  }
  inline double d2dy2(int i, int j, int k) const {
    return ((*this)(i, j + 1, k) - 2 * (*this)(i, j, k) + (*this)(i, j - 1, k)) *
           inv_dy2;
  }
  inline double d2dz2(int i, int j, int k) const {
    return ((*this)(i, j, k + 1) - 2 * (*this)(i, j, k) + (*this)(i, j, k - 1)) *
           inv_dz2;
  }
  inline double Ix_F2C(int i, int j, int k) const {
    return ((*this)(i, j, k) + (*this)(i + 1, j, k)) * 0.5;
  }
  inline double Iy_F2C(int i, int j, int k) const {
    return ((*this)(i, j, k) + (*this)(i, j + 1, k)) * 0.5;
  }
  inline double Iz_F2C(int i, int j, int k) const {
    return ((*this)(i, j, k) + (*this)(i, j, k + 1)) * 0.5;
  }
  inline double Ix_C2F(int i, int j, int k) const {
    return ((*this)(i, j, k) + (*this)(i - 1, j, k)) * 0.5;
  }
  inline double Iy_C2F(int i, int j, int k) const {
    return ((*this)(i, j, k) + (*this)(i, j - 1, k)) * 0.5;
  }
  inline double Iz_C2F(int i, int j, int k) const {
    return ((*this)(i, j, k) + (*this)(i, j, k - 1)) * 0.5;
  }
  // stencils with their reference write ranges
  void Equal_Div_F2C(const struct F1& T);  // tensor0.cpp:303-310
  void Equal_Del2(const F0& T);            // tensor0.cpp:312-319
  void Equal_Ix_C2F(const F0& T);          // tensor0.cpp:353-360
  void Equal_Iy_C2F(const F0& T);          // tensor0.cpp:362-369
  void Equal_Iz_C2F(const F0& T);          // tensor0.cpp:371-378
  void Equal_Ix_F2C(const F0& T);          // tensor0.cpp:381-388
  void Equal_Iy_F2C(const F0& T);          // tensor0.cpp:390-397
  void Equal_Iz_F2C(const F0& T);          // tensor0.cpp:399-406
  void Equal_I_F2C(const struct F1& T);    // tensor0.cpp:408-416
  // reductions (tensor0.cpp:418-516; one rank: Reduce+Bcast is identity)
  double sum() const;
  double mean() const;
  void make_mean_zero();
  double max() const;
  double min() const;
  double max_abs() const;
  double sum_squares() const;
  void reduceX();  // communicator.cpp:635-646
  // post-process (tensor0.cpp:575-662)
  void Equal_SijSij(const struct F1& U);
  void Equal_nu(double mu, const F0& Rho);
  void Equal_dissipation(const F0& nu, const F0& sijsij);
  void Equal_WijWij(const struct F1& U);           // tensor0.cpp:597-607 (+=)
  void Equal_eta(const F0& nu, const F0& eps);     // tensor0.cpp:633-640
  void Equal_ST(double Tp, const F0& sijsij);      // tensor0.cpp:642-650
  void Equal_Reduced_sd(F0& T);                    // tensor0.cpp:652-663
};

// ---------------------------------------------------------------- tensor1
// tensor1.h:6-90: three staggered components.
struct F1 {
  F0 x, y, z;
  void init(const Geom* g) {
    x.init(g);
    y.init(g);
    z.init(g);
  }
  void Update_Ghosts() {
    x.Update_Ghosts();
    y.Update_Ghosts();
    z.Update_Ghosts();
  }
  void Update_Ghosts_CUM() {
    x.Update_Ghosts_CUM();
    y.Update_Ghosts_CUM();
    z.Update_Ghosts_CUM();
  }
  void set(double a) { x.set(a); y.set(a); z.set(a); }
  void copy(const F1& a) { x.copy(a.x); y.copy(a.y); z.copy(a.z); }
  void add(const F1& a) { x.add(a.x); y.add(a.y); z.add(a.z); }
  void mul(const F1& a) { x.mul(a.x); y.mul(a.y); z.mul(a.z); }
  void PlusEqual_Mult(double a, const F1& b) {
    x.PlusEqual_Mult(a, b.x); y.PlusEqual_Mult(a, b.y); z.PlusEqual_Mult(a, b.z);
  }
  void Equal_Divide(const F1& a, const F1& b) {
    x.Equal_Divide(a.x, b.x); y.Equal_Divide(a.y, b.y); z.Equal_Divide(a.z, b.z);
  }
  void Equal_Divide(double a, const F1& b) {
    x.Equal_Divide(a, b.x); y.Equal_Divide(a, b.y); z.Equal_Divide(a, b.z);
  }
  void Equal_Divide(const F1& a, const F0& b) {
    x.Equal_Divide(a.x, b); y.Equal_Divide(a.y, b); z.Equal_Divide(a.z, b);
  }
  void Equal_LinComb(double a, const F1& T, double b, const F1& R) {
    x.Equal_LinComb(a, T.x, b, R.x);
    y.Equal_LinComb(a, T.y, b, R.y);
    z.Equal_LinComb(a, T.z, b, R.z);
  }
  void Equal_LinComb(double a, const F1& T, double b, const F1& R, double c,
                     const F1& S) {
    x.Equal_LinComb(a, T.x, b, R.x, c, S.x);
    y.Equal_LinComb(a, T.y, b, R.y, c, S.y);
    z.Equal_LinComb(a, T.z, b, R.z, c, S.z);
  }
  void Equal_Grad_C2F(const F0& T);  // tensor1.cpp:242-260
  void Equal_Del2(const F1& T) {     // tensor1.cpp:262-267
    x.Equal_Del2(T.x); y.Equal_Del2(T.y); z.Equal_Del2(T.z);
  }
  void Equal_I_C2F(const F0& T) {  // tensor1.cpp:269-274
    x.Equal_Ix_C2F(T); y.Equal_Iy_C2F(T); z.Equal_Iz_C2F(T);
  }
  void Equal_Ix_C2F(const F1& T) {  // tensor1.cpp:283-288
    x.Equal_Ix_C2F(T.x); y.Equal_Ix_C2F(T.y); z.Equal_Ix_C2F(T.z);
  }
  void Equal_Iy_C2F(const F1& T) {  // tensor1.cpp:290-295
    x.Equal_Iy_C2F(T.x); y.Equal_Iy_C2F(T.y); z.Equal_Iy_C2F(T.z);
  }
  void Equal_Iz_C2F(const F1& T) {  // tensor1.cpp:297-302
    x.Equal_Iz_C2F(T.x); y.Equal_Iz_C2F(T.y); z.Equal_Iz_C2F(T.z);
  }
  void make_mean_zero() { x.make_mean_zero(); y.make_mean_zero(); z.make_mean_zero(); }
  void reduceX() { x.reduceX(); y.reduceX(); z.reduceX(); }
  double max() const;  // tensor1.cpp:318-326
  double min() const;  // tensor1.cpp:328-336
  double max_cfl(double dt) const;  // tensor1.cpp:356-365
  void Equal_I_F2C(const F1& T) {  // tensor1.cpp:276-281
    x.Equal_Ix_F2C(T.x); y.Equal_Iy_F2C(T.y); z.Equal_Iz_F2C(T.z);
  }
  void Equal_TKE(const F1& U, const F0& R);  // tensor1.cpp:415-426
  void Equal_TKE2(const F1& RU, const F1& U, const F0& R);  // tensor1.cpp:375-386
  void Equal_UiUi2(const F1& RU, const F1& U, const F0& R);  // tensor1.cpp:388-399
  void Equal_Cor_Rho_U(const F1& U, const F0& R);  // tensor1.cpp:401-412
  void Equal_UiUi(const F1& U, const F0& R);  // tensor1.cpp:428-439
};

// ---------------------------------------------------------------- FFT
// Unnormalised complex DFT, sign s: X_k = sum_n x_n exp(s*2*pi*i*n*k/N).
// In-repo replacement for the Sandia fft_3d/fft_2d over FFTW 2.1.5
// (poisson.cpp:87,431,453; poisson_IO.cpp:41,168,196).  Power-of-two
// lengths use an iterative radix-2 transform; other lengths a direct DFT.
void fft1d(double* re, double* im, int n, long stride, int sign);
void fft3d(std::vector<double>& re, std::vector<double>& im, int nx, int ny,
           int nz, int sign);
void fft2d(double* re, double* im, int ny, int nz, int sign);
// Real transform of n (even, >= 4) contiguous reals: X_k, k = 0..n/2, of the
// sign -1 DFT, via an n/2-point complex transform of z_m = x_2m + i x_2m+1.
void rfft1d(const double* x, int n, double* Xre, double* Xim);
// Inverse of rfft1d (sign +1, unnormalised: returns n * x) from X_0..X_n/2.
void irfft1d(const double* Xre, const double* Xim, int n, double* x);

// ---------------------------------------------------------------- particles
struct Particles {  // particle.h:12-155 (one rank)
  const Params* par = nullptr;
  const Geom* g = nullptr;
  double idx, idy, idz;
  int Np = 0;
  std::vector<double> x, x_int, x_new, x_np1;
  std::vector<double> y, y_int, y_new, y_np1;
  std::vector<double> z, z_int, z_new, z_np1;
  std::vector<double> u, u_int, u_new, u_np1;
  std::vector<double> v, v_int, v_new, v_np1;
  std::vector<double> w, w_int, w_new, w_np1;
  std::vector<double> T, T_int, T_new, T_np1;
  std::vector<double> RHS_u, RHS_v, RHS_w, RHS_T, ug, vg, wg, Tg;
  std::vector<int> flag;
  std::vector<double> Sbuf0, Sbuf1;  // migration send lists (29 per particle)
  void init(const Params* p, const Geom* gg);
  void RESIZE();
  void update_position(int k, double pre, double post);  // particle.cpp:314-338
  void update_velocity(int k, double pre, double post);  // particle.cpp:362-399
  void update_Temp(int k, double pre, double post);      // particle.cpp:480-496
  void gas2part_velocity(const F1& U);                   // particle.cpp:340-360
  void gas2part_Temp(const F0& T, bool use_new);         // particle.cpp:402-456
  void Compute_RHS_Temp(bool use_new);                   // particle.cpp:460-478
  void part2gas_velocity(F1& A);                         // particle.cpp:498-524
  void part2gas_concentration(F0& A);                    // particle.cpp:526-559
  void part2gas_Temp(F0& A, bool use_new);               // particle.cpp:586-658
  void Send_Recv();                                      // particle.cpp:853-858
  // duct rank that is inlet and outlet: y,z periodic, x leavers dropped, the
  // inlet particles appended (particle.cpp:814-851, 860-865); consumes inbuf
  void Send_Recv_duct(std::vector<double>& inbuf);
  void SortValid(int axis, bool keep = true);            // particle.cpp:150-250
  void load_random();                                    // particle.cpp:919-958
  void copy_stage(int dst, int src);  // whole-vector stage copies (grid.cpp)
  std::vector<double>* stage(int q, int s);  // q in 0..6 (x..T), s in 0..3
};

// glibc rand() (TYPE_3 additive feedback generator, default seed 1),
// restated from its published description so the sequence does not depend on
// other rand() consumers in the process.
struct GlibcRand {
  uint32_t ring[34];
  int head = 0;  // index of r[i-34] in the ring
  void seed(unsigned s);
  int next();
};

// ---------------------------------------------------------------- poisson
struct BoxPoisson {  // poisson.h:19-81, FFT branch (Solver=0)
  const Params* par = nullptr;
  const Geom* g = nullptr;
  F0 center_coeff, RHS_new;
  F1 dummy;
  int num_iteration = 0;
  std::vector<double> re, im;
  void init(const Params* p, const Geom* gg);
  void Solve(const F1& C, F0& P, const F0& RHS);  // poisson.cpp:115-117,391-411
  void CCP_FFT(const F0& RHS, F0& P);             // poisson.cpp:413-462
  void CCP_FFT_complex(const F0& RHS, F0& P);     // full complex, odd Nx
};

struct DuctPoisson {  // poisson_IO.h:20-88
  const Params* par = nullptr;
  const Geom* g = nullptr;
  F0 center_coeff, RHS_new, fft_real, fft_imag;
  F1 dummy, thomas_c, thomas_d;
  int num_iteration1 = 0;
  double num_iteration2 = 0;
  void init(const Params* p, const Geom* gg);
  void Solve(const F1& C, F0& P, const F0& RHS);  // poisson_IO.cpp:81-114
  void CCP_FFT(const F0& RHS, F0& P);             // poisson_IO.cpp:116-206
  void CCP_FFT_complex(const F0& RHS, F0& P);     // full complex, odd Ny
  void thomas(F0& val, F1& tc, F1& td, int comp, int j, int k);  // 208-452
  void thomas_line(double* v, long stride, long off, int j, int k) const;
};

// ---------------------------------------------------------------- PCG + GMG
// Preconditioned CG on S = -div(C grad), C = 1/Rho_face, with one FP32
// geometric-multigrid V-cycle as preconditioner (stands in for the Hypre
// PCG/BoomerAMG branch, poisson.cpp:115-248).  Levels hold interior cells
// only, face coefficients C*inv_d2 (left face of each cell), periodic wrap,
// and closed x ends on the duct.
struct MGLevel {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> cx, cy, cz, e, r, t;
  size_t at(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
};

struct MGPCG {
  const Geom* g = nullptr;
  bool duct = false;
  std::vector<MGLevel> L;
  int pre = 2, post = 2, coarse_sweeps = 20;
  int max_levels = 0;  // 0: coarsen while every size is even and >= 8
  void build(const Geom* gg);
  void coefficients(const F1& Rho_face);
  void smooth(MGLevel& v, int color);
  void residual(MGLevel& v);
  void vcycle(size_t level);
  // P: initial guess in, solution out (ghosts filled); RHS zero-mean.
  // Stops when ||r||_2 <= tol ||RHS||_2 or after maxit iterations.
  int solve(const F1& Rho_face, F0& P, const F0& RHS, int maxit, double tol, double* relres);
};

// ---------------------------------------------------------------- duct inflow
// What the duct copies from the box grid g each substep (grid.cpp:351-364,
// 482-504), as ghosted [Nz][Ny] planes (j fastest): rho_new = g.Rho_new(1,j,k),
// rho_np1 = g.Rho_np1(1,j,k), ru_new[0] = g.RU_new.x(2,j,k), ru_new[1..2] =
// g.RU_new.{y,z}(1,j,k), ru_np1 likewise; and the particles entering through
// the inlet (the box's Rbuf1, particle.cpp:814-851), 29 values each in ToBuf
// order (particle.cpp:252-262), consumed by the next duct migration.
struct Inflow {
  std::vector<double> rho_new, rho_np1, ru_new[3], ru_np1[3];
  std::vector<double> pbuf;
};

// ---------------------------------------------------------------- grid
struct Grid {  // grid.h:16-115
  Params par;
  Geom g;
  double RK4_pre[4], RK4_post[4];
  double mean_energy_transferred = 0, C_BC = 0, U_avg_in = 0, U_avg_out = 0;
  Particles part;
  F1 RU, RU_int, RU_new, RU_np1, RHS_RU, U, Rho_face, dummy, dummy2, dummy3;
  F0 P, RHS_Pois, C, Rho, Rho_int, Rho_new, Rho_np1, RHS_Rho, T, divergence,
      RHS_Part_Temp;
  double P0 = 0, P0_int = 0, P0_new = 0, P0_np1 = 0, dP0_dt = 0, T_cur = 0;
  int num_timestep = 0, RK4_count = 0;
  BoxPoisson PS;
  DuctPoisson PS_IO;
  Inflow in;  // duct only
  MGPCG mg;   // Solver = 2
  double last_relres = 0;
  int last_iterations = 0;
  void setup(const Params& p);
  // Initialize (grid.cpp:110-231), Initial = 0/2 forms; the velocity field of
  // Initial=0 is passed in as three interior blocks [Nz][Ny][Nx].
  void Initialize(const double* Ublk, const double* Vblk, const double* Wblk);
  void Finish_Initialize();  // grid.cpp:197-229 (after fields/particles set)
  void Update_Rho();          // grid.cpp:327-333
  void Update_P0();           // grid.cpp:539-554
  void Update_Particle();     // grid.cpp:659-680
  void Update_RU_WOP();       // grid.cpp:387-425
  void Compute_Div_U_new();   // grid.cpp:556-563
  void Compute_RHS_Pois();    // grid.cpp:629-637
  void Solve_Poisson();       // grid.cpp:639-643
  void Solve_Poisson_IO();    // grid.cpp:645-649
  void Update_RU_WP();        // grid.cpp:651-657
  void TimeAdvance_RK4();     // grid.cpp:705-712
  void Correct_mean_RU();     // grid.cpp:565-577
  void TimeAdvance();         // grid.cpp:315-325
  void box_step();            // simulation.cpp:34-71, box calls only
  // duct (IsBox = 0) forms, with the box grid replaced by `in`
  void Update_Rho_duct();       // grid.cpp:335-385
  void Update_RU_WOP_duct();    // grid.cpp:427-537
  void Update_Particle_duct();  // grid.cpp:682-703
  void Correct_OutFlow();       // grid.cpp:579-627
  void duct_step();             // simulation.cpp:36-70, duct calls
  // step diagnostics (grid.cpp:714-866 subset) + mean particle T (extension)
  void diagnostics(double* out);
  // run files, relative to the working directory (io.cpp): touch.check and
  // the stat files at creation when `outputs` is set (grid.cpp:25-49),
  // Statistics (714-866), Store (233-313), Touch (77-98), the Initial = 1
  // restart (124-182), Write_info (869-973)
  bool outputs = false;
  std::string stat_files[15];
  int Is_touch = 0;
  void open_outputs();
  void Statistics();
  bool Store();
  bool Touch();
  bool Restart();
  void Write_info();
  F0* field0(const std::string& name);
  F1* field1(const std::string& name);
};

// The coupled run of simulation.cpp:19-73 on one rank: a periodic box `b`
// feeding a duct `d` of the same Ny, Nz (grid.cpp:110-182 and 335-537 with
// the box's planes, particle.cpp:814-851 with the box's '>hi' list).
void pair_inflow(Grid& b, Grid& d, bool init);  // the duct's `in` from the box
void pair_initialize(Grid& b, Grid& d, const double* const ic_b[3],
                     const double* const ic_d[3]);
void pair_step(Grid& b, Grid& d);   // simulation.cpp:34-62
bool pair_touch(Grid& b, Grid& d);  // grid::Touch(grid& g_IO), grid.cpp:77-98

}  // namespace orc
