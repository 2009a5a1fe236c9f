// Host side of the box solver: the reference classes (params, gridsize,
// tensor0, tensor1, particle, poisson, grid) driving the gfx950 kernels, and
// the fr_sim_* C entry points.  Compiled by hipcc with the kernels into
// libfr_amd.so.
#include <array>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "fr_host.h"

using namespace fr;

// ---------------------------------------------------------------- params
// params.cpp:6-64: lines whose first character is NUL or '#' are dropped,
// the rest are read as whitespace-separated tokens in a fixed order.
params::params(const char* path) {
  std::ifstream in(path);
  if (!in) return;
  std::ostringstream trimmed;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    trimmed << line << "\n";
  }
  std::istringstream s(trimmed.str());
  s >> IsBox_;
  s >> Nx_tot_ >> Ny_tot_ >> Nz_tot_;
  s >> Lx_ >> Ly_ >> Lz_;
  s >> bs_;
  s >> Np_track_;
  s >> T_final_ >> dt_ >> threshold_;
  s >> Statistics_ >> Stat_print_;
  s >> Data_freq_fast_ >> Data_freq_slow_;
  s >> Data_dir_ >> Stat_dir_;
  s >> Initial_;
  s >> PreCond_ >> Solver_;
  s >> cooling_;
  s >> Iteration1_ >> Iteration2_ >> epsilon_;
  s >> TWC_ >> ParticleGravity_;
  s >> A1_ >> A2_;
  s >> Rho0_;
  s >> U0_;
  s >> T0_;
  s >> Mu0_ >> k_;
  s >> gx1_ >> gy1_ >> gz1_ >> gx2_ >> gy2_ >> gz2_;
  s >> Cp_ >> Cv_;
  s >> np0_;
  s >> Dp_ >> Rhop_;
  s >> Cvp_;
  s >> epsilonp_;
  s >> Nu_;
  s >> I01_ >> I02_;
  if (s.fail()) return;
  // extension fields after the reference's last one (its parser stops
  // before them); a reference params.in has none and keeps the single Tp
  double lo = 0, hi = 0;
  if (s >> lo >> hi) {
    Tp_lo_ = lo;
    Tp_hi_ = hi;
  }
  // params.cpp:50-63
  R_ = Cp_ - Cv_;
  P0_ = Rho0_ * T0_ * R_;
  Tp_ = Rhop_ * Dp_ * Dp_ / (18 * Mu0_);
  mp_ = Rhop_ * PI * Dp_ * Dp_ * Dp_ / 6.;
  W_ = Ly_;
  N0_ = (int)(np0_ * Vol());
  Iteration_ = Iteration1_;
  A_ = A1_;
  gx_ = gx1_;
  gy_ = gy1_;
  gz_ = gz1_;
  I0_ = I01_;
  ok = true;
}

void params::update(double time) {  // params.cpp:66-77
  if (time > threshold_) {
    Iteration_ = Iteration2_;
    A_ = A2_;
    gx_ = gx2_;
    gy_ = gy2_;
    gz_ = gz2_;
    I0_ = I02_;
  }
}

// gridsize.cpp:22-28, 83-88 on one rank.
gridsize::gridsize(const params& p) {
  Nx_tot_ = p.Nx_tot();
  Ny_tot_ = p.Ny_tot();
  Nz_tot_ = p.Nz_tot();
  g.bs = p.bs();
  g.Nx = Nx_tot_ + 2 * g.bs;
  g.Ny = Ny_tot_ + 2 * g.bs;
  g.Nz = Nz_tot_ + 2 * g.bs;
  g.Lx = p.Lx();
  g.Ly = p.Ly();
  g.Lz = p.Lz();
  g.dx = g.Lx / Nx_tot_;
  g.dy = g.Ly / Ny_tot_;
  g.dz = g.Lz / Nz_tot_;
  g.is_box = p.IsBox() ? 1 : 0;
  xl_ = 0 * g.dx;
  xh_ = (Nx_tot_ - 1 + 1) * g.dx;
  yl_ = 0 * g.dy;
  yh_ = (Ny_tot_ - 1 + 1) * g.dy;
  zl_ = 0 * g.dz;
  zh_ = (Nz_tot_ - 1 + 1) * g.dz;
}

// ---------------------------------------------------------------- tensor0
tensor0::tensor0(gridsize* p, hipStream_t st) { init(p, st); }

void tensor0::init(gridsize* p, hipStream_t st, bool fp32) {
  p_ = p;
  st_ = st;
  f32 = fp32;
  const size_t bytes = es() * (size_t)p_->size();
  if (hipMalloc(&ptr, bytes) != hipSuccess) {
    fprintf(stderr, "fr_amd: tensor0 allocation of %d values failed\n", p_->size());
    ptr = nullptr;
    return;
  }
  hipMemsetAsync(ptr, 0, bytes, st_);
}

tensor0::~tensor0() {
  if (ptr) hipFree(ptr);
}

// The FP32 storage of the precision form and the FP64 storage share one code
// path through these helpers; reductions accumulate in double either way.
#define FR_T0_DISPATCH(expr_f, expr_d) \
  do {                                 \
    if (f32) {                         \
      expr_f;                          \
    } else {                           \
      expr_d;                          \
    }                                  \
  } while (0)

void tensor0::Update_Ghosts() {
  const Dev d = p_->dev();
  FR_T0_DISPATCH(ghost_fill(d, p<float>(), st_), ghost_fill(d, p<double>(), st_));
}
void tensor0::Update_Ghosts_CUM() {
  const Dev d = p_->dev();
  FR_T0_DISPATCH(cum_fold(d, p<float>(), st_), cum_fold(d, p<double>(), st_));
}
void tensor0::assign(double a) {
  const long n = p_->size();
  FR_T0_DISPATCH(elementwise<float>(n, 5, p<float>(), a, nullptr, 0, nullptr, 0, nullptr, st_),
                 elementwise<double>(n, 5, p<double>(), a, nullptr, 0, nullptr, 0, nullptr, st_));
}
void tensor0::add(double a) {
  const long n = p_->size();
  FR_T0_DISPATCH(elementwise<float>(n, 6, p<float>(), a, nullptr, 0, nullptr, 0, nullptr, st_),
                 elementwise<double>(n, 6, p<double>(), a, nullptr, 0, nullptr, 0, nullptr, st_));
}
// Operands of one operation share one storage type; a mixed call would read
// past the end of the narrower buffer, so it is refused.
static bool same_type(bool a, bool b, const char* what) {
  if (a == b) return true;
  fprintf(stderr, "fr_amd: %s on fields of different precision refused\n", what);
  return false;
}
void tensor0::copy(const tensor0& a) {
  if (!same_type(f32, a.f32, "copy")) return;
  hipMemcpyAsync(ptr, a.ptr, es() * (size_t)p_->size(), hipMemcpyDeviceToDevice, st_);
}
void tensor0::Equal_Divide(double a, const tensor0& b) {
  const long n = p_->size();
  if (!same_type(f32, b.f32, "Equal_Divide")) return;
  FR_T0_DISPATCH(
      elementwise<float>(n, 3, p<float>(), a, b.p<float>(), 0, nullptr, 0, nullptr, st_),
      elementwise<double>(n, 3, p<double>(), a, b.p<double>(), 0, nullptr, 0, nullptr, st_));
}
void tensor0::Equal_Divide(const tensor0& a, const tensor0& b) {
  const long n = p_->size();
  if (!same_type(f32, a.f32, "Equal_Divide") || !same_type(f32, b.f32, "Equal_Divide")) return;
  FR_T0_DISPATCH(
      elementwise<float>(n, 4, p<float>(), 0, a.p<float>(), 0, b.p<float>(), 0, nullptr, st_),
      elementwise<double>(n, 4, p<double>(), 0, a.p<double>(), 0, b.p<double>(), 0, nullptr,
                          st_));
}
double tensor0::reduce(int op) {
  double r = 0;
  const Dev d = p_->dev();
  FR_T0_DISPATCH(reduce_sync(d, (const float*)ptr, op, &r, st_),
                 reduce_sync(d, (const double*)ptr, op, &r, st_));
  return r;
}
double tensor0::sum() { return reduce(0); }
double tensor0::mean() { return reduce(1); }
void tensor0::make_mean_zero() {  // tensor0.cpp:437-441: (*this) -= mean
  double m = mean();
  add(-m);
}
void tensor0::make_mean_zero_async(double* ws) {
  const Dev d = p_->dev();
  const long n = p_->size();
  FR_T0_DISPATCH(fr::make_mean_zero_async(d, p<float>(), n, ws, st_),
                 fr::make_mean_zero_async(d, p<double>(), n, ws, st_));
}
double tensor0::max() { return reduce(2); }
double tensor0::min() { return reduce(3); }
double tensor0::max_abs() { return reduce(4); }
void tensor0::download(double* host) const {
  const size_t n = (size_t)p_->size();
  if (f32) {
    std::vector<float> h(n);
    hipMemcpyAsync(h.data(), ptr, sizeof(float) * n, hipMemcpyDeviceToHost, st_);
    hipStreamSynchronize(st_);
    for (size_t i = 0; i < n; i++) host[i] = (double)h[i];
    return;
  }
  hipMemcpyAsync(host, ptr, sizeof(double) * n, hipMemcpyDeviceToHost, st_);
  hipStreamSynchronize(st_);
}
void tensor0::upload(const double* host) {
  const size_t n = (size_t)p_->size();
  if (f32) {
    std::vector<float> h(n);
    for (size_t i = 0; i < n; i++) h[i] = (float)host[i];
    hipMemcpyAsync(ptr, h.data(), sizeof(float) * n, hipMemcpyHostToDevice, st_);
    hipStreamSynchronize(st_);
    return;
  }
  hipMemcpyAsync(ptr, host, sizeof(double) * n, hipMemcpyHostToDevice, st_);
  hipStreamSynchronize(st_);
}

// ---------------------------------------------------------------- particle
// glibc random_r TYPE_3 (default seed 1): r[i] = 16807 r[i-1] mod (2^31-1)
// for i < 31, r[i] = r[i-31] for i < 34, then r[i] = r[i-31] + r[i-3] mod
// 2^32 with the first 310 outputs discarded; rand() = r >> 1.  Restated from
// its published description so that placement does not depend on any other
// rand() caller in the process.
namespace {
struct GlibcRand {
  uint32_t ring[34];
  int head = 0;
  explicit GlibcRand(unsigned seed) {
    int32_t r[34];
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
      long hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
      long word = 16807 * lo - 2836 * hi;
      if (word < 0) word += 2147483647;
      r[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 0; i < 34; i++) ring[i] = (uint32_t)r[i];
    for (int i = 34; i < 344; i++) next();
  }
  int next() {
    uint32_t v = ring[(head + 3) % 34] + ring[(head + 31) % 34];
    ring[head] = v;
    head = (head + 1) % 34;
    return (int)(v >> 1);
  }
};
}  // namespace

int particle::init(params* p, gridsize* s, hipStream_t st, int capacity) {
  param_ = p;
  size_ = s;
  st_ = st;
  Np = 0;
  return dev.alloc(capacity, s->dev());
}

// particle.cpp:919-958 on one rank: x,y,z uniform from rand(), u = U0,
// v = w = 0, T = T0, flag = running index; all other stages zero.
int particle::load_random() {
  const int n = param_->N0();
  if (n > dev.cap) return FR_EOVERFLOW;
  Np = n;
  dev.np = n;
  GlibcRand rng(1);
  const double RM = 2147483647.0;
  std::vector<double> X(n), Y(n), Z(n), Uu(n, param_->U0()), zero(n, 0.0), T(n, param_->T0());
  std::vector<int> fl(n);
  const double xl = size_->xl_, xh = size_->xh_, yl = size_->yl_, yh = size_->yh_,
               zl = size_->zl_, zh = size_->zh_;
  for (int i = 0; i < n; i++) {
    X[i] = ((double)rng.next() / RM) * (xh - xl) + xl;
    Y[i] = ((double)rng.next() / RM) * (yh - yl) + yl;
    Z[i] = ((double)rng.next() / RM) * (zh - zl) + zl;
    fl[i] = i;
  }
  const std::vector<double>* base[7] = {&X, &Y, &Z, &Uu, &zero, &zero, &T};
  for (int q = 0; q < 7; q++)
    for (int t = 0; t < 4; t++)
      FR_CHECK(hipMemcpyAsync(dev.s[q][t], t == 0 ? base[q]->data() : zero.data(),
                              sizeof(double) * n, hipMemcpyHostToDevice, st_));
  FR_CHECK(hipMemcpyAsync(dev.flag, fl.data(), sizeof(int) * n, hipMemcpyHostToDevice, st_));
  int rc = dev.reset_order(st_);
  if (rc) return rc;
  FR_CHECK(hipStreamSynchronize(st_));
  return FR_OK;
}

int particle::advance(int k, double pre, double post, tensor1& U, bool set0, bool eq) {
  const params& p = *param_;
  AdvCoef c;
  const double dt = p.dt();
  c.dt_post = dt * post;
  c.dt_pre = dt * pre;
  c.iTp = (p.Tp() != 0) ? 1. / p.Tp() : 0.0;
  c.per_particle = (p.Tp_lo_ != 0 || p.Tp_hi_ != 0);
  c.tp_lo = p.Tp_lo_;
  c.tp_span = p.Tp_hi_ - p.Tp_lo_;
  c.gx = p.ParticleGravity() * p.gx();
  c.gy = p.ParticleGravity() * p.gy();
  c.gz = p.ParticleGravity() * p.gz();
  const double iS = 1. / (p.mp() * p.Cvp());
  c.dtpost_iS = dt * post * iS;
  c.dtpre_iS = dt * pre * iS;
  c.alpha = 0.25 * p.epsilonp() * PI * p.Dp() * p.Dp() * p.I0();
  c.last = (k == 3);
  c.set0 = set0 && c.last;
  if (U.x.is_f32()) {
    const float* u[3] = {U.x.p<float>(), U.y.p<float>(), U.z.p<float>()};
    return dev.advance(size_->dev(), c, u, st_, eq);
  }
  const double* u[3] = {U.x.pointer(), U.y.pointer(), U.z.pointer()};
  return dev.advance(size_->dev(), c, u, st_, eq);
}

// The rounded-index buckets were built on the _new positions at the end of
// the previous substep (or at initialisation), which are this substep's _int
// positions: particle order and positions are unchanged since.
int particle::part2gas_concentration(tensor0& A, hipStream_t on) {
  const hipStream_t s = on ? on : st_;
  if (A.is_f32()) return dev.deposit_cic(size_->dev(), false, A.p<float>(), false, s);
  return dev.deposit_cic(size_->dev(), false, A.pointer(), false, s);
}

int particle::part2gas_velocity(tensor1& A, const int* rk, bool halo, hipEvent_t head) {
  int rc;
  if (head) {
    FR_CHECK(hipStreamWaitEvent(st_, head, 0));
    rc = dev.build_tail(size_->dev(), 1, 1, st_, true, rk);
  } else {
    rc = dev.build(size_->dev(), 1, 1, st_, true, rk);
  }
  if (rc) return rc;
  if (A.x.is_f32()) {
    float* o[3] = {A.x.p<float>(), A.y.p<float>(), A.z.p<float>()};
    return dev.deposit_vel(size_->dev(), o, st_, halo);
  }
  double* o[3] = {A.x.pointer(), A.y.pointer(), A.z.pointer()};
  return dev.deposit_vel(size_->dev(), o, st_, halo);
}

int particle::Send_Recv(hipStream_t on) {
  const double lo[3] = {size_->xl_, size_->yl_, size_->zl_};
  const double hi[3] = {size_->xh_, size_->yh_, size_->zh_};
  const double L[3] = {size_->Lx(), size_->Ly(), size_->Lz()};
  return dev.migrate(lo, hi, L, on ? on : st_);
}

int particle::Send_Recv_duct(const double* inlet, int n) {
  const double lo[3] = {size_->xl_, size_->yl_, size_->zl_};
  const double hi[3] = {size_->xh_, size_->yh_, size_->zh_};
  const double L[3] = {size_->Lx(), size_->Ly(), size_->Lz()};
  int rc = dev.migrate(lo, hi, L, st_, true, inlet, n);
  Np = dev.np;
  return rc;
}

int particle::Send_Recv_duct_dev(const double* rows, int n) {
  const double lo[3] = {size_->xl_, size_->yl_, size_->zl_};
  const double hi[3] = {size_->xh_, size_->yh_, size_->zh_};
  const double L[3] = {size_->Lx(), size_->Ly(), size_->Lz()};
  int rc = dev.migrate(lo, hi, L, st_, true, rows, n, true);
  Np = dev.np;
  return rc;
}

int particle::temp(int stage, double c0, const tensor0& Rho, hipStream_t on) {
  const double iTth = param_->Nu() * PI * param_->Dp() * param_->k();
  const hipStream_t s = on ? on : st_;
  if (Rho.is_f32()) return dev.temp(size_->dev(), stage, c0, Rho.p<float>(), iTth, s);
  return dev.temp(size_->dev(), stage, c0, Rho.pointer(), iTth, s);
}

// Requires rebuild_cells(stage) after the RHS_T of that stage was formed.
int particle::part2gas_Temp(int stage, tensor0& A, hipStream_t on, double* sum, bool* summed) {
  (void)stage;
  const hipStream_t s = on ? on : st_;
  if (A.is_f32()) return dev.deposit_cic(size_->dev(), true, A.p<float>(), true, s, sum, summed);
  return dev.deposit_cic(size_->dev(), true, A.pointer(), true, s, sum, summed);
}

int particle::rebuild_cells(int stage, hipStream_t on) {
  return dev.build(size_->dev(), stage, 0, on ? on : st_);
}

// ---------------------------------------------------------------- poisson
int poisson::init(params* p, gridsize* g, hipStream_t st) {
  par = p;
  st_ = st;
  if (p->Solver() == 2) return fr_pcg_create(&g->g, &sp);
  return g->IsBox() ? fr_box_poisson_create(&g->g, &s) : fr_duct_poisson_create(&g->g, &sd);
}
poisson::~poisson() {
  fr_box_poisson_destroy(s);
  fr_duct_poisson_destroy(sd);
  fr_pcg_destroy(sp);
}
int poisson::Solve(tensor1& C, tensor0& P, tensor0& RHS) {
  if (sp) return FR_EINVAL;  // the PCG solver takes Rho_face (Solve_face_rho)
  if (sd)
    return fr_duct_poisson_solve(sd, C.x.pointer(), C.y.pointer(), C.z.pointer(), P.pointer(),
                                 RHS.pointer(), par->Iteration(), par->epsilon(),
                                 &num_iteration_, st_);
  return fr_box_poisson_solve(s, C.x.pointer(), C.y.pointer(), C.z.pointer(), P.pointer(),
                              RHS.pointer(), par->Iteration(), par->epsilon(),
                              &num_iteration_, st_);
}
int poisson::Solve_face_rho(tensor1& Rho_face, tensor0& P, tensor0& RHS, const double* rsum) {
  if (rsum && !s) return FR_EINVAL;  // a pending mean shift only in the box FFT solve
  if (sp)
    return pcg_solve(sp, Rho_face.x.pointer(), Rho_face.y.pointer(), Rho_face.z.pointer(),
                     P.pointer(), RHS.pointer(), par->Iteration(), par->epsilon(),
                     &num_iteration_, &relres_, st_);
  if (s && Rho_face.x.is_f32())
    return box_poisson_solve<float>(s, Rho_face.x.p<float>(), Rho_face.y.p<float>(),
                                    Rho_face.z.p<float>(), true, P.pointer(), RHS.pointer(),
                                    par->Iteration(), par->epsilon(), &num_iteration_, st_,
                                    rsum);
  if (sd)
    return duct_poisson_solve(sd, Rho_face.x.pointer(), Rho_face.y.pointer(),
                              Rho_face.z.pointer(), true, P.pointer(), RHS.pointer(),
                              par->Iteration(), par->epsilon(), &num_iteration_, st_);
  return box_poisson_solve<double>(s, Rho_face.x.pointer(), Rho_face.y.pointer(),
                                   Rho_face.z.pointer(), true, P.pointer(), RHS.pointer(),
                                   par->Iteration(), par->epsilon(), &num_iteration_, st_, rsum);
}
int poisson::Solve_cell_rho(tensor0& Rho_new, tensor0& P, tensor0& RHS, const double* rsum) {
  if (!s) return FR_EINVAL;
  if (Rho_new.is_f32())
    return box_poisson_solve<float>(s, Rho_new.p<float>(), nullptr, nullptr, 2, P.pointer(),
                                    RHS.pointer(), par->Iteration(), par->epsilon(),
                                    &num_iteration_, st_, rsum);
  return box_poisson_solve<double>(s, Rho_new.pointer(), nullptr, nullptr, 2, P.pointer(),
                                   RHS.pointer(), par->Iteration(), par->epsilon(),
                                   &num_iteration_, st_, rsum);
}

// ---------------------------------------------------------------- grid
#define FR_TRY(x)                \
  do {                           \
    int rc_ = (x);               \
    if (rc_) {                   \
      if (!err) err = rc_;       \
      return rc_;                \
    }                            \
  } while (0)

grid::~grid() {
  if (dred) hipFree(dred);
  if (dprof) hipFree(dprof);
  if (inflow) hipFree(inflow);
  // a grid whose setup stopped early holds null handles; destroying those
  // would leave an error behind for the next launch check
  for (auto& e : ev)
    if (e) hipEventDestroy(e);
  for (auto& e : evk)
    if (e) hipEventDestroy(e);
  if (evf) hipEventDestroy(evf);
  if (evj) hipEventDestroy(evj);
  if (evj2) hipEventDestroy(evj2);
  if (evc) hipEventDestroy(evc);
  if (evm) hipEventDestroy(evm);
  if (evm0) hipEventDestroy(evm0);
  if (evh) hipEventDestroy(evh);
  if (st3) hipStreamDestroy(st3);
  if (evq) hipEventDestroy(evq);
  if (dq) hipFree(dq);
  if (dmz) hipFree(dmz);
  if (dpart) hipFree(dpart);
  if (hq) hipHostFree(hq);
  if (st2) hipStreamDestroy(st2);
  if (st) hipStreamDestroy(st);
}

int grid::setup(const char* params_path, const char* ic) {
  param = params(params_path);
  if (!param.ok) return FR_EINVAL;
  if (param.bs() != 1) return FR_EINVAL;
  // the precision form: the box with the FFT solve (FP32 inner transform) or
  // with PCG and the FP32 multigrid preconditioner (Solver = 2)
  if (f32 && (!param.IsBox() || (param.Solver() != 0 && param.Solver() != 2))) return FR_EINVAL;
  size = gridsize(param);
  ic_dir = ic ? ic : ".";
  if (const char* e = getenv("FR_RESORT_EVERY")) resort_every = atoi(e) < 1 ? 1 : atoi(e);
  FR_CHECK(hipStreamCreate(&st));
  FR_CHECK(hipStreamCreate(&st2));
  FR_CHECK(hipStreamCreate(&st3));
  FR_CHECK(hipEventCreateWithFlags(&evm, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evm0, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evh, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evf, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evj, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evj2, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evc, hipEventDisableTiming));
  FR_CHECK(hipEventCreateWithFlags(&evq, hipEventDisableTiming));
  for (auto& e : ev) FR_CHECK(hipEventCreate(&e));
  for (auto& e : evk) FR_CHECK(hipEventCreate(&e));
  // grid.cpp:16-23
  RK4_preCoeff[0] = 0.5;
  RK4_preCoeff[1] = 0.5;
  RK4_preCoeff[2] = 1.;
  RK4_preCoeff[3] = 1. / 6.;
  RK4_postCoeff[0] = 1. / 6.;
  RK4_postCoeff[1] = 1. / 3.;
  RK4_postCoeff[2] = 1. / 3.;
  RK4_postCoeff[3] = 1. / 6.;
  for (tensor1* f : {&RU, &RU_int, &RU_new, &RU_np1, &RHS_RU, &U, &Rho_face})
    f->init(&size, st, f32);
  dummy.init(&size, st);
  if (f32) {
    d64a.init(&size, st);
    d64b.init(&size, st);
    d64r.init(&size, st);
  }
  for (tensor0* f : {&C, &Rho, &Rho_int, &Rho_new, &Rho_np1, &RHS_Part_Temp})
    f->init(&size, st, f32);
  for (tensor0* f : {&P, &RHS_Pois, &scratch, &divergence}) f->init(&size, st);
  if (!divergence.pointer()) return FR_ENOMEM;
  if (!param.IsBox()) {
    const size_t m = (size_t)size.Ny() * size.Nz();
    FR_CHECK(hipMalloc(&inflow, sizeof(double) * 8 * m));
    FR_CHECK(hipMemset(inflow, 0, sizeof(double) * 8 * m));
  }
  FR_TRY(part.init(&param, &size, st, param.N0() > 0 ? param.N0() : 1));
  FR_TRY(PS_.init(&param, &size, st));
  FR_CHECK(hipMalloc(&dred, sizeof(double) * (4 * (size_t)size.Nx() + 512)));
  FR_CHECK(hipMalloc(&dq, sizeof(double) * (RED_PARTS + 8)));
  FR_CHECK(hipMalloc(&dmz, sizeof(double) * (RED_PARTS + 16)));
  // dpart: per-workgroup sums of k_div_rhs, or of the lean last stage's
  // momentum update, one at a time on the main stream
  if (param.IsBox())
    FR_CHECK(hipMalloc(&dpart, sizeof(double) * std::max(box_div_parts(size.dev()),
                                                         box_wp_parts(size.dev()))));
  FR_CHECK(hipHostMalloc(&hq, sizeof(double) * 2, hipHostMallocDefault));
  return FR_OK;
}

static int read_block(const std::string& path, const gridsize& s, std::vector<double>& ghosted) {
  const int nx = s.Nx_tot_, ny = s.Ny_tot_, nz = s.Nz_tot_;
  std::vector<double> blk((size_t)nx * ny * nz);
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return FR_EINVAL;
  size_t got = fread(blk.data(), sizeof(double), blk.size(), f);
  fclose(f);
  if (got != blk.size()) return FR_EINVAL;
  ghosted.assign((size_t)s.size(), 0.0);
  size_t n = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        ghosted[((size_t)(k + 1) * s.Ny() + (j + 1)) * s.Nx() + (i + 1)] = blk[n++];
  return FR_OK;
}

int grid::reserve(int capacity) {
  part.dev.release();
  part.dev = PartDev();
  part.Np = 0;
  return part.dev.alloc(capacity, size.dev());
}

// grid.cpp:110-231 (Initial = 0 and 2); particles are loaded in the box only.
int grid::Initialize() {
  if (param.Initial() == 0) {
    const char* names[3] = {"U.bin", "V.bin", "W.bin"};
    std::vector<double> h;
    for (int m = 0; m < 3; m++) {
      FR_TRY(read_block(ic_dir + "/" + names[m], size, h));
      RU.c(m).upload(h.data());
      RU.c(m).Update_Ghosts();  // communicator::read -> Update_Ghosts
    }
    RU.x.add(param.Rho0() * param.U0());
  } else if (param.Initial() == 2) {
    RU.assign(0);
    RU.x.assign(param.Rho0() * param.U0());
  } else if (param.Initial() == 1) {  // grid.cpp:124-182
    FR_TRY(Restart());
    FR_TRY(Finish_Initialize());
    if (outputs) FR_TRY(Write_info());
    return FR_OK;
  } else {
    return FR_EINVAL;
  }
  Rho.assign(param.Rho0());
  P.assign(0);
  if (param.IsBox()) {
    FR_TRY(part.load_random());
  } else {
    part.Np = part.dev.np = 0;
  }
  P0 = param.P0();
  T_cur = 0;
  num_timestep = 0;
  FR_TRY(Finish_Initialize());
  if (outputs) FR_TRY(Write_info());
  return FR_OK;
}

// grid.cpp:197-229: the time-integration stages, Rho_face, the particle
// stages and the first particle heat source, and C_BC.
int grid::Finish_Initialize() {
  rho_np1_stale = ru_np1_stale = ru_lean = false;
  face_src = 0;
  face_stored = true;
  RU_int.copy(RU);
  RU_np1.copy(RU);
  Rho_int.copy(Rho);
  Rho_np1.copy(Rho);
  P0_int = P0;
  P0_np1 = P0;
  for (int m = 0; m < 3; m++) {
    if (f32)
      FR_TRY(interp_c2f(&size.g, Rho_face.c(m).p<float>(), Rho.p<float>(), m, st));
    else
      FR_TRY(interp_c2f(&size.g, Rho_face.c(m).pointer(), Rho.pointer(), m, st));
  }
  FR_TRY(part.dev.copy_stage(1, 0, st));
  FR_TRY(part.dev.copy_stage(3, 0, st));
  // T = (P0_int/R)/Rho_int; gas2part_Temp_int; Compute_RHS_Temp_int;
  // part2gas_Temp_int (grid.cpp:223-226)
  FR_TRY(part.temp(1, P0_int / param.R(), Rho_int));
  FR_TRY(part.rebuild_cells(1));
  FR_TRY(part.part2gas_Temp(1, RHS_Part_Temp));
  qmean_valid = false;
  U.x.Equal_Divide(RU.x, Rho_face.x);
  U.y.Equal_Divide(RU.y, Rho_face.y);
  U.z.Equal_Divide(RU.z, Rho_face.z);
  C_BC = U.x.max();
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

int grid::set_inflow(const double* planes, int n, const double* particles) {
  if (param.IsBox() || n < 0) return FR_EINVAL;
  if (planes) {
    const size_t n = 8 * (size_t)size.Ny() * size.Nz();
    FR_CHECK(hipMemcpy(inflow, planes, sizeof(double) * n, hipMemcpyHostToDevice));
    inflow_h.assign(planes, planes + n);
  }
  if (n > 0 && particles) inlet.insert(inlet.end(), particles, particles + (size_t)n * 29);
  return FR_OK;
}

static BoxCoef box_coef(const grid& g) {
  BoxCoef c;
  const double dt = g.param.dt(), pre = g.RK4_preCoeff[g.RK4_count],
               post = g.RK4_postCoeff[g.RK4_count];
  c.cpost_neg = -(dt * post);
  c.cpre_neg = -dt * pre;
  c.cpost = dt * post;
  c.cpre = dt * pre;
  c.last = g.RK4_count == 3;
  return c;
}

// grid.cpp:327-333; the same kernel also forms U = RU_int/Rho_face, the first
// statement of grid::Update_Particle (grid.cpp:662), from unchanged inputs.
int grid::Update_Rho() {
  const double* ri[3] = {RU_int.x.pointer(), RU_int.y.pointer(), RU_int.z.pointer()};
  const double* rf[3] = {Rho_face.x.pointer(), Rho_face.y.pointer(), Rho_face.z.pointer()};
  double* u[3] = {U.x.pointer(), U.y.pointer(), U.z.pointer()};
  if (!param.IsBox()) {  // grid.cpp:335-385
    const size_t m = (size_t)size.Ny() * size.Nz();
    FR_TRY(duct_update_rho(size.dev(), box_coef(*this), C_BC, ri, Rho.pointer(),
                           Rho_int.pointer(), Rho_np1.pointer(), Rho_new.pointer(), rf, u,
                           inflow, inflow + m, st));
    return FR_OK;
  }
  rho_set = in_step && RK4_count == 3;
  // Rho_new is rewritten: faces formed from it must be stored first; faces
  // formed from Rho_int are formed here from Rho_int
  if (face_src == 1) {
    FR_TRY(fill_faces());
    face_src = 0;
  }
  const bool fi = face_src == 2;
  if (!fi) FR_TRY(fill_faces());
  BoxCoef cf = box_coef(*this);
  if (rho_np1_stale && RK4_count == 0)
    cf.np1_base = 1;  // Rho_np1 = Rho: read Rho in its place
  else
    FR_TRY(fill_rho_np1());
  cf.lean = lean_stage();
  if (f32) {
    const float* ri_[3] = {RU_int.x.p<float>(), RU_int.y.p<float>(), RU_int.z.p<float>()};
    const float* rf_[3] = {Rho_face.x.p<float>(), Rho_face.y.p<float>(), Rho_face.z.p<float>()};
    float* u_[3] = {U.x.p<float>(), U.y.p<float>(), U.z.p<float>()};
    FR_TRY(box_update_rho<float>(size.dev(), cf, ri_, Rho.p<float>(), Rho_np1.p<float>(),
                                 Rho_new.p<float>(), rf_, u_, st,
                                 rho_set ? Rho.p<float>() : nullptr,
                                 fi ? Rho_int.p<float>() : nullptr));
  } else {
    FR_TRY(box_update_rho<double>(size.dev(), cf, ri, Rho.pointer(), Rho_np1.pointer(),
                                  Rho_new.pointer(), rf, u, st,
                                  rho_set ? Rho.pointer() : nullptr,
                                  fi ? Rho_int.pointer() : nullptr));
  }
  rho_np1_stale = cf.lean != 0;  // written unless lean (then equal to Rho)
  return FR_OK;
}

int grid::fill_np1() {
  FR_TRY(fill_rho_np1());
  return fill_ru_np1();
}
int grid::fill_rho_np1() {
  if (rho_np1_stale) Rho_np1.copy(Rho);
  rho_np1_stale = false;
  return FR_OK;
}
int grid::fill_ru_np1() {
  if (ru_lean) return FR_EINVAL;  // inside a full step's last stage only
  if (ru_np1_stale) RU_np1.copy(RU);
  ru_np1_stale = false;
  return FR_OK;
}

int grid::fill_faces() {
  if (face_stored) return FR_OK;
  if (face_src != 1 && face_src != 2) return FR_EINVAL;
  tensor0& src = face_src == 1 ? Rho_new : Rho_int;
  if (f32) {
    float* rf[3] = {Rho_face.x.p<float>(), Rho_face.y.p<float>(), Rho_face.z.p<float>()};
    FR_TRY(box_faces_of<float>(size.dev(), src.p<float>(), rf, st));
  } else {
    double* rf[3] = {Rho_face.x.pointer(), Rho_face.y.pointer(), Rho_face.z.pointer()};
    FR_TRY(box_faces_of<double>(size.dev(), src.pointer(), rf, st));
  }
  face_stored = true;
  return FR_OK;
}

int grid::Update_P0() {  // grid.cpp:539-554
  FR_TRY(settle_side());
  if (!qmean_valid) {
    qmean = RHS_Part_Temp.mean();
    qmean_valid = true;
  }
  mean_energy_transferred = qmean;
  if (param.IsBox()) {
    if (param.Is_Cooling())
      dP0_dt = 0;
    else
      dP0_dt = -param.R() / param.Cv() * mean_energy_transferred;
    P0_np1 += param.dt() * RK4_postCoeff[RK4_count] * dP0_dt;
    if (RK4_count != 3)
      P0_new = P0 + param.dt() * RK4_preCoeff[RK4_count] * dP0_dt;
    else
      P0_new = P0_np1;
  } else {
    P0_new = P0;
    P0_np1 = P0;
  }
  return FR_OK;
}

int grid::Update_Particle() {  // grid.cpp:659-680
  hipEventRecord(ev[2], st);
  // storage back in cell order every resort_every steps (reference order is
  // kept in the particles' ranks, so no result depends on it)
  if (RK4_count == 0 && num_timestep % resort_every == 0)
    FR_TRY(part.dev.resort(size.dev(), st));
  // the coupled box's '>hi' rows carry stage 0 as the reference has it before
  // TimeAdvance, so the copy stays in TimeAdvance there
  part_set = in_step && param.IsBox() && !pair_out && RK4_count == 3;
  // the first stage of a step after a full step: slots 1 and 3 hold slot 0
  const bool eq = in_step && RK4_count == 0 && stages_equal;
  stages_equal = false;
  FR_TRY(part.advance(RK4_count, RK4_preCoeff[RK4_count], RK4_postCoeff[RK4_count], U,
                      part_set, eq));
  // The concentration deposit (rounded-index buckets, output C) and the
  // momentum deposit (truncating-index buckets, output RHS_RU) read the same
  // particles and write disjoint data: both are latency-bound, so they run
  // side by side on two streams.  In the box the migration (new-stage
  // positions and reference indices, which neither deposit reads) runs on a
  // third stream beside them; the momentum deposit orders its buckets by a
  // copy of the reference indices taken before the migration starts, which
  // is the order of the reference, whose deposits precede Send_Recv.
  FR_CHECK(hipEventRecord(evf, st));
  FR_CHECK(hipStreamWaitEvent(st2, evf, 0));
  const bool front = step_front;
  step_front = false;
  if (!front) FR_TRY(part.part2gas_concentration(C, st2));
  const bool box = param.IsBox();
  const int* rk = nullptr;
  if (box) {
    if (param.TWC() && part.dev.np > 0) {
      FR_CHECK(hipMemcpyAsync(part.dev.rank_dep, part.dev.rank, sizeof(int) * part.dev.np,
                              hipMemcpyDeviceToDevice, st));
      rk = part.dev.rank_dep;
    }
    FR_CHECK(hipEventRecord(evm0, st));
  }
  // the momentum deposit is the longer branch: its launches go first
  if (param.TWC())
    FR_TRY(part.part2gas_velocity(RHS_RU, rk, !front, vel_head ? evh : nullptr));
  else
    RHS_RU.assign(0);
  vel_head = false;
  // Heat branch on st2 (gas temperature at the particles, buckets of the new
  // positions, heat deposit, its sum): it follows the concentration deposit
  // on st2 and needs the migrated positions and reference indices.  Nothing
  // on st reads it before join_side() (Update_RU_WOP reads none of it).
  if (box) {
    FR_CHECK(hipStreamWaitEvent(st3, evm0, 0));
    FR_TRY(part.Send_Recv(st3));
    if (pair_out) {  // the coupled duct's Rbuf1 (particle.cpp:836-838)
      PairLink& l = *pair_out;
      FR_CHECK(hipStreamWaitEvent(st3, l.ev_duct, 0));
      FR_TRY(part.dev.gather_hi(l.buf, l.cap, l.d_n, st3));
      FR_CHECK(hipMemcpyAsync(l.h_n, l.d_n, sizeof(int), hipMemcpyDeviceToHost, st3));
      FR_CHECK(hipEventRecord(l.ev_rows, st3));
    }
    FR_CHECK(hipEventRecord(evm, st3));
    // box: it starts as soon as the migration is done, beside the momentum
    // deposit's tail and the momentum update
    FR_CHECK(hipStreamWaitEvent(st2, evm, 0));
    side_held = true;
    FR_TRY(launch_side());
  } else {
    // duct: the migration (outflow drop, inlet append) reorders the storage,
    // so it follows both deposits on st; the branch's launches are issued
    // after the momentum kernel's (launch_side from Update_RU_WOP), so that
    // kernel is queued first
    FR_CHECK(hipEventRecord(evj, st2));
    FR_CHECK(hipStreamWaitEvent(st, evj, 0));
    if (pair_in) {  // the coupled box's '>hi' list of this substep
      FR_CHECK(hipEventSynchronize(pair_in->ev_rows));
      const int n = *pair_in->h_n;
      if (n > pair_in->cap) {
        fprintf(stderr, "fr_amd: %d box particles enter the duct, capacity %d\n", n,
                pair_in->cap);
        return FR_EOVERFLOW;
      }
      FR_TRY(part.Send_Recv_duct_dev(pair_in->buf, n));
    } else {
      FR_TRY(part.Send_Recv_duct(inlet.data(), (int)(inlet.size() / 29)));
      inlet.clear();
    }
    FR_CHECK(hipEventRecord(evf, st));
    FR_CHECK(hipStreamWaitEvent(st2, evf, 0));
    side_held = true;
  }
  side_pending = true;
  qmean_valid = false;
  return FR_OK;
}

// Issues the held heat branch on st2 (its stream waits are already queued).
int grid::launch_side() {
  if (!side_held) return FR_OK;
  side_held = false;
  FR_TRY(part.temp(2, P0_new / param.R(), Rho_new, st2));
  FR_TRY(part.rebuild_cells(2, st2));
  // the heat sum (Update_P0's mean) from the deposit's own partials where
  // its form provides them, else a reduction of the folded field
  bool summed = false;
  FR_TRY(part.part2gas_Temp(2, RHS_Part_Temp, st2, dq + RED_PARTS, &summed));
  FR_CHECK(hipEventRecord(ev[3], st2));
  if (!summed && RHS_Part_Temp.is_f32())
    FR_TRY(reduce_async_ws(size.dev(), RHS_Part_Temp.p<float>(), 0, dq + RED_PARTS, dq, st2));
  else if (!summed)
    FR_TRY(reduce_async_ws(size.dev(), RHS_Part_Temp.pointer(), 0, dq + RED_PARTS, dq, st2));
  FR_CHECK(hipEventRecord(evj2, st2));
  return FR_OK;
}

// Joins the heat branch into the main stream (a stream wait: the host does
// not wait) and starts the copy of its sum to pinned host memory.
int grid::join_side() {
  if (!side_pending) return FR_OK;
  FR_TRY(launch_side());
  side_pending = false;
  FR_CHECK(hipStreamWaitEvent(st, evj2, 0));
  FR_CHECK(hipMemcpyAsync(hq, dq + RED_PARTS, sizeof(double), hipMemcpyDeviceToHost, st));
  FR_CHECK(hipEventRecord(evq, st));
  q_inflight = true;
  return FR_OK;
}

// join_side(), then the mean of RHS_Part_Temp on the host (the tensor0::mean
// of grid.cpp:677) and the particle time.  Its first reader is the next
// stage's Update_P0, after the pressure solve has synchronised the stream,
// so the copy has normally landed.
int grid::settle_side() {
  int rc = join_side();
  if (rc) return rc;
  if (!q_inflight) return FR_OK;
  q_inflight = false;
  FR_CHECK(hipEventSynchronize(evq));
  qmean = hq[0] / ((double)(size.Nx() - 2) * (size.Ny() - 2) * (size.Nz() - 2));
  qmean_valid = true;
  float ms = 0;
  if (hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) t_particles += ms;
  return FR_OK;
}

int grid::Update_RU_WOP() {  // grid.cpp:387-425
  WopCoef w;
  w.a_s = -param.mp() / size.Vcell();
  w.b_g = param.Mu0() / 3.;
  w.c_l = param.Mu0();
  w.A = param.A();
  w.sx_forc = param.A() * param.Rho0() * param.U0();
  w.rho0 = param.Rho0();
  w.gx = param.gx();
  w.gy = param.gy();
  w.gz = param.gz();
  w.gravity = (w.gx != 0 || w.gy != 0 || w.gz != 0);
  const double* u[3] = {U.x.pointer(), U.y.pointer(), U.z.pointer()};
  const double* ri[3] = {RU_int.x.pointer(), RU_int.y.pointer(), RU_int.z.pointer()};
  const double* S[3] = {RHS_RU.x.pointer(), RHS_RU.y.pointer(), RHS_RU.z.pointer()};
  const double* r0[3] = {RU.x.pointer(), RU.y.pointer(), RU.z.pointer()};
  double* n1[3] = {RU_np1.x.pointer(), RU_np1.y.pointer(), RU_np1.z.pointer()};
  double* nw[3] = {RU_new.x.pointer(), RU_new.y.pointer(), RU_new.z.pointer()};
  const double* rf[3] = {Rho_face.x.pointer(), Rho_face.y.pointer(), Rho_face.z.pointer()};
  if (w.gravity) FR_TRY(fill_faces());
  BoxCoef cf = box_coef(*this);
  if (param.IsBox() && ru_np1_stale && RK4_count == 0)
    cf.np1_base = 1;  // RU_np1 = RU: read RU in its place
  else
    FR_TRY(fill_ru_np1());
  if (param.IsBox()) cf.lean = lean_stage();
  hipEventRecord(evk[2 * (RK4_count & 3)], st);
  if (param.IsBox() && f32) {
    auto P3 = [](tensor1& t) { return std::array<float*, 3>{t.x.p<float>(), t.y.p<float>(), t.z.p<float>()}; };
    auto a = P3(U), b = P3(RU_int), c = P3(RHS_RU), e = P3(RU), f = P3(RU_np1), g = P3(RU_new),
         h = P3(Rho_face);
    const float* u_[3] = {a[0], a[1], a[2]};
    const float* ri_[3] = {b[0], b[1], b[2]};
    const float* S_[3] = {c[0], c[1], c[2]};
    const float* r0_[3] = {e[0], e[1], e[2]};
    float* n1_[3] = {f[0], f[1], f[2]};
    float* nw_[3] = {g[0], g[1], g[2]};
    const float* rf_[3] = {h[0], h[1], h[2]};
    FR_TRY(box_update_ru_wop<float>(size.dev(), cf, w, u_, ri_, S_, r0_, n1_, nw_, rf_, st));
  } else if (param.IsBox()) {
    FR_TRY(box_update_ru_wop<double>(size.dev(), cf, w, u, ri, S, r0, n1, nw, rf, st));
  } else {  // grid.cpp:427-537
    const size_t m = (size_t)size.Ny() * size.Nz();
    FR_TRY(duct_update_ru_wop(size.dev(), box_coef(*this), w, C_BC, u, ri, S, r0, n1, nw,
                              inflow + 2 * m, inflow + 5 * m, st));
  }
  hipEventRecord(evk[2 * (RK4_count & 3) + 1], st);
  ru_np1_stale = false;
  ru_lean = cf.lean != 0;
  // the particle heat branch, held by Update_Particle, now runs beside it
  FR_TRY(launch_side());
  return FR_OK;
}

// grid.cpp:556-563 is evaluated inside the Compute_RHS_Pois kernel, where its
// only consumer is; this call has nothing to launch.
static DivCoef div_coef(grid& g) {
  DivCoef q;
  q.kc = g.param.k() / g.param.Cp();
  q.bq = -g.param.R() / (g.param.Cp() * g.P0_new * g.size.Vcell());
  q.cooling = g.param.Is_Cooling();
  if (q.cooling) g.settle_side();
  if (q.cooling && !g.qmean_valid) {
    g.qmean = g.RHS_Part_Temp.mean();
    g.qmean_valid = true;
  }
  q.qmean = g.qmean;
  q.s = (1. / (g.param.dt() * g.RK4_preCoeff[g.RK4_count]));
  return q;
}

// Box: grid.cpp:556-563 is evaluated inside the Compute_RHS_Pois kernel,
// where its only consumer is.  Duct: the divergence field is stored (its sum
// enters Correct_OutFlow) and Rho_face = I_C2F(Rho_new) of grid.cpp:582 is
// formed in the same pass (nothing reads Rho_face in between).
int grid::Compute_Div_U_new() {
  FR_TRY(join_side());
  if (param.IsBox()) return FR_OK;
  double* rf[3] = {Rho_face.x.pointer(), Rho_face.y.pointer(), Rho_face.z.pointer()};
  FR_TRY(duct_div(size.dev(), div_coef(*this), Rho_new.pointer(), RHS_Part_Temp.pointer(), rf,
                  divergence.pointer(), st));
  return FR_OK;
}

// grid.cpp:579-627 for a device that holds inlet and outlet: mass balance of
// the interior divergence and the inlet/outlet x-face fluxes, outlet RU_new.x
// rescaled, C_BC = mean outlet velocity.
int grid::Correct_OutFlow() {
  if (param.IsBox()) return FR_OK;
  const Dev d = size.dev();
  FR_TRY(reduce_async(d, divergence.pointer(), 0, dred, st));
  FR_TRY(duct_planes(d, RU_new.x.pointer(), Rho_face.x.pointer(), dred + 1, st));
  double h[3];
  FR_CHECK(hipMemcpyAsync(h, dred, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  const double sum_div = h[0] * size.dx();
  const double sum_out = h[1], sum_in = h[2];
  const double ratio = (sum_div + sum_in) / sum_out;
  U_avg_in = sum_in / (size.Ny_tot_ * size.Nz_tot_);
  U_avg_out = sum_out / (size.Ny_tot_ * size.Nz_tot_);
  C_BC = U_avg_out;
  FR_TRY(duct_scale(d, ratio, RK4_count == 3, RU_new.x.pointer(), RU_np1.x.pointer(), st));
  return FR_OK;
}

int grid::Compute_RHS_Pois() {  // grid.cpp:556-563 + 629-637
  FR_TRY(join_side());
  const DivCoef q = div_coef(*this);
  const double* rn[3] = {RU_new.x.pointer(), RU_new.y.pointer(), RU_new.z.pointer()};
  double* rf[3] = {Rho_face.x.pointer(), Rho_face.y.pointer(), Rho_face.z.pointer()};
  rhs_mean_pending = false;
  // sum and minus the mean, in slots no other mean's reduction writes
  double* rsum = dmz + RED_PARTS + 4;
  // in a full step with the FFT solve and no gravity the faces stay unstored
  const bool keep = !(param.IsBox() && in_step && PS_.s && param.gx() == 0 &&
                      param.gy() == 0 && param.gz() == 0);
  if (param.IsBox() && !keep) rf[0] = nullptr;
  if (param.IsBox() && f32) {
    const float* rn_[3] = {RU_new.x.p<float>(), RU_new.y.p<float>(), RU_new.z.p<float>()};
    float* rf_[3] = {keep ? Rho_face.x.p<float>() : nullptr, Rho_face.y.p<float>(),
                     Rho_face.z.p<float>()};
    FR_TRY(box_div_rhs<float>(size.dev(), q, Rho_new.p<float>(), RHS_Part_Temp.p<float>(), rn_,
                              rf_, RHS_Pois.pointer(), st, dpart, rsum));
  } else if (param.IsBox()) {
    FR_TRY(box_div_rhs<double>(size.dev(), q, Rho_new.pointer(), RHS_Part_Temp.pointer(), rn,
                               rf, RHS_Pois.pointer(), st, dpart, rsum));
  } else {
    const double* rfc[3] = {rf[0], rf[1], rf[2]};
    FR_TRY(duct_rhs(size.dev(), q.s, rn, rfc, divergence.pointer(), RHS_Pois.pointer(), st));
    RHS_Pois.make_mean_zero_async(dmz);
    return FR_OK;
  }
  face_src = 1;
  face_stored = keep;
  // box: the interior sum was formed by the kernel (tensor0::make_mean_zero).
  // In the time step with the FFT solve the field keeps its mean: the
  // solve's residual, the field's only reader there, adds the same shift
  // itself, and the field takes it only when read (fill_rhs_pois).
  rhs_mean_pending = true;
  if (!(in_step && PS_.s)) return fill_rhs_pois();
  return FR_OK;
}

// RHS_Pois with its mean removed, as tensor0::make_mean_zero leaves it.
int grid::fill_rhs_pois() {
  if (!rhs_mean_pending) return FR_OK;
  rhs_mean_pending = false;
  return add_neg_mean_async(size.dev(), RHS_Pois.pointer(), (long)size.size(),
                            dmz + RED_PARTS + 4, st);
}

int grid::Solve_Poisson() {  // grid.cpp:639-643
  // C = 1./Rho_face is formed inside the solver's residual
  hipEventRecord(ev[0], st);
  if (rhs_mean_pending && !PS_.s) FR_TRY(fill_rhs_pois());
  const double* rsum = rhs_mean_pending ? dmz + RED_PARTS + 4 : nullptr;
  if (f32 && PS_.sp) {
    FR_TRY(fill_faces());
    // the PCG's FP64 operator reads the FP32 faces widened (the FFT solve
    // widens them inside its residual); its V-cycles run in FP32 either way
    for (int m = 0; m < 3; m++)
      to_double(Rho_face.c(m).p<float>(), d64b.c(m).pointer(), size.size(), st);
    FR_TRY(PS_.Solve_face_rho(d64b, P, RHS_Pois));
  } else if (face_src == 1 && PS_.s) {
    FR_TRY(PS_.Solve_cell_rho(Rho_new, P, RHS_Pois, rsum));
  } else {
    FR_TRY(fill_faces());
    FR_TRY(PS_.Solve_face_rho(Rho_face, P, RHS_Pois, rsum));
  }
  hipEventRecord(ev[1], st);
  hipEventSynchronize(ev[1]);
  float ms = 0;
  hipEventElapsedTime(&ms, ev[0], ev[1]);
  t_poisson += ms;
  iters_step += PS_.num_iteration();
  return FR_OK;
}

int grid::Update_RU_WP() {  // grid.cpp:651-657
  double* n1[3] = {RU_np1.x.pointer(), RU_np1.y.pointer(), RU_np1.z.pointer()};
  double* nw[3] = {RU_new.x.pointer(), RU_new.y.pointer(), RU_new.z.pointer()};
  BoxCoef cf = box_coef(*this);
  if (ru_lean && RK4_count == 3)
    cf.lean = 1;  // RU_new holds RU_np1; the means' sums go to dmz + RED_PARTS + 8
  else if (ru_lean)
    return FR_EINVAL;
  else
    FR_TRY(fill_ru_np1());
  double* sums = dmz + RED_PARTS + 8;
  if (param.IsBox() && f32) {
    float* n1_[3] = {RU_np1.x.p<float>(), RU_np1.y.p<float>(), RU_np1.z.p<float>()};
    float* nw_[3] = {RU_new.x.p<float>(), RU_new.y.p<float>(), RU_new.z.p<float>()};
    FR_TRY(box_update_ru_wp<float>(size.dev(), cf, P.pointer(), n1_, nw_, st, dpart, sums));
  } else if (param.IsBox())
    FR_TRY(box_update_ru_wp<double>(size.dev(), cf, P.pointer(), n1, nw, st, dpart, sums));
  else
    FR_TRY(duct_update_ru_wp(size.dev(), box_coef(*this), P.pointer(), n1, nw, st));
  return FR_OK;
}

// grid.cpp:705-712.  Rho_new and RU_new are rewritten whole before their
// next read, so exchanging buffers equals the copies.
int grid::TimeAdvance_RK4() {
  Rho_int.swap(Rho_new);
  if (face_src) face_src = 3 - face_src;  // the faces' source moved with the buffer
  RU_int.swap(RU_new);
  P0_int = P0_new;
  part.dev.swap_stages(1, 2);
  return FR_OK;
}

int grid::Correct_mean_RU() {  // grid.cpp:565-577
  if (ru_lean) {
    // RU_int holds the last stage's RU_np1 after the stage exchange; RU takes
    // it with the means removed, and RU_np1 equals RU from here on
    const Dev d = size.dev();
    const double add = param.Rho0() * param.U0();
    const double* sums = dmz + RED_PARTS + 8;
    if (f32) {
      float* r[3] = {RU.x.p<float>(), RU.y.p<float>(), RU.z.p<float>()};
      const float* s[3] = {RU_int.x.p<float>(), RU_int.y.p<float>(), RU_int.z.p<float>()};
      FR_TRY(box_set_mean_free<float>(d, r, s, sums, add, st));
    } else {
      double* r[3] = {RU.x.pointer(), RU.y.pointer(), RU.z.pointer()};
      const double* s[3] = {RU_int.x.pointer(), RU_int.y.pointer(), RU_int.z.pointer()};
      FR_TRY(box_set_mean_free<double>(d, r, s, sums, add, st));
    }
    ru_lean = false;
    ru_np1_stale = true;
    ru_set = true;
    return FR_OK;
  }
  FR_TRY(fill_np1());
  if (param.IsBox() && in_step) {
    // with RU = RU_np1 of TimeAdvance in the same pass
    const Dev d = size.dev();
    const long n = size.size();
    for (int m = 0; m < 3; m++) {
      tensor0 &f = RU_np1.c(m), &g = RU.c(m);
      const double add = param.Rho0() * param.U0();
      if (f.is_f32())
        FR_TRY(make_mean_zero_set_async(d, f.p<float>(), g.p<float>(), n, dmz, m == 0, add, st));
      else
        FR_TRY(make_mean_zero_set_async(d, f.p<double>(), g.p<double>(), n, dmz, m == 0, add,
                                        st));
    }
    ru_set = true;
  } else if (param.IsBox()) {
    RU_np1.make_mean_zero_async(dmz);
    RU_np1.x.add(param.Rho0() * param.U0());
  } else {
    RU_np1.y.make_mean_zero_async(dmz);
    RU_np1.z.make_mean_zero_async(dmz);
  }
  return FR_OK;
}

int grid::TimeAdvance() {  // grid.cpp:315-325
  if (!rho_set || !ru_set) FR_TRY(fill_np1());
  if (!rho_set) Rho.copy(Rho_np1);
  if (!ru_set) RU.copy(RU_np1);
  P0 = P0_np1;
  if (!part_set) FR_TRY(part.dev.copy_stage(0, 3, st));
  // a full step's last stage wrote its result to slots 0, 2 and 3, the
  // migration wrapped them alike and the stage exchange made slot 1 slot 2
  stages_equal = part_set;
  rho_set = ru_set = part_set = false;
  T_cur += param.dt();
  num_timestep++;
  return FR_OK;
}

int grid::begin_step() {
  param.update(T_cur);
  in_step = true;
  rho_set = ru_set = part_set = false;
  step_front = false;
  t_poisson = t_particles = 0;
  iters_step = 0;
  box_poisson_xtimer(PS_.s, nullptr, nullptr, true);
  part.dev.t_vel.collect();  // launches outside a step are not counted
  part.dev.t_cic.collect();
  part.dev.t_vel.reset();
  part.dev.t_cic.reset();
  t0_step = std::chrono::steady_clock::now();
  return FR_OK;
}

int grid::substep_front(int k) {
  RK4_count = k;
  if (param.IsBox()) {
    FR_CHECK(hipEventRecord(evc, st));
    FR_CHECK(hipStreamWaitEvent(st2, evc, 0));
    FR_TRY(part.part2gas_concentration(C, st2));
    step_front = true;
    // The momentum deposit's buckets are of the stage-1 positions, which
    // this substep does not change before the deposit (the advance writes
    // the other stages; the storage re-sort, when due here, reorders, so
    // then the whole build waits for it): count, scan and unsorted lists go
    // beside the density update and the advance on st3, ahead of the
    // migration there; the placement, which takes the advance's forces,
    // follows the advance on st.
    const bool resort_now = k == 0 && num_timestep % resort_every == 0;
    vel_head = param.TWC() && part.dev.np > 0 && !resort_now;
    if (vel_head) {
      FR_CHECK(hipStreamWaitEvent(st3, evc, 0));
      FR_TRY(part.dev.build_head(size.dev(), 1, 1, st3));
      FR_CHECK(hipEventRecord(evh, st3));
    }
  }
  FR_TRY(Update_Rho());
  FR_TRY(Update_P0());
  FR_TRY(Update_Particle());
  FR_TRY(Update_RU_WOP());
  FR_TRY(Compute_Div_U_new());
  FR_TRY(Correct_OutFlow());  // duct only
  FR_TRY(Compute_RHS_Pois());
  FR_TRY(Solve_Poisson());
  return FR_OK;
}

int grid::step() {  // simulation.cpp:36-70 (box calls)
  FR_TRY(begin_step());
  for (int k = 0; k < 4; k++) {
    FR_TRY(substep_front(k));
    FR_TRY(Update_RU_WP());
    FR_TRY(TimeAdvance_RK4());
  }
  FR_TRY(Correct_mean_RU());
  FR_TRY(TimeAdvance());
  return end_step();
}

int grid::end_step() {
  in_step = false;
  FR_CHECK(hipStreamSynchronize(st));
  FR_TRY(settle_side());
  t_step = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_step)
               .count();
  t_momentum = 0;
  for (int k = 0; k < 4; k++) {
    float ms = 0;
    hipEventElapsedTime(&ms, evk[2 * k], evk[2 * k + 1]);
    t_momentum += ms;
  }
  n_momentum = 4;
  box_poisson_xtimer(PS_.s, &t_xpass, &n_xpass, false);
  part.dev.t_vel.collect();
  part.dev.t_cic.collect();
  return FR_OK;
}

// grid.cpp:714-811 subset; see fr_sim_diagnostics.
int grid::diagnostics(double* out) {
  FR_TRY(fill_np1());
  FR_TRY(fill_faces());
  const Dev d = size.dev();
  out[1] = Rho.mean();
  out[2] = Rho.max();
  out[3] = Rho.min();
  out[4] = C.max();
  out[5] = C.min();
  out[6] = C.mean();
  double* rred = dred;
  double* ured = dred + size.Nx();
  double* prof = dred + 2 * size.Nx();
  double* psum = dred + 4 * size.Nx();
  // The FP32 form widens its fields into FP64 copies; the sums are the same.
  const long n = size.size();
  auto wide0 = [&](tensor0& f, tensor0& tmp) -> tensor0* {
    if (!f32) return &f;
    to_double(f.p<float>(), tmp.pointer(), n, st);
    return &tmp;
  };
  auto wide1 = [&](tensor1& f, tensor1& tmp) -> tensor1* {
    if (!f32) return &f;
    for (int m = 0; m < 3; m++) to_double(f.c(m).p<float>(), tmp.c(m).pointer(), n, st);
    return &tmp;
  };
  tensor0* rho = wide0(Rho, d64r);
  tensor1* ru = wide1(RU, d64a);
  FR_TRY(reducex(d, rho->pointer(), rred, st));
  double tke = 0;
  for (int m = 0; m < 3; m++) {
    scratch.Equal_Divide(ru->c(m), *rho);  // dummy2 = RU/Rho (grid.cpp:738)
    FR_TRY(reducex(d, scratch.pointer(), ured, st));
    FR_TRY(diag_tke(d, scratch.pointer(), rho->pointer(), ured, scratch.pointer(), st));
    tke = tke + scratch.mean();
  }
  out[0] = tke;
  // U = RU_np1/Rho_face (grid.cpp:723); dissipation profile (794-811)
  tensor1* rn = wide1(RU_np1, d64a);
  tensor1* rf = wide1(Rho_face, d64b);
  for (int m = 0; m < 3; m++) dummy.c(m).Equal_Divide(rn->c(m), rf->c(m));
  const double* uf[3] = {dummy.x.pointer(), dummy.y.pointer(), dummy.z.pointer()};
  FR_TRY(diag_diss(d, uf, rho->pointer(), param.Mu0(), scratch.pointer(), st));
  FR_TRY(reducex(d, scratch.pointer(), prof, st));
  std::vector<double> h(size.Nx());
  FR_CHECK(hipMemcpyAsync(h.data(), prof, sizeof(double) * size.Nx(), hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  double s = 0;
  for (int i = 1; i < size.Nx() - 1; i++) s += h[i];
  out[7] = s / size.Nx_tot_;
  double ts = 0;
  if (part.dev.np > 0) FR_TRY(vsum(part.dev.s[6][0], part.dev.np, psum, &ts, st));
  out[8] = part.dev.np > 0 ? ts / part.dev.np : 0.0;
  out[9] = PS_.num_iteration();
  return FR_OK;
}

tensor0* grid::field0(const std::string& n) {
  if (fill_np1() || fill_faces()) return nullptr;
  face_src = 0;  // the caller may change the fields
  if (n == "P") return &P;
  if (n == "RHS_Pois") {
    (void)fill_rhs_pois();
    return &RHS_Pois;
  }
  if (n == "C") return &C;
  if (n == "Rho") return &Rho;
  if (n == "Rho_int") return &Rho_int;
  if (n == "Rho_new") return &Rho_new;
  if (n == "Rho_np1") return &Rho_np1;
  if (n == "RHS_Part_Temp") return &RHS_Part_Temp;
  if (n == "divergence") return &divergence;
  return nullptr;
}

tensor1* grid::field1(const std::string& n) {
  if (fill_np1() || fill_faces()) return nullptr;
  face_src = 0;
  if (n == "RU") return &RU;
  if (n == "RU_int") return &RU_int;
  if (n == "RU_new") return &RU_new;
  if (n == "RU_np1") return &RU_np1;
  if (n == "RHS_RU") return &RHS_RU;
  if (n == "U") return &U;
  if (n == "Rho_face") return &Rho_face;
  return nullptr;
}

// ---------------------------------------------------------------- C ABI
struct fr_sim {
  grid g;
};

extern "C" {

int fr_sim_create(const char* params_path, const char* ic_dir, fr_sim** out) {
  return fr_sim_create_ex(params_path, ic_dir, 0, out);
}

int fr_sim_create_ex(const char* params_path, const char* ic_dir, int flags, fr_sim** out) {
  if (!params_path || !out || (flags & ~(FR_SIM_FP32 | FR_SIM_OUTPUT))) return FR_EINVAL;
  fr_sim* s = new fr_sim();
  s->g.f32 = (flags & FR_SIM_FP32) != 0;
  s->g.outputs = (flags & FR_SIM_OUTPUT) != 0;
  int rc = s->g.setup(params_path, ic_dir);
  if (!rc && s->g.outputs) rc = s->g.open_outputs();
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return FR_OK;
}

int fr_sim_destroy(fr_sim* s) {
  delete s;
  return FR_OK;
}

int fr_sim_initialize(fr_sim* s) { return s ? s->g.Initialize() : FR_EINVAL; }

// a grid of a coupled run steps only through fr_pair_step
static bool coupled(const fr_sim* s) { return s->g.pair_in || s->g.pair_out; }

int fr_sim_step(fr_sim* s) { return s && !coupled(s) ? s->g.step() : FR_EINVAL; }

#define FR_TRY_S(x)     \
  do {                  \
    int rc_ = (x);      \
    if (rc_) return rc_; \
  } while (0)

int fr_sim_call(fr_sim* s, const char* m, int k) {
  if (!s || !m || coupled(s)) return FR_EINVAL;
  grid& g = s->g;
  std::string n(m);
  g.RK4_count = k;
  g.in_step = false;  // methods one by one: every copy where the reference makes it
  int rc;
  if (n == "update") {
    g.param.update(g.T_cur);
    rc = 0;
  } else if (n == "Update_Rho") rc = g.Update_Rho();
  else if (n == "Update_P0") rc = g.Update_P0();
  else if (n == "Update_Particle") rc = g.Update_Particle();
  else if (n == "Update_RU_WOP") rc = g.Update_RU_WOP();
  else if (n == "Compute_Div_U_new") rc = g.Compute_Div_U_new();
  else if (n == "Compute_RHS_Pois") rc = g.Compute_RHS_Pois();
  else if (n == "Solve_Poisson" || n == "Solve_Poisson_IO") rc = g.Solve_Poisson();
  else if (n == "Correct_OutFlow") rc = g.Correct_OutFlow();
  else if (n == "Finish_Initialize") rc = g.Finish_Initialize();
  else if (n == "Update_RU_WP") rc = g.Update_RU_WP();
  else if (n == "TimeAdvance_RK4") rc = g.TimeAdvance_RK4();
  else if (n == "Correct_mean_RU") rc = g.Correct_mean_RU();
  else if (n == "TimeAdvance") rc = g.TimeAdvance();
  else if (n == "Send_Recv") {
    if (g.param.IsBox()) {
      rc = g.part.Send_Recv();
    } else {
      rc = g.part.Send_Recv_duct(g.inlet.data(), (int)(g.inlet.size() / 29));
      g.inlet.clear();
    }
  }
  else return FR_EINVAL;
  if (!rc) rc = g.settle_side();
  if (!rc) rc = (int)hipStreamSynchronize(g.st);
  return rc;
}

static tensor0* pick(fr_sim* s, const char* name, int comp) {
  if (comp < 0) return s->g.field0(name);
  tensor1* t = s->g.field1(name);
  return (t && comp < 3) ? &t->c(comp) : nullptr;
}

int fr_sim_get_field(fr_sim* s, const char* name, int comp, double* host) {
  if (!s || !name || !host) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  tensor0* t = pick(s, name, comp);
  if (!t) return FR_EINVAL;
  t->download(host);
  return (int)hipGetLastError();
}

int fr_sim_set_field(fr_sim* s, const char* name, int comp, const double* host) {
  if (!s || !name || !host) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  tensor0* t = pick(s, name, comp);
  if (!t) return FR_EINVAL;
  t->upload(host);
  if (std::string(name) == "RHS_Part_Temp") s->g.qmean_valid = false;
  return (int)hipGetLastError();
}

static double* scal(fr_sim* s, const char* name) {
  grid& g = s->g;
  std::string n(name);
  if (n == "P0") return &g.P0;
  if (n == "P0_int") return &g.P0_int;
  if (n == "P0_new") return &g.P0_new;
  if (n == "P0_np1") return &g.P0_np1;
  if (n == "T_cur") return &g.T_cur;
  if (n == "C_BC") return &g.C_BC;
  if (n == "U_avg_in") return &g.U_avg_in;
  if (n == "U_avg_out") return &g.U_avg_out;
  if (n == "relres") return &g.PS_.relres_;
  return nullptr;
}

int fr_sim_get_scalar(fr_sim* s, const char* name, double* v) {
  if (!s || !name || !v) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  double* p = scal(s, name);
  if (!p) return FR_EINVAL;
  *v = *p;
  return FR_OK;
}

int fr_sim_set_scalar(fr_sim* s, const char* name, double v) {
  if (!s || !name) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  double* p = scal(s, name);
  if (!p) return FR_EINVAL;
  *p = v;
  return FR_OK;
}

int fr_sim_geom(fr_sim* s, fr_geom* out) {
  if (!s || !out) return FR_EINVAL;
  *out = s->g.size.g;
  return FR_OK;
}

int fr_sim_np(fr_sim* s, int* np) {
  if (!s || !np) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  *np = s->g.part.dev.np;
  return FR_OK;
}

int fr_sim_get_particles(fr_sim* s, int q, int st, double* host) {
  if (!s || !host || q < 0 || q > 7 || st < 0 || st > 3) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  PartDev& p = s->g.part.dev;
  const double* src = q == 7 ? p.rhs_T : p.s[q][st];
  return p.get(host, src, s->g.st);
}

int fr_sim_set_particles(fr_sim* s, int np, int q, int st, const double* host) {
  if (!s || !host || q < 0 || q > 7 || st < 0 || st > 3) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  s->g.stages_equal = false;
  PartDev& p = s->g.part.dev;
  if (np > p.cap || np < 0) return FR_EOVERFLOW;
  if (np != p.np) {  // a new particle set: storage in reference order
    p.np = np;
    s->g.part.Np = np;
    int rc = p.reset_order(s->g.st);
    if (rc) return rc;
  }
  double* dst = q == 7 ? p.rhs_T : p.s[q][st];
  return p.set(host, dst, s->g.st);
}

int fr_sim_get_flags(fr_sim* s, int* host) {
  if (!s || !host) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  return s->g.part.dev.get_flags(host, s->g.st);
}

int fr_sim_set_flags(fr_sim* s, int np, const int* host) {
  if (!s || !host) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  s->g.stages_equal = false;
  PartDev& p = s->g.part.dev;
  if (np > p.cap || np < 0) return FR_EOVERFLOW;
  if (np != p.np) {
    p.np = np;
    s->g.part.Np = np;
    int rc = p.reset_order(s->g.st);
    if (rc) return rc;
  }
  return p.set_flags(host, s->g.st);
}

int fr_sim_particles_changed(fr_sim* s) {
  if (!s) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  s->g.stages_equal = false;
  int rc = s->g.part.rebuild_cells(1);
  if (!rc) rc = (int)hipStreamSynchronize(s->g.st);
  return rc;
}

int fr_sim_diagnostics(fr_sim* s, double* out) {
  if (!s || !out) return FR_EINVAL;
  FR_TRY_S(s->g.settle_side());
  return s->g.diagnostics(out);
}

int fr_sim_set_inflow(fr_sim* s, const double* planes, int n_inlet, const double* inlet) {
  if (!s) return FR_EINVAL;
  return s->g.set_inflow(planes, n_inlet, inlet);
}

int fr_sim_reserve_particles(fr_sim* s, int capacity) {
  if (!s || capacity < 1) return FR_EINVAL;
  return s->g.reserve(capacity);
}

int fr_sim_set_learned(fr_sim* s, fr_unet* net, int mode) {
  if (!s || !s->g.PS_.sp) return FR_EINVAL;
  return fr_pcg_set_learned(s->g.PS_.sp, net, mode);
}

int fr_sim_set_collect(fr_sim* s, fr_unet_data* d) {
  if (!s || !s->g.PS_.sp) return FR_EINVAL;
  return fr_pcg_set_collect(s->g.PS_.sp, d);
}

int fr_sim_learned_stats(fr_sim* s, int* guesses_kept, int* fallbacks, double* log10_gain) {
  if (!s || !s->g.PS_.sp) return FR_EINVAL;
  return fr_pcg_learned_stats(s->g.PS_.sp, guesses_kept, fallbacks, log10_gain);
}

int fr_sim_statistics(fr_sim* s) { return s ? s->g.Statistics() : FR_EINVAL; }

int fr_sim_store(fr_sim* s) { return s ? s->g.Store() : FR_EINVAL; }

int fr_sim_touch(fr_sim* s, int* stop) {
  if (!s || !stop) return FR_EINVAL;
  bool b = false;
  int rc = s->g.Touch(&b);
  *stop = b ? 1 : 0;
  return rc;
}

int fr_sim_write_info(fr_sim* s) { return s ? s->g.Write_info() : FR_EINVAL; }

int fr_sim_num_timestep(fr_sim* s, int* n) {
  if (!s || !n) return FR_EINVAL;
  *n = s->g.num_timestep;
  return FR_OK;
}

int fr_sim_set_resort_interval(fr_sim* s, int steps) {
  if (!s || steps < 1) return FR_EINVAL;
  s->g.resort_every = steps;
  return FR_OK;
}

int fr_sim_timing(fr_sim* s, double* out) { return fr_sim_timing_ex(s, out, 6); }

int fr_sim_timing_ex(fr_sim* s, double* out, int n) {
  if (!s || !out || n < 0) return FR_EINVAL;
  const PartDev& p = s->g.part.dev;
  const double v[12] = {s->g.t_step,     s->g.t_poisson, s->g.t_particles,
                        (double)s->g.iters_step, s->g.t_momentum, (double)s->g.n_momentum,
                        s->g.t_xpass,    (double)s->g.n_xpass, p.t_vel.ms, (double)p.t_vel.n,
                        p.t_cic.ms,      (double)p.t_cic.n};
  for (int i = 0; i < n && i < 12; i++) out[i] = v[i];
  return FR_OK;
}

// ---------------------------------------------------------------- coupled run
// simulation.cpp:19-73 on one device: a periodic box feeding a duct.
struct fr_pair {
  fr_sim* box = nullptr;
  fr_sim* duct = nullptr;
  PairLink link;
  bool ready = false;
};

// The duct's 8 inflow planes from the box's first interior slice (x index 1,
// the RU x faces at 2).  At initialisation the box's RU and Rho stand in for
// the _new/_np1 fields: the duct's restart reads the box's RU.y, RU.z there
// (grid.cpp:155-164), which the ru_np1 y, z planes carry.
static int pair_copy_planes(grid& b, grid& d, bool init) {
  const double* now[8] = {b.Rho_new.pointer(),  b.Rho_np1.pointer(),  b.RU_new.x.pointer(),
                          b.RU_new.y.pointer(), b.RU_new.z.pointer(), b.RU_np1.x.pointer(),
                          b.RU_np1.y.pointer(), b.RU_np1.z.pointer()};
  const double* ini[8] = {b.Rho.pointer(),  b.Rho.pointer(),  b.RU.x.pointer(),
                          b.RU.y.pointer(), b.RU.z.pointer(), b.RU.x.pointer(),
                          b.RU.y.pointer(), b.RU.z.pointer()};
  const int xi[8] = {1, 1, 2, 1, 1, 2, 1, 1};
  return pair_planes(b.size.dev(), init ? ini : now, xi, d.inflow, b.st);
}

int fr_pair_create(fr_sim* box, fr_sim* duct, fr_pair** out) {
  if (!box || !duct || !out || box == duct) return FR_EINVAL;
  grid& b = box->g;
  grid& d = duct->g;
  const auto same = [](double x, double y) { return std::fabs(x - y) <= 1e-12 * std::fabs(y); };
  if (!b.param.IsBox() || d.param.IsBox() || b.f32 || d.f32 || coupled(box) || coupled(duct) ||
      b.size.Ny() != d.size.Ny() || b.size.Nz() != d.size.Nz() ||
      !same(b.size.Ly(), d.size.Ly()) || !same(b.size.Lz(), d.size.Lz())) {
    fprintf(stderr,
            "fr_amd: a coupled run takes an FP64 box and an FP64 duct of equal Ny, Nz, "
            "Ly, Lz, neither already coupled\n");
    return FR_EINVAL;
  }
  fr_pair* p = new fr_pair();
  p->box = box;
  p->duct = duct;
  PairLink& l = p->link;
  int rc = (int)hipMalloc(&l.d_n, sizeof(int));
  if (!rc) rc = (int)hipHostMalloc(&l.h_n, sizeof(int), hipHostMallocDefault);
  hipEvent_t* ev[3] = {&l.ev_rows, &l.ev_planes, &l.ev_duct};
  for (int i = 0; i < 3 && !rc; i++) rc = (int)hipEventCreateWithFlags(ev[i], hipEventDisableTiming);
  if (rc) {
    fr_pair_destroy(p);
    return rc;
  }
  *l.h_n = 0;
  b.pair_out = &l;
  d.pair_in = &l;
  *out = p;
  return FR_OK;
}

int fr_pair_destroy(fr_pair* p) {
  if (!p) return FR_OK;
  PairLink& l = p->link;
  if (p->box) {
    (void)hipStreamSynchronize(p->box->g.st3);
    (void)hipStreamSynchronize(p->box->g.st);
    if (p->box->g.pair_out == &l) p->box->g.pair_out = nullptr;
  }
  if (p->duct) {
    (void)hipStreamSynchronize(p->duct->g.st);
    if (p->duct->g.pair_in == &l) p->duct->g.pair_in = nullptr;
  }
  (void)hipFree(l.buf);
  (void)hipFree(l.d_n);
  (void)hipHostFree(l.h_n);
  for (hipEvent_t e : {l.ev_rows, l.ev_planes, l.ev_duct})
    if (e) (void)hipEventDestroy(e);
  delete p;
  return FR_OK;
}

// simulation.cpp:29-30: GRID.Initialize(GRID), GRID_IO.Initialize(GRID).
// The inlet row buffer holds the duct's particle capacity (reserve it first).
int fr_pair_initialize(fr_pair* p) {
  if (!p) return FR_EINVAL;
  grid& b = p->box->g;
  grid& d = p->duct->g;
  PairLink& l = p->link;
  FR_TRY_S(b.Initialize());
  const int cap = d.part.dev.cap;
  if (cap > l.cap) {
    (void)hipFree(l.buf);
    l.buf = nullptr;
    l.cap = 0;
    FR_CHECK(hipMalloc(&l.buf, sizeof(double) * 29 * (size_t)cap));
    l.cap = cap;
  }
  FR_TRY_S(pair_copy_planes(b, d, true));
  const size_t n = 8 * (size_t)d.size.Ny() * d.size.Nz();
  d.inflow_h.resize(n);
  FR_CHECK(hipMemcpyAsync(d.inflow_h.data(), d.inflow, sizeof(double) * n,
                          hipMemcpyDeviceToHost, b.st));
  FR_CHECK(hipStreamSynchronize(b.st));
  FR_TRY_S(d.Initialize());
  FR_CHECK(hipEventRecord(l.ev_duct, d.st));
  p->ready = true;
  return FR_OK;
}

// simulation.cpp:34-62.  Per substep the box's eight methods, its planes
// into the duct's inflow, the duct's eight (its migration takes the box's
// '>hi' rows), then the box's TimeAdvance_RK4 and the duct's last two.
int fr_pair_step(fr_pair* p) {
  if (!p || !p->ready) return FR_EINVAL;
  grid& b = p->box->g;
  grid& d = p->duct->g;
  PairLink& l = p->link;
  FR_TRY_S(b.begin_step());
  FR_TRY_S(d.begin_step());
  for (int k = 0; k < 4; k++) {
    FR_TRY_S(b.substep_front(k));
    FR_TRY_S(b.Update_RU_WP());
    FR_CHECK(hipStreamWaitEvent(b.st, l.ev_duct, 0));
    FR_TRY_S(pair_copy_planes(b, d, false));
    FR_CHECK(hipEventRecord(l.ev_planes, b.st));
    FR_CHECK(hipStreamWaitEvent(d.st, l.ev_planes, 0));
    FR_TRY_S(d.substep_front(k));
    FR_TRY_S(b.TimeAdvance_RK4());
    FR_TRY_S(d.Update_RU_WP());
    FR_TRY_S(d.TimeAdvance_RK4());
    FR_CHECK(hipEventRecord(l.ev_duct, d.st));
  }
  FR_TRY_S(b.Correct_mean_RU());
  FR_TRY_S(d.Correct_mean_RU());
  FR_TRY_S(b.TimeAdvance());
  FR_TRY_S(d.TimeAdvance());
  FR_TRY_S(b.end_step());
  return d.end_step();
}

// grid::Touch(grid& g_IO) (grid.cpp:77-98): one read of touch.check; on a
// stop both grids store.
int fr_pair_touch(fr_pair* p, int* stop) {
  if (!p || !stop) return FR_EINVAL;
  bool s = false;
  FR_TRY_S(p->box->g.Touch(&s));
  p->duct->g.Is_touch = s;
  *stop = s ? 1 : 0;
  return s ? p->duct->g.Store() : FR_OK;
}

}  // extern "C"
