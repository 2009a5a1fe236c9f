// The reference's member surface of tensor0, tensor1, particle and
// poisson_IO (tensor0.h:17-131, tensor1.h:6-90, particle.h:126-154,
// poisson_IO.h:76-88) on the device classes of fr_host.h.  Every member is
// queued on its object's stream; the stencil members go through the C-ABI
// operators (fr_t0_*), the element-wise ones through one kernel that keeps
// the reference's operand order (k_arith), the particle members through the
// single phases of the fused particle kernels.  FP64 storage only.
#include <cstdio>

#include "fr_host.h"

using namespace fr;

namespace fr {
static int g_status = 0;
int host_status(bool reset) {
  const int r = g_status;
  if (reset) g_status = 0;
  return r;
}
void host_note(int rc) {
  if (rc && !g_status) g_status = rc;
}
}  // namespace fr

// ---------------------------------------------------------------- tensor0
void tensor0::arith(int op, double a, const tensor0* A, const tensor0* B, double b,
                    const tensor0* C, double c, const tensor0* D) {
  for (const tensor0* t : {(const tensor0*)this, A, B, C, D})
    if (t && (t->f32 || !t->ptr)) {
      fprintf(stderr, "fr_amd: tensor0 members take FP64 fields\n");
      host_note(FR_EINVAL);
      return;
    }
  auto P = [](const tensor0* t) { return t ? (const double*)t->ptr : nullptr; };
  host_note(t0_arith(p_->size(), op, (double*)ptr, a, P(A), P(B), b, P(C), c, P(D), st_));
}

tensor0& tensor0::operator=(double a) {
  arith(AR_SET_S, a, nullptr, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator+=(double a) {
  arith(AR_ADD_S, a, nullptr, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator-=(double a) {
  arith(AR_SUB_S, a, nullptr, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator*=(double a) {
  arith(AR_MUL_S, a, nullptr, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator/=(double a) {
  arith(AR_DIV_S, a, nullptr, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator=(const tensor0& a) {
  if (&a != this) arith(AR_SET_T, 0, &a, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator+=(const tensor0& a) {
  arith(AR_ADD_T, 0, &a, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator-=(const tensor0& a) {
  arith(AR_SUB_T, 0, &a, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator*=(const tensor0& a) {
  arith(AR_MUL_T, 0, &a, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
tensor0& tensor0::operator/=(const tensor0& a) {
  arith(AR_DIV_T, 0, &a, nullptr, 0, nullptr, 0, nullptr);
  return *this;
}
void tensor0::PlusEqual_Mult(tensor0& a, tensor0& b) {
  arith(AR_PEM_TT, 0, &a, &b, 0, nullptr, 0, nullptr);
}
void tensor0::PlusEqual_Mult(tensor0& a, double b) {
  arith(AR_PEM_ST, b, &a, nullptr, 0, nullptr, 0, nullptr);
}
void tensor0::PlusEqual_Mult(double a, tensor0& b) {
  arith(AR_PEM_ST, a, &b, nullptr, 0, nullptr, 0, nullptr);
}
void tensor0::Equal_Divide(tensor0& a, double b) {
  arith(AR_DIV_TS, b, &a, nullptr, 0, nullptr, 0, nullptr);
}
void tensor0::Equal_Mult(tensor0& a, tensor0& b) {
  arith(AR_MUL_TT, 0, &a, &b, 0, nullptr, 0, nullptr);
}
void tensor0::Equal_Mult(tensor0& a, double b) {
  arith(AR_MUL_ST, b, &a, nullptr, 0, nullptr, 0, nullptr);
}
void tensor0::Equal_Mult(double a, tensor0& b) {
  arith(AR_MUL_ST, a, &b, nullptr, 0, nullptr, 0, nullptr);
}
void tensor0::Equal_LinComb(double a, tensor0& T, double b, tensor0& R) {
  arith(AR_LC_STST, a, &T, nullptr, b, &R, 0, nullptr);
}
void tensor0::Equal_LinComb(tensor0& A, tensor0& T, double b, tensor0& R) {
  arith(AR_LC_TTST, 0, &A, &T, b, &R, 0, nullptr);
}
void tensor0::Equal_LinComb(tensor0& A, tensor0& T, tensor0& B, tensor0& R) {
  arith(AR_LC_TTTT, 0, &A, &T, 0, &B, 0, &R);
}
void tensor0::Equal_LinComb(double a, tensor0& T, tensor0& B, tensor0& R) {
  arith(AR_LC_STTT, a, &T, nullptr, 0, &B, 0, &R);
}
void tensor0::Equal_LinComb(double a, tensor0& T, double b, tensor0& R, double c, tensor0& S) {
  arith(AR_LC_STSTST, a, &T, nullptr, b, &R, c, &S);
}

static bool fp64(const tensor0& t) {
  if (t.is_f32() || !t.pointer()) {
    fprintf(stderr, "fr_amd: tensor0 members take FP64 fields\n");
    host_note(FR_EINVAL);
    return false;
  }
  return true;
}

void tensor0::Equal_Div_F2C(tensor1& T) {  // tensor0.cpp:303-310
  if (fp64(*this) && fp64(T.x) && fp64(T.y) && fp64(T.z))
    host_note(fr_t0_div_f2c(&p_->g, pointer(), T.x.pointer(), T.y.pointer(), T.z.pointer(),
                            st_));
}
void tensor0::Equal_Del2(tensor0& T) {  // tensor0.cpp:312-319
  if (fp64(*this) && fp64(T)) host_note(fr_t0_del2(&p_->g, pointer(), T.pointer(), st_));
}
static void interp(tensor0& out, tensor0& in, int axis, int dir) {  // tensor0.cpp:353-406
  if (fp64(out) && fp64(in))
    host_note(fr_t0_interp(&out.parameter()->g, out.pointer(), in.pointer(), axis, dir,
                           out.stream()));
}
void tensor0::Equal_Ix_C2F(tensor0& T) { interp(*this, T, 0, 0); }
void tensor0::Equal_Iy_C2F(tensor0& T) { interp(*this, T, 1, 0); }
void tensor0::Equal_Iz_C2F(tensor0& T) { interp(*this, T, 2, 0); }
void tensor0::Equal_Ix_F2C(tensor0& T) { interp(*this, T, 0, 1); }
void tensor0::Equal_Iy_F2C(tensor0& T) { interp(*this, T, 1, 1); }
void tensor0::Equal_Iz_F2C(tensor0& T) { interp(*this, T, 2, 1); }
void tensor0::Equal_I_F2C(tensor1& T) {  // tensor0.cpp:408-416
  if (fp64(*this) && fp64(T.x) && fp64(T.y) && fp64(T.z))
    host_note(
        fr_t0_i_f2c(&p_->g, pointer(), T.x.pointer(), T.y.pointer(), T.z.pointer(), st_));
}

// ---------------------------------------------------------------- tensor1
tensor1& tensor1::operator=(double a) {
  x = a;
  y = a;
  z = a;
  return *this;
}
tensor1& tensor1::operator+=(double a) {
  x += a;
  y += a;
  z += a;
  return *this;
}
tensor1& tensor1::operator-=(double a) {
  x -= a;
  y -= a;
  z -= a;
  return *this;
}
tensor1& tensor1::operator*=(double a) {
  x *= a;
  y *= a;
  z *= a;
  return *this;
}
tensor1& tensor1::operator/=(double a) {
  x /= a;
  y /= a;
  z /= a;
  return *this;
}
tensor1& tensor1::operator=(const tensor1& a) {
  x = a.x;
  y = a.y;
  z = a.z;
  return *this;
}
tensor1& tensor1::operator+=(const tensor1& a) {
  x += a.x;
  y += a.y;
  z += a.z;
  return *this;
}
tensor1& tensor1::operator-=(const tensor1& a) {
  x -= a.x;
  y -= a.y;
  z -= a.z;
  return *this;
}
tensor1& tensor1::operator*=(const tensor1& a) {
  x *= a.x;
  y *= a.y;
  z *= a.z;
  return *this;
}
tensor1& tensor1::operator/=(const tensor1& a) {
  x /= a.x;
  y /= a.y;
  z /= a.z;
  return *this;
}
tensor1& tensor1::operator=(const tensor0& a) {
  x = a;
  y = a;
  z = a;
  return *this;
}
tensor1& tensor1::operator+=(const tensor0& a) {
  x += a;
  y += a;
  z += a;
  return *this;
}
tensor1& tensor1::operator-=(const tensor0& a) {
  x -= a;
  y -= a;
  z -= a;
  return *this;
}
tensor1& tensor1::operator*=(const tensor0& a) {
  x *= a;
  y *= a;
  z *= a;
  return *this;
}
tensor1& tensor1::operator/=(const tensor0& a) {
  x /= a;
  y /= a;
  z /= a;
  return *this;
}
void tensor1::Update_Ghosts() {
  x.Update_Ghosts();
  y.Update_Ghosts();
  z.Update_Ghosts();
}
void tensor1::Update_Ghosts_CUM() {
  x.Update_Ghosts_CUM();
  y.Update_Ghosts_CUM();
  z.Update_Ghosts_CUM();
}
double tensor1::sum() { return x.sum() + y.sum() + z.sum(); }
double tensor1::max() {  // tensor1.cpp:320-328
  double m = x.max();
  const double my = y.max(), mz = z.max();
  m = (my > m) ? my : m;
  return (mz > m) ? mz : m;
}
double tensor1::min() {
  double m = x.min();
  const double my = y.min(), mz = z.min();
  m = (my < m) ? my : m;
  return (mz < m) ? mz : m;
}
void tensor1::PlusEqual_Mult(tensor1& a, tensor1& b) {
  x.PlusEqual_Mult(a.x, b.x);
  y.PlusEqual_Mult(a.y, b.y);
  z.PlusEqual_Mult(a.z, b.z);
}
void tensor1::PlusEqual_Mult(tensor1& a, double b) {
  x.PlusEqual_Mult(a.x, b);
  y.PlusEqual_Mult(a.y, b);
  z.PlusEqual_Mult(a.z, b);
}
void tensor1::PlusEqual_Mult(double a, tensor1& b) {
  x.PlusEqual_Mult(a, b.x);
  y.PlusEqual_Mult(a, b.y);
  z.PlusEqual_Mult(a, b.z);
}
void tensor1::Equal_Divide(tensor1& a, tensor1& b) {
  x.Equal_Divide(a.x, b.x);
  y.Equal_Divide(a.y, b.y);
  z.Equal_Divide(a.z, b.z);
}
void tensor1::Equal_Divide(tensor1& a, double b) {
  x.Equal_Divide(a.x, b);
  y.Equal_Divide(a.y, b);
  z.Equal_Divide(a.z, b);
}
void tensor1::Equal_Divide(tensor1& a, tensor0& b) {
  x.Equal_Divide(a.x, b);
  y.Equal_Divide(a.y, b);
  z.Equal_Divide(a.z, b);
}
void tensor1::Equal_Divide(tensor0& a, tensor1& b) {
  x.Equal_Divide(a, b.x);
  y.Equal_Divide(a, b.y);
  z.Equal_Divide(a, b.z);
}
void tensor1::Equal_Mult(tensor1& a, tensor1& b) {
  x.Equal_Mult(a.x, b.x);
  y.Equal_Mult(a.y, b.y);
  z.Equal_Mult(a.z, b.z);
}
void tensor1::Equal_Mult(tensor1& a, double b) {
  x.Equal_Mult(a.x, b);
  y.Equal_Mult(a.y, b);
  z.Equal_Mult(a.z, b);
}
void tensor1::Equal_Mult(double a, tensor1& b) {
  x.Equal_Mult(a, b.x);
  y.Equal_Mult(a, b.y);
  z.Equal_Mult(a, b.z);
}
void tensor1::Equal_LinComb(double a, tensor1& T, double b, tensor1& R) {
  x.Equal_LinComb(a, T.x, b, R.x);
  y.Equal_LinComb(a, T.y, b, R.y);
  z.Equal_LinComb(a, T.z, b, R.z);
}
void tensor1::Equal_LinComb(tensor1& A, tensor1& T, double b, tensor1& R) {
  x.Equal_LinComb(A.x, T.x, b, R.x);
  y.Equal_LinComb(A.y, T.y, b, R.y);
  z.Equal_LinComb(A.z, T.z, b, R.z);
}
void tensor1::Equal_LinComb(tensor1& A, tensor1& T, tensor1& B, tensor1& R) {
  x.Equal_LinComb(A.x, T.x, B.x, R.x);
  y.Equal_LinComb(A.y, T.y, B.y, R.y);
  z.Equal_LinComb(A.z, T.z, B.z, R.z);
}
void tensor1::Equal_LinComb(double a, tensor1& T, tensor1& B, tensor1& R) {
  x.Equal_LinComb(a, T.x, B.x, R.x);
  y.Equal_LinComb(a, T.y, B.y, R.y);
  z.Equal_LinComb(a, T.z, B.z, R.z);
}
void tensor1::Equal_LinComb(double a, tensor1& T, double b, tensor1& R, double c, tensor1& S) {
  x.Equal_LinComb(a, T.x, b, R.x, c, S.x);
  y.Equal_LinComb(a, T.y, b, R.y, c, S.y);
  z.Equal_LinComb(a, T.z, b, R.z, c, S.z);
}
void tensor1::Equal_I_C2F(tensor0& T) {  // tensor1.cpp:269-274
  x.Equal_Ix_C2F(T);
  y.Equal_Iy_C2F(T);
  z.Equal_Iz_C2F(T);
}
void tensor1::Equal_Ix_C2F(tensor1& T) {
  x.Equal_Ix_C2F(T.x);
  y.Equal_Ix_C2F(T.y);
  z.Equal_Ix_C2F(T.z);
}
void tensor1::Equal_Iy_C2F(tensor1& T) {
  x.Equal_Iy_C2F(T.x);
  y.Equal_Iy_C2F(T.y);
  z.Equal_Iy_C2F(T.z);
}
void tensor1::Equal_Iz_C2F(tensor1& T) {
  x.Equal_Iz_C2F(T.x);
  y.Equal_Iz_C2F(T.y);
  z.Equal_Iz_C2F(T.z);
}
void tensor1::Equal_I_F2C(tensor1& T) {
  x.Equal_Ix_F2C(T.x);
  y.Equal_Iy_F2C(T.y);
  z.Equal_Iz_F2C(T.z);
}
void tensor1::Equal_Grad_C2F(tensor0& T) {  // tensor1.cpp:242-260
  if (fp64(x) && fp64(y) && fp64(z) && fp64(T))
    host_note(fr_t1_grad_c2f(&x.parameter()->g, x.pointer(), y.pointer(), z.pointer(),
                             T.pointer(), x.stream()));
}
void tensor1::Equal_Del2(tensor1& T) {
  x.Equal_Del2(T.x);
  y.Equal_Del2(T.y);
  z.Equal_Del2(T.z);
}

// ---------------------------------------------------------------- particle
particle::~particle() {
  (void)hipFree(rows);
  (void)hipFree(rows_n);
  dev.release();
}

// The stage coefficients of particle.cpp:314-399 and 480-496.
AdvCoef particle::adv_coef(int k, double pre, double post) const {
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
  return c;
}

// phase bits of the fused particle kernel (particles.hip)
enum { PHASE_GATHER = 1, PHASE_POS = 2, PHASE_VEL = 4, PHASE_TEMP = 8 };

int particle::gas2part_velocity(tensor1& U) {  // particle.cpp:340-360
  if (U.x.is_f32()) return FR_EINVAL;
  const double* u[3] = {U.x.pointer(), U.y.pointer(), U.z.pointer()};
  return dev.advance_phase(size_->dev(), adv_coef(0, 0, 0), u, PHASE_GATHER, st_);
}
int particle::update_position(int k, double pre, double post) {  // particle.cpp:314-338
  return dev.advance_phase(size_->dev(), adv_coef(k, pre, post), nullptr, PHASE_POS, st_);
}
int particle::update_velocity(int k, double pre, double post) {  // particle.cpp:362-399
  return dev.advance_phase(size_->dev(), adv_coef(k, pre, post), nullptr, PHASE_VEL, st_);
}
int particle::update_Temp(int k, double pre, double post) {  // particle.cpp:480-496
  return dev.advance_phase(size_->dev(), adv_coef(k, pre, post), nullptr, PHASE_TEMP, st_);
}
static double i_tth(const params& p) { return p.Nu() * PI * p.Dp() * p.k(); }
int particle::gas2part_Temp_int(tensor0& T) {  // particle.cpp:402-421
  if (T.is_f32()) return FR_EINVAL;
  return dev.temp_split(size_->dev(), 1, 0, T.pointer(), i_tth(*param_), st_);
}
int particle::gas2part_Temp_new(tensor0& T) {  // particle.cpp:422-441
  if (T.is_f32()) return FR_EINVAL;
  return dev.temp_split(size_->dev(), 2, 0, T.pointer(), i_tth(*param_), st_);
}
int particle::Compute_RHS_Temp_int() {  // particle.cpp:443-460
  return dev.temp_split(size_->dev(), 1, 1, nullptr, i_tth(*param_), st_);
}
int particle::Compute_RHS_Temp_new() {  // particle.cpp:461-478
  return dev.temp_split(size_->dev(), 2, 1, nullptr, i_tth(*param_), st_);
}
// The heat deposits read RHS_T in bucket order: the buckets of the stage's
// positions are built with it first.
int particle::part2gas_Temp_int(tensor0& A) {  // particle.cpp:586-620
  FR_RET(rebuild_cells(1));
  return part2gas_Temp(1, A);
}
int particle::part2gas_Temp_new(tensor0& A) {  // particle.cpp:624-658
  FR_RET(rebuild_cells(2));
  return part2gas_Temp(2, A);
}

// particle::Send_Recv(particle& pbox) (particle.cpp:814-851, 860-865) on a
// duct rank that is inlet and outlet: pbox's '>hi' list of its last
// Send_Recv() is gathered on pbox's stream, then appended here.
int particle::Send_Recv(particle& pbox) {
  if (param_->IsBox() || !pbox.param_->IsBox()) return FR_EINVAL;
  if (!rows_n) FR_CHECK(hipMalloc(&rows_n, sizeof(int)));
  if (rows_cap < dev.cap) {
    (void)hipFree(rows);
    rows = nullptr;
    rows_cap = 0;
    FR_CHECK(hipMalloc(&rows, sizeof(double) * 29 * (size_t)dev.cap));
    rows_cap = dev.cap;
  }
  FR_RET(pbox.dev.gather_hi(rows, rows_cap, rows_n, pbox.st_));
  int n = 0;
  FR_CHECK(hipMemcpyAsync(&n, rows_n, sizeof(int), hipMemcpyDeviceToHost, pbox.st_));
  FR_CHECK(hipStreamSynchronize(pbox.st_));
  if (n > rows_cap) {
    fprintf(stderr, "fr_amd: %d box particles enter the duct, capacity %d\n", n, rows_cap);
    return FR_EOVERFLOW;
  }
  return Send_Recv_duct_dev(rows, n);
}

// ---------------------------------------------------------------- poisson_IO
int poisson_IO::init(params* p, gridsize* g, hipStream_t st) {
  if (g->IsBox()) return FR_EINVAL;
  FR_RET(PS.init(p, g, st));
  if (PS.sd) FR_RET(duct_poisson_timing(PS.sd, p->Iteration(), nullptr, nullptr));
  return FR_OK;
}

int poisson_IO::account() {
  if (PS.sp) {  // one PCG solve: a single outer iteration
    n1_ = 1;
    n2_ = PS.num_iteration();
    return FR_OK;
  }
  n1_ = PS.num_iteration();
  n2_ = 0;
  double f = 0, s = 0;
  FR_RET(duct_poisson_timing(PS.sd, PS.par->Iteration(), &f, &s));
  FFT_TIME = f * 1e-3;
  SOLVER_TIME = s * 1e-3;
  return FR_OK;
}

int poisson_IO::Solve(tensor1& C, tensor0& P, tensor0& RHS) {  // poisson_IO.cpp:81-114
  FR_RET(PS.Solve(C, P, RHS));
  return account();
}

int poisson_IO::Solve_face_rho(tensor1& Rho_face, tensor0& P, tensor0& RHS) {
  FR_RET(PS.Solve_face_rho(Rho_face, P, RHS));
  return account();
}
