// Host classes of the MI355X solver library, mirroring the reference's
// plug-in surface: params (params.h:8-130), gridsize (gridsize.h:5-59),
// tensor0 (tensor0.h:6-132), tensor1 (tensor1.h:6-90), particle
// (particle.h:12-155), poisson (poisson.h:19-81) and grid (grid.h:16-115).
// Field storage is device memory with the reference's ghosted layout; the
// member functions launch the gfx950 kernels of csrc/.  One process drives
// one device; MPI is replaced by single-rank semantics (every neighbour is
// self), so halos are periodic wraps and reductions are device trees.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "../csrc/fr_internal.hpp"

const double PI(3.141592653589793);  // params.h:5

// params.cpp:6-77: fixed-order '#'-commented parameter file.
class params {
 public:
  int IsBox_ = 1, Nx_tot_ = 0, Ny_tot_ = 0, Nz_tot_ = 0, bs_ = 1, Np_track_ = 0;
  double Lx_ = 0, Ly_ = 0, Lz_ = 0, T_final_ = 0, dt_ = 0, threshold_ = 0;
  int Statistics_ = 0, Stat_print_ = 0, Data_freq_fast_ = 1, Data_freq_slow_ = 1;
  std::string Data_dir_, Stat_dir_;
  int Initial_ = 0, PreCond_ = 0, Solver_ = 0, cooling_ = 0;
  int Iteration_ = 30, Iteration1_ = 30, Iteration2_ = 30;
  double epsilon_ = 1e-7;
  int TWC_ = 1, ParticleGravity_ = 0;
  double A_ = 0, A1_ = 0, A2_ = 0, Rho0_ = 1, U0_ = 0, T0_ = 1, Mu0_ = 0, k_ = 0;
  double gx_ = 0, gy_ = 0, gz_ = 0, gx1_ = 0, gy1_ = 0, gz1_ = 0, gx2_ = 0, gy2_ = 0, gz2_ = 0;
  double Cp_ = 0, Cv_ = 0, R_ = 0, np0_ = 0, Dp_ = 0, Rhop_ = 0, Cvp_ = 0, epsilonp_ = 0;
  double Nu_ = 0, I0_ = 0, I01_ = 0, I02_ = 0, Tp_ = 0, mp_ = 0, P0_ = 0, W_ = 0;
  int N0_ = 0;
  // extension fields appended after I0_1 I0_2 (absent in reference files):
  // per-particle response times uniform on [Tp_lo, Tp_hi], drawn from the
  // particle's flag; 0 0 keeps the single Tp of params.cpp:53
  double Tp_lo_ = 0, Tp_hi_ = 0;
  bool ok = false;

  params() {}
  explicit params(const char* path);
  void update(double time);
  double Lx() const { return Lx_; }
  double Ly() const { return Ly_; }
  double Lz() const { return Lz_; }
  double Tp() const { return Tp_; }
  double R() const { return R_; }
  double P0() const { return P0_; }
  double mp() const { return mp_; }
  double Nu() const { return Nu_; }
  double I0() const { return I0_; }
  double epsilonp() const { return epsilonp_; }
  double Cvp() const { return Cvp_; }
  double Dp() const { return Dp_; }
  double k() const { return k_; }
  double Cv() const { return Cv_; }
  double Cp() const { return Cp_; }
  double Mu0() const { return Mu0_; }
  double T0() const { return T0_; }
  double U0() const { return U0_; }
  double Rho0() const { return Rho0_; }
  double A() const { return A_; }
  double gx() const { return gx_; }
  double gy() const { return gy_; }
  double gz() const { return gz_; }
  double Vol() const { return Lx_ * Ly_ * Lz_; }
  bool IsBox() const { return IsBox_ != 0; }
  int N0() const { return N0_; }
  int Nx_tot() const { return Nx_tot_; }
  int Ny_tot() const { return Ny_tot_; }
  int Nz_tot() const { return Nz_tot_; }
  int bs() const { return bs_; }
  double T_final() const { return T_final_; }
  double dt() const { return dt_; }
  int Initial() const { return Initial_; }
  int Iteration() const { return Iteration_; }
  double epsilon() const { return epsilon_; }
  bool Is_Cooling() const { return cooling_ != 0; }
  bool TWC() const { return TWC_ != 0; }
  bool ParticleGravity() const { return ParticleGravity_ != 0; }
  int Solver() const { return Solver_; }
};

// gridsize.cpp:4-88 on a single rank.
class gridsize {
 public:
  fr_geom g;
  int Nx_tot_, Ny_tot_, Nz_tot_;
  double xl_, xh_, yl_, yh_, zl_, zh_;
  gridsize() {}
  explicit gridsize(const params& p);
  int Nx() const { return g.Nx; }
  int Ny() const { return g.Ny; }
  int Nz() const { return g.Nz; }
  int bs() const { return g.bs; }
  double dx() const { return g.dx; }
  double dy() const { return g.dy; }
  double dz() const { return g.dz; }
  double Lx() const { return g.Lx; }
  double Ly() const { return g.Ly; }
  double Lz() const { return g.Lz; }
  double Vcell() const { return g.dx * g.dy * g.dz; }
  int size() const { return g.Nx * g.Ny * g.Nz; }
  long size_tot() const { return (long)Nx_tot_ * Ny_tot_ * Nz_tot_; }
  bool IsBox() const { return g.is_box != 0; }
  fr::Dev dev() const { return fr::make_dev(&g); }
};

class tensor1;

// tensor0: a ghosted scalar field in device memory (tensor0.h:6-132), FP64
// or, in the FP32 precision form, FP32 storage.
class tensor0 {
  gridsize* p_ = nullptr;
  void* ptr = nullptr;
  hipStream_t st_ = 0;
  bool f32 = false;
  size_t es() const { return f32 ? sizeof(float) : sizeof(double); }
  double reduce(int op);

 public:
  tensor0() {}
  tensor0(gridsize* p, hipStream_t st);
  ~tensor0();
  tensor0(const tensor0&) = delete;
  void init(gridsize* p, hipStream_t st, bool fp32 = false);
  double* pointer() const { return (double*)ptr; }  // FP64 storage
  template <typename T>
  T* p() const {
    return (T*)ptr;
  }
  bool is_f32() const { return f32; }
  hipStream_t stream() const { return st_; }
  int size() const { return p_->size(); }
  void swap(tensor0& o) {
    std::swap(ptr, o.ptr);
    std::swap(f32, o.f32);
  }
  void Update_Ghosts();
  void Update_Ghosts_CUM();
  void assign(double a);        // operator=(double)
  void add(double a);           // operator+=(double)
  void copy(const tensor0& a);  // operator=(tensor0)
  void Equal_Divide(double a, const tensor0& b);
  void Equal_Divide(const tensor0& a, const tensor0& b);
  double sum();
  double mean();
  void make_mean_zero();
  // make_mean_zero with the mean kept on the device (ws: RED_PARTS + 1
  // doubles), no host synchronisation
  void make_mean_zero_async(double* ws);
  double max();
  double min();
  double max_abs();
  void download(double* host) const;  // converted to double
  void upload(const double* host);

  // ---- the reference's member surface (tensor0.h:17-131), FP64 storage,
  // each call queued on the tensor's stream (members.cpp).  The reference's
  // members return nothing: a call that fails leaves its status in
  // fr::host_status().
  int Nx() const { return p_->Nx(); }
  int Ny() const { return p_->Ny(); }
  int Nz() const { return p_->Nz(); }
  int bs() const { return p_->bs(); }
  double dx() const { return p_->dx(); }
  double dy() const { return p_->dy(); }
  double dz() const { return p_->dz(); }
  double Lx() const { return p_->Lx(); }
  double Ly() const { return p_->Ly(); }
  double Lz() const { return p_->Lz(); }
  gridsize* parameter() const { return p_; }
  tensor0& operator=(double a);
  tensor0& operator+=(double a);
  tensor0& operator-=(double a);
  tensor0& operator*=(double a);
  tensor0& operator/=(double a);
  tensor0& operator=(const tensor0& a);  // values; the storage stays this tensor's
  tensor0& operator+=(const tensor0& a);
  tensor0& operator-=(const tensor0& a);
  tensor0& operator*=(const tensor0& a);
  tensor0& operator/=(const tensor0& a);
  void PlusEqual_Mult(tensor0& a, tensor0& b);
  void PlusEqual_Mult(tensor0& a, double b);
  void PlusEqual_Mult(double a, tensor0& b);
  void Equal_Divide(tensor0& a, double b);
  void Equal_Mult(tensor0& a, tensor0& b);
  void Equal_Mult(tensor0& a, double b);
  void Equal_Mult(double a, tensor0& b);
  void Equal_LinComb(double a, tensor0& T, double b, tensor0& R);
  void Equal_LinComb(tensor0& A, tensor0& T, double b, tensor0& R);
  void Equal_LinComb(tensor0& A, tensor0& T, tensor0& B, tensor0& R);
  void Equal_LinComb(double a, tensor0& T, tensor0& B, tensor0& R);
  void Equal_LinComb(double a, tensor0& T, double b, tensor0& R, double c, tensor0& S);
  void Equal_Div_F2C(tensor1& T);
  void Equal_Del2(tensor0& T);
  void Equal_Ix_C2F(tensor0& T);
  void Equal_Iy_C2F(tensor0& T);
  void Equal_Iz_C2F(tensor0& T);
  void Equal_Ix_F2C(tensor0& T);
  void Equal_Iy_F2C(tensor0& T);
  void Equal_Iz_F2C(tensor0& T);
  void Equal_I_F2C(tensor1& T);

 private:
  void arith(int op, double a, const tensor0* A, const tensor0* B, double b, const tensor0* C,
             double c, const tensor0* D);
};

// tensor1: three staggered components (tensor1.h:6-90).
class tensor1 {
 public:
  tensor0 x, y, z;
  tensor1() {}
  void init(gridsize* p, hipStream_t st, bool fp32 = false) {
    x.init(p, st, fp32);
    y.init(p, st, fp32);
    z.init(p, st, fp32);
  }
  tensor0& c(int m) { return m == 0 ? x : (m == 1 ? y : z); }
  void swap(tensor1& o) {
    x.swap(o.x);
    y.swap(o.y);
    z.swap(o.z);
  }
  void copy(const tensor1& a) {
    x.copy(a.x);
    y.copy(a.y);
    z.copy(a.z);
  }
  void assign(double a) {
    x.assign(a);
    y.assign(a);
    z.assign(a);
  }
  void Equal_Divide(double a, const tensor1& b) {
    x.Equal_Divide(a, b.x);
    y.Equal_Divide(a, b.y);
    z.Equal_Divide(a, b.z);
  }
  void make_mean_zero() {
    x.make_mean_zero();
    y.make_mean_zero();
    z.make_mean_zero();
  }
  void make_mean_zero_async(double* ws) {
    x.make_mean_zero_async(ws);
    y.make_mean_zero_async(ws);
    z.make_mean_zero_async(ws);
  }
  double* const* ptrs3(double* out[3]) {
    out[0] = x.pointer();
    out[1] = y.pointer();
    out[2] = z.pointer();
    return out;
  }

  // ---- the reference's member surface (tensor1.h:6-90), per component
  // through the tensor0 members (members.cpp)
  int Nx() const { return x.Nx(); }
  int Ny() const { return x.Ny(); }
  int Nz() const { return x.Nz(); }
  int bs() const { return x.bs(); }
  double dx() const { return x.dx(); }
  double dy() const { return x.dy(); }
  double dz() const { return x.dz(); }
  double Lx() const { return x.Lx(); }
  double Ly() const { return x.Ly(); }
  double Lz() const { return x.Lz(); }
  gridsize* parameter() const { return x.parameter(); }
  tensor1& operator=(double a);
  tensor1& operator+=(double a);
  tensor1& operator-=(double a);
  tensor1& operator*=(double a);
  tensor1& operator/=(double a);
  tensor1& operator=(const tensor1& a);
  tensor1& operator+=(const tensor1& a);
  tensor1& operator-=(const tensor1& a);
  tensor1& operator*=(const tensor1& a);
  tensor1& operator/=(const tensor1& a);
  tensor1& operator=(const tensor0& a);
  tensor1& operator+=(const tensor0& a);
  tensor1& operator-=(const tensor0& a);
  tensor1& operator*=(const tensor0& a);
  tensor1& operator/=(const tensor0& a);
  void Update_Ghosts();
  void Update_Ghosts_CUM();
  double sum();
  double max();
  double min();
  void PlusEqual_Mult(tensor1& a, tensor1& b);
  void PlusEqual_Mult(tensor1& a, double b);
  void PlusEqual_Mult(double a, tensor1& b);
  void Equal_Divide(tensor1& a, tensor1& b);
  void Equal_Divide(tensor1& a, double b);
  void Equal_Divide(tensor1& a, tensor0& b);
  void Equal_Divide(tensor0& a, tensor1& b);
  void Equal_Mult(tensor1& a, tensor1& b);
  void Equal_Mult(tensor1& a, double b);
  void Equal_Mult(double a, tensor1& b);
  void Equal_LinComb(double a, tensor1& T, double b, tensor1& R);
  void Equal_LinComb(tensor1& A, tensor1& T, double b, tensor1& R);
  void Equal_LinComb(tensor1& A, tensor1& T, tensor1& B, tensor1& R);
  void Equal_LinComb(double a, tensor1& T, tensor1& B, tensor1& R);
  void Equal_LinComb(double a, tensor1& T, double b, tensor1& R, double c, tensor1& S);
  void Equal_I_C2F(tensor0& T);
  void Equal_Ix_C2F(tensor1& T);
  void Equal_Iy_C2F(tensor1& T);
  void Equal_Iz_C2F(tensor1& T);
  void Equal_I_F2C(tensor1& T);
  void Equal_Grad_C2F(tensor0& T);
  void Equal_Del2(tensor1& T);
};

// particle (particle.h:12-155): device SoA set on one rank.
class particle {
 public:
  params* param_ = nullptr;
  gridsize* size_ = nullptr;
  hipStream_t st_ = 0;
  fr::PartDev dev;
  int Np = 0;
  particle() {}
  int init(params* p, gridsize* s, hipStream_t st, int capacity);
  ~particle();
  int load_random();  // particle.cpp:919-958
  // particle.cpp:340-399 + 314-338 + 480-496 (gather and the three ODE updates);
  // set0 (last stage of a full step): stage 0 takes the stage-3 values too
  int advance(int k, double pre, double post, tensor1& U, bool set0 = false, bool eq = false);
  int part2gas_concentration(tensor0& A, hipStream_t on = 0);  // particle.cpp:526-559
  // rk: reference indices to order the deposit by (a copy taken before a
  // migration that runs beside it); default the current ones
  // halo = false leaves the periodic ghost shell of A unformed (the fold into
  // the interior is done either way)
  // head: the bucket build's first half (PartDev::build_head of the stage-1
  // positions) was issued on another stream and recorded in this event
  int part2gas_velocity(tensor1& A, const int* rk = nullptr, bool halo = true,
                        hipEvent_t head = nullptr);  // particle.cpp:498-524
  int Send_Recv(hipStream_t on = 0);                            // particle.cpp:853-858
  // duct rank that is inlet and outlet: Send_Recv(particle& pbox)
  // (particle.cpp:860-865, 814-851) with the box's Rbuf1 given as n inlet
  // particles of 29 values each
  int Send_Recv_duct(const double* inlet, int n);
  // the same with the n rows in device memory (the coupled box's '>hi' list)
  int Send_Recv_duct_dev(const double* rows, int n);
  // gas2part_Temp_{new,int} + Compute_RHS_Temp_{new,int} (402-478)
  int temp(int stage, double c0, const tensor0& Rho, hipStream_t on = 0);
  // sum, summed: see PartDev::deposit_cic (the interior sum of A after the
  // fold, formed by the deposit itself where its form allows)
  int part2gas_Temp(int stage, tensor0& A, hipStream_t on = 0, double* sum = nullptr,
                    bool* summed = nullptr);  // particle.cpp:586-658
  int rebuild_cells(int stage, hipStream_t on = 0);

  // ---- the reference's member surface (particle.h:126-154), one member at
  // a time (FP64; members.cpp).  The time step itself runs them fused
  // (advance, temp).  part2gas_velocity, part2gas_concentration, Send_Recv
  // and load_random above keep the reference's names and arguments.
  int update_position(int k, double pre, double post);
  int update_velocity(int k, double pre, double post);
  int update_Temp(int k, double pre, double post);
  int gas2part_velocity(tensor1& U);
  int gas2part_Temp_int(tensor0& T);
  int gas2part_Temp_new(tensor0& T);
  int Compute_RHS_Temp_int();
  int Compute_RHS_Temp_new();
  int part2gas_Temp_int(tensor0& A);
  int part2gas_Temp_new(tensor0& A);
  // the duct rank that is inlet and outlet: the box's '>hi' list of its last
  // Send_Recv() is appended (particle.cpp:814-851, 860-865)
  int Send_Recv(particle& pbox);
  double* rows = nullptr;  // that list in device memory, 29 values a row
  int rows_cap = 0;
  int* rows_n = nullptr;
  fr::AdvCoef adv_coef(int k, double pre, double post) const;
};

// poisson (poisson.h:19-81), FFT solver branch, and poisson_IO
// (poisson_IO.h:20-88) for the duct.
class poisson {
 public:
  fr_box_poisson* s = nullptr;
  fr_duct_poisson* sd = nullptr;
  fr_pcg* sp = nullptr;  // Solver = 2: PCG with the FP32 multigrid preconditioner
  double relres_ = 0;
  params* par = nullptr;
  int num_iteration_ = 0;
  hipStream_t st_ = 0;
  poisson() {}
  int init(params* p, gridsize* g, hipStream_t st);
  ~poisson();
  int Solve(tensor1& C, tensor0& P, tensor0& RHS);  // poisson.cpp:115-117
  // Solve with C = 1./Rho_face (grid.cpp:641) without storing C.  rsum
  // (device, box FFT solve only): RHS's interior sum while its mean is still
  // in the field; the residual removes it as make_mean_zero would.
  int Solve_face_rho(tensor1& Rho_face, tensor0& P, tensor0& RHS,
                     const double* rsum = nullptr);  // P: FP64 storage
  // The same (box FFT solve only) when Rho_face = I_C2F(Rho_new), as
  // Compute_Div_U_new leaves it: the residual forms the face values from
  // Rho_new itself (one array read per iteration instead of three).
  int Solve_cell_rho(tensor0& Rho_new, tensor0& P, tensor0& RHS, const double* rsum = nullptr);
  int num_iteration() const { return num_iteration_; }
};

// poisson_IO (poisson_IO.h:20-88): the duct's pressure solve through the
// reference's interface.  Solve(C, P, RHS) solves div(C grad P) = RHS with P
// the initial guess (FFT in y, z and Thomas in x, defect correction to
// max|dp| <= epsilon; or, with Solver = 2, PCG with the FP32 multigrid
// preconditioner, C = 1/Rho_face given as Rho_face through Solve_face_rho).
// num_iteration1() is the outer iteration count; num_iteration2() the mean
// inner iterations per outer one: 0 for the direct FFT + Thomas inner solve
// (the reference leaves it unset on that branch), the PCG count with
// Solver = 2 (one outer iteration).  FFT_TIME and SOLVER_TIME sum seconds
// of the FFT passes (with the residual formed in the first) and of the
// Thomas sweeps, from HIP events.
class poisson_IO {
 public:
  poisson PS;
  double FFT_TIME = 0, SOLVER_TIME = 0;
  int n1_ = 0;
  double n2_ = 0;
  poisson_IO() {}
  int init(params* p, gridsize* g, hipStream_t st);
  int Solve(tensor1& C, tensor0& P, tensor0& RHS);
  int Solve_face_rho(tensor1& Rho_face, tensor0& P, tensor0& RHS);
  int num_iteration1() const { return n1_; }
  double num_iteration2() const { return n2_; }

 private:
  int account();
};

namespace fr {
// First nonzero status of a reference-surface member that returns nothing
// (tensor0/tensor1 members); reset clears it.
int host_status(bool reset = false);
void host_note(int rc);
}  // namespace fr

// grid (grid.h:16-115): the box and duct time-step methods.  For the duct
// the box grid g of the reference's grid::*(grid& g) methods is replaced by
// the inflow planes and inlet particles set with set_inflow.
// The coupling of a periodic box to a duct (simulation.cpp:19-73).  The
// box writes the duct's inputs of each substep: the '>hi' list of its x pass
// (device rows, 29 values each, at most cap; count on the device and in
// pinned host memory) after its migration, and the duct's inflow planes after
// its Update_RU_WP.  Both writes wait for ev_duct, which the duct records
// when a substep no longer reads the previous ones.
struct PairLink {
  double* buf = nullptr;
  int cap = 0;
  int* d_n = nullptr;
  int* h_n = nullptr;
  hipEvent_t ev_rows = nullptr;    // box: rows and count written (its st3)
  hipEvent_t ev_planes = nullptr;  // box: planes written (its st)
  hipEvent_t ev_duct = nullptr;    // duct: substep done with both (its st)
};

class grid {
 public:
  params param;
  gridsize size;
  hipStream_t st = 0;
  double RK4_preCoeff[4], RK4_postCoeff[4];
  double mean_energy_transferred = 0, C_BC = 0;
  particle part;
  tensor1 RU, RU_int, RU_new, RU_np1, RHS_RU, U, Rho_face, dummy;
  tensor0 P, RHS_Pois, C, Rho, Rho_int, Rho_new, Rho_np1, RHS_Part_Temp, scratch, divergence;
  double U_avg_in = 0, U_avg_out = 0;
  double* inflow = nullptr;    // duct: 8 ghosted [Nz][Ny] planes (device)
  // FP32 precision form (box): every field but P and RHS_Pois in FP32 storage
  // and FP32 stencil arithmetic; P, the Poisson right-hand side and residual
  // and all sums in FP64
  bool f32 = false;
  tensor1 d64a, d64b;  // FP64 copies for the diagnostics of the FP32 form
  tensor0 d64r;
  std::vector<double> inlet;   // duct: inlet particles not yet received
  PairLink* pair_out = nullptr;  // box of a coupled run: writes the duct's rows
  PairLink* pair_in = nullptr;   // duct of a coupled run: its inlet rows
  double P0 = 0, P0_int = 0, P0_new = 0, P0_np1 = 0, dP0_dt = 0, T_cur = 0;
  int num_timestep = 0, RK4_count = 0;
  // Steps between re-sorts of the particle storage into cell order.  A
  // re-sort costs about 1 ms at 256^3 / 4M and particles move well under a
  // cell per step, so the order stays close to cell order in between; no
  // result depends on it (the reference order lives in the ranks).  Set per
  // solver (fr_sim_set_resort_interval) or by FR_RESORT_EVERY at creation.
  int resort_every = 4;
  poisson PS_;
  std::string ic_dir;
  // mean of RHS_Part_Temp after its last deposit (tensor0::mean)
  double qmean = 0;
  bool qmean_valid = false;
  int err = 0;  // first nonzero status of a method
  // device timers
  hipEvent_t ev[4] = {};
  hipEvent_t evk[8] = {};  // momentum kernel start/stop per substep
  // second stream (default priority) for the concentration deposit, which is
  // independent of the momentum deposit; fork/join events
  hipStream_t st2 = 0;
  // third stream (default priority): the box migration beside the deposits
  hipStream_t st3 = 0;
  hipEvent_t evm = nullptr, evm0 = nullptr;
  // the momentum deposit's bucket head, issued on st3 at the top of a box
  // substep beside the density update and the advance (vel_head: pending)
  hipEvent_t evh = nullptr;
  bool vel_head = false;
  hipEvent_t evf = nullptr, evj = nullptr, evj2 = nullptr;
  // the heat branch of Update_Particle (gas temperature at the particles,
  // buckets, heat deposit, its mean) runs on st2 beside Update_RU_WOP and is
  // joined by join_side() before anything reads RHS_Part_Temp or qmean
  bool side_pending = false;
  bool side_held = false;  // branch recorded but its launches not yet issued
  // In a full time step of the box (substep_front), Update_Particle runs as
  // a part of the step: the substep's concentration deposit was issued on
  // st2 at the top of the substep, beside the density update and the
  // particle advance (memory-bound kernels beside an issue-bound one) -- its
  // inputs, the rounded-index buckets of the previous heat branch, are final
  // there, and nothing reads C before the step ends -- and the momentum
  // deposit's periodic ghost shell is not formed: its one reader in the
  // step, Update_RU_WOP, reads RHS_RU at interior faces only.  Called on its
  // own, Update_Particle does both as the reference does.
  bool step_front = false;
  // Inside a full time step (begin_step .. end_step) of the box, the last
  // stage's kernels also make grid::TimeAdvance's copies (Rho = Rho_np1 in
  // the density update, RU = RU_np1 in Correct_mean_RU's mean removal,
  // particle stage 0 = stage 3 in the advance), each in the pass that
  // writes the source values; the flags tell TimeAdvance which are made.
  bool in_step = false;
  // RHS_Pois still holds its mean (removed inside the box FFT solve's
  // residual in a time step); fill_rhs_pois() removes it from the field
  bool rhs_mean_pending = false;
  int fill_rhs_pois();
  bool rho_set = false, ru_set = false, part_set = false;
  // particle stage slots 1 and 3 hold slot 0's values (set by a full box
  // step's end, cleared by the next advance and by any outside change)
  bool stages_equal = false;
  // A full box step (not coupled) ends with Rho_np1 = Rho and RU_np1 = RU
  // without writing either: its last stage leaves them out (the density
  // update writes Rho directly; the momentum updates finish RU_np1 in RU_new
  // and take the sums of its means on the way), and the next first stage
  // reads Rho and RU in their place.  The flags say which _np1 field has
  // not been written; fill_np1() writes them before any other reader.
  // ru_lean: from the lean last stage to Correct_mean_RU, RU_new holds the
  // stage's RU_np1.
  bool rho_np1_stale = false, ru_np1_stale = false, ru_lean = false;
  // face_src: Rho_face = I_C2F(Rho_new) (1) or I_C2F(Rho_int) (2) in value,
  // as the box's Compute_RHS_Pois formed it (1; the stage exchange makes it
  // 2), with neither field changed since; 0: only the stored Rho_face.
  // face_stored: Rho_face's memory holds it.  Inside a full box step with
  // the FFT solve and no gravity the faces are not stored: their readers
  // there (the solve's residual, the density update's U = RU_int/Rho_face)
  // form them from the density; fill_faces() stores them for any other.
  int face_src = 0;
  bool face_stored = true;
  int fill_faces();
  bool lean_stage() const { return in_step && param.IsBox() && !pair_out && RK4_count == 3; }
  int fill_np1();  // both; each method fills only the one it reads
  int fill_rho_np1();
  int fill_ru_np1();
  hipEvent_t evc = nullptr;
  int launch_side();
  double* dq = nullptr;  // reduction workspace + result of that branch
  double* dmz = nullptr;  // workspace of the main stream's device-side means
  double* dpart = nullptr;  // box: per-workgroup sums of RHS_Pois (k_div_rhs)
  // join_side() waits for the branch on the device and copies its sum to
  // pinned host memory (hq) behind an event; settle_side() also takes the
  // mean on the host, waiting for that copy only if it has not landed
  double* hq = nullptr;
  hipEvent_t evq = nullptr;
  bool q_inflight = false;
  int join_side();
  int settle_side();
  double t_poisson = 0, t_particles = 0, t_step = 0, t_momentum = 0;
  int n_momentum = 0;
  double t_xpass = 0;  // forward x pass of the box FFT solve: summed ms, launches
  long n_xpass = 0;
  int iters_step = 0;
  double* dred = nullptr;  // device scratch for diagnostics (Nx doubles x 4)
  // ---- run files (host_io.cpp), relative to the working directory as in
  // the reference: outputs enables touch.check at creation and info.txt at
  // initialisation; Statistics writes the stat files when params.in asks
  bool outputs = false;
  int Is_touch = 0;
  tensor1 st_a, st_b, st_c;  // Statistics scratch (U, dummy / dummy3, dummy2)
  tensor0 io_wide[12];       // FP64 copies of FP32 fields for the run files
  double* dprof = nullptr;   // device x-profiles of Statistics (10 x Nx)
  std::vector<double> inflow_h;  // duct: host copy of the last inflow planes
  int open_outputs();        // grid.cpp:25-49
  int Statistics();          // grid.cpp:714-866
  int Store();               // grid.cpp:233-313, particle.cpp:867-917
  int Touch(bool* stop);     // grid.cpp:77-98
  int Restart();             // grid.cpp:124-182, particle.cpp:961-1025
  int Write_info();          // grid.cpp:869-973
  const double* wide(tensor0& f, int slot);
  int download(tensor0& f, std::vector<double>& h, int slot);

  grid() {}
  ~grid();
  int setup(const char* params_path, const char* ic_dir_);
  int reserve(int capacity);  // particle capacity (before Initialize)
  int Initialize();
  int Finish_Initialize();  // grid.cpp:197-229
  // duct inflow: rho_new, rho_np1, RU_new x,y,z, RU_np1 x,y,z planes of the
  // box (grid.cpp:351-364, 482-504) and n inlet particles (29 values each)
  int set_inflow(const double* planes, int n, const double* particles);
  int Correct_OutFlow();  // grid.cpp:579-627
  int Update_Rho();
  int Update_P0();
  int Update_Particle();
  int Update_RU_WOP();
  int Compute_Div_U_new();
  int Compute_RHS_Pois();
  int Solve_Poisson();
  int Update_RU_WP();
  int TimeAdvance_RK4();
  int Correct_mean_RU();
  int TimeAdvance();
  int step();  // simulation.cpp:36-70, box or duct (GRID_IO) calls
  // step() in parts, for the coupled run: the parameter update and timers;
  // one substep's Update_Rho .. Solve_Poisson; the end-of-step wait/timers
  int begin_step();
  int substep_front(int k);
  int end_step();
  std::chrono::steady_clock::time_point t0_step;
  int diagnostics(double* out);
  tensor0* field0(const std::string& name);
  tensor1* field1(const std::string& name);
};
