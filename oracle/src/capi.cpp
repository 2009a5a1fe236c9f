// C entry points of the oracle for the test suite and the CPU baseline in
// bench.py.  TEST INFRASTRUCTURE: only tests/, the smoke check and the CPU
// baseline leg load this library.
#include <algorithm>
#include <cstring>
#include <string>

#include "orc.h"

using namespace orc;

extern "C" {

// Geometry block shared with the product C-ABI (include/fr_amd.h).
typedef struct orc_geom {
  int Nx, Ny, Nz, bs;
  double dx, dy, dz, Lx, Ly, Lz;
  int is_box;
} orc_geom;

static Geom mk(const orc_geom* gg) {
  Geom g;
  g.setup(gg->Nx - 2 * gg->bs, gg->Ny - 2 * gg->bs, gg->Nz - 2 * gg->bs, gg->bs,
          gg->Lx, gg->Ly, gg->Lz, gg->is_box != 0);
  return g;
}
static void load(F0& f, const double* p) { std::memcpy(f.v.data(), p, f.v.size() * 8); }
static void store(const F0& f, double* p) { std::memcpy(p, f.v.data(), f.v.size() * 8); }

// ---- tensor0/tensor1 operators (each includes its reference halo) ---------
void orc_div_f2c(const orc_geom* gg, double* out, const double* tx,
                 const double* ty, const double* tz) {
  Geom g = mk(gg);
  F1 T;
  T.init(&g);
  load(T.x, tx);
  load(T.y, ty);
  load(T.z, tz);
  F0 o;
  o.init(&g);
  load(o, out);
  o.Equal_Div_F2C(T);
  store(o, out);
}
void orc_del2(const orc_geom* gg, double* out, const double* in) {
  Geom g = mk(gg);
  F0 a, o;
  a.init(&g);
  o.init(&g);
  load(a, in);
  load(o, out);
  o.Equal_Del2(a);
  store(o, out);
}
void orc_grad_c2f(const orc_geom* gg, double* ox, double* oy, double* oz,
                  const double* in) {
  Geom g = mk(gg);
  F0 a;
  a.init(&g);
  load(a, in);
  F1 o;
  o.init(&g);
  load(o.x, ox);
  load(o.y, oy);
  load(o.z, oz);
  o.Equal_Grad_C2F(a);
  store(o.x, ox);
  store(o.y, oy);
  store(o.z, oz);
}
// axis 0/1/2, dir 0 = C2F, 1 = F2C
void orc_interp(const orc_geom* gg, double* out, const double* in, int axis,
                int dir) {
  Geom g = mk(gg);
  F0 a, o;
  a.init(&g);
  o.init(&g);
  load(a, in);
  load(o, out);
  if (dir == 0) {
    if (axis == 0) o.Equal_Ix_C2F(a);
    if (axis == 1) o.Equal_Iy_C2F(a);
    if (axis == 2) o.Equal_Iz_C2F(a);
  } else {
    if (axis == 0) o.Equal_Ix_F2C(a);
    if (axis == 1) o.Equal_Iy_F2C(a);
    if (axis == 2) o.Equal_Iz_F2C(a);
  }
  store(o, out);
}
void orc_i_f2c(const orc_geom* gg, double* out, const double* tx,
               const double* ty, const double* tz) {
  Geom g = mk(gg);
  F1 T;
  T.init(&g);
  load(T.x, tx);
  load(T.y, ty);
  load(T.z, tz);
  F0 o;
  o.init(&g);
  load(o, out);
  o.Equal_I_F2C(T);
  store(o, out);
}
void orc_update_ghosts(const orc_geom* gg, double* f, int cumulative) {
  Geom g = mk(gg);
  F0 a;
  a.init(&g);
  load(a, f);
  if (cumulative)
    a.Update_Ghosts_CUM();
  else
    a.Update_Ghosts();
  store(a, f);
}
// Element-wise forms (tensor0.cpp:93-241) under the op codes of
// fr_t0_elementwise (include/fr_amd.h); NULL operands are absent.
void orc_elementwise(const orc_geom* gg, int op, double* out, double a, const double* A,
                     double b, const double* B, double c, const double* Cc) {
  Geom g = mk(gg);
  F0 o, fa, fb, fc;
  o.init(&g);
  fa.init(&g);
  fb.init(&g);
  fc.init(&g);
  load(o, out);
  if (A) load(fa, A);
  if (B) load(fb, B);
  if (Cc) load(fc, Cc);
  switch (op) {
    case 0:
      if (B && Cc) o.Equal_LinComb(a, fa, b, fb, c, fc);
      else if (B) o.Equal_LinComb(a, fa, b, fb);
      else o.Equal_Mult(a, fa);
      break;
    case 1: o.PlusEqual_Mult(a, fa); break;
    case 2: o.Equal_LinComb(fa, fb, b, fc); break;
    case 3: o.Equal_Divide(a, fa); break;
    case 4: o.Equal_Divide(fa, fb); break;
    case 5: o.set(a); break;
    case 6: o.add(a); break;
    case 7: o.mul(a); break;
    case 8: o.copy(fa); break;
    case 9: o.mul(fa); break;
    case 10: o.sub(fa); break;
    case 11: o.add(fa); break;
  }
  store(o, out);
}
// op: 0 sum, 1 mean, 2 max, 3 min, 4 max_abs, 5 sum_squares
double orc_reduce(const orc_geom* gg, const double* f, int op) {
  Geom g = mk(gg);
  F0 a;
  a.init(&g);
  load(a, f);
  switch (op) {
    case 0: return a.sum();
    case 1: return a.mean();
    case 2: return a.max();
    case 3: return a.min();
    case 4: return a.max_abs();
    case 5: return a.sum_squares();
  }
  return 0;
}
void orc_reducex(const orc_geom* gg, const double* f, double* out) {
  Geom g = mk(gg);
  F0 a;
  a.init(&g);
  load(a, f);
  a.reduceX();
  std::memcpy(out, a.red.data(), a.red.size() * 8);
}

// ---- FFT and pressure solves -----------------------------------------------
void orc_fft3d(int nx, int ny, int nz, double* re, double* im, int sign) {
  size_t n = (size_t)nx * ny * nz;
  std::vector<double> r(re, re + n), i(im, im + n);
  fft3d(r, i, nx, ny, nz, sign);
  std::memcpy(re, r.data(), n * 8);
  std::memcpy(im, i.data(), n * 8);
}
void orc_fft2d(int ny, int nz, double* re, double* im, int sign) {
  fft2d(re, im, ny, nz, sign);
}
// PCG with the FP32 multigrid preconditioner on div(C grad P) = RHS with
// C = 1/Rho_face (face densities rfx, rfy, rfz).  Returns the iteration count;
// *relres = ||r||/||RHS|| at exit.
int orc_pcg(const orc_geom* gg, const double* rfx, const double* rfy, const double* rfz,
            double* P, const double* RHS, int maxit, double tol, double* relres) {
  Geom g = mk(gg);
  F1 rf;
  rf.init(&g);
  load(rf.x, rfx);
  load(rf.y, rfy);
  load(rf.z, rfz);
  F0 p, r;
  p.init(&g);
  r.init(&g);
  load(p, P);
  load(r, RHS);
  MGPCG s;
  int it = s.solve(rf, p, r, maxit, tol, relres);
  std::memcpy(P, p.v.data(), p.v.size() * 8);
  return it;
}
void orc_rfft1d(const double* x, int n, double* Xre, double* Xim) { rfft1d(x, n, Xre, Xim); }
void orc_irfft1d(const double* Xre, const double* Xim, int n, double* x) {
  irfft1d(Xre, Xim, n, x);
}
// Box (is_box=1, poisson.cpp:391-462) or duct (is_box=0, poisson_IO.cpp:
// 81-452) defect-correction solve of div(C grad P) = RHS.  Returns the
// iteration count.
int orc_poisson(const orc_geom* gg, const double* cx, const double* cy,
                const double* cz, double* P, const double* RHS, int iterations,
                double epsilon) {
  Geom g = mk(gg);
  Params par;
  par.Iteration = iterations;
  par.epsilon = epsilon;
  F1 C;
  C.init(&g);
  load(C.x, cx);
  load(C.y, cy);
  load(C.z, cz);
  F0 p, r;
  p.init(&g);
  r.init(&g);
  load(p, P);
  load(r, RHS);
  int it;
  if (g.box) {
    BoxPoisson s;
    s.init(&par, &g);
    s.Solve(C, p, r);
    it = s.num_iteration;
  } else {
    DuctPoisson s;
    s.init(&par, &g);
    s.Solve(C, p, r);
    it = s.num_iteration1;
  }
  store(p, P);
  return it;
}
// One V-cycle of the PCG preconditioner applied to the interior residual r
// ([Nz][Ny][Nx] floats) with e = 0 on entry, the hierarchy cut at `levels`
// (0: the solver's own), pre/post red-black sweeps and coarse_sweeps on the
// coarsest level; e receives the fine-level correction.
void orc_mg_vcycle(const orc_geom* gg, const double* rfx, const double* rfy, const double* rfz,
                   const float* r, float* e, int levels, int pre, int post, int coarse_sweeps) {
  Geom g = mk(gg);
  F1 C;
  C.init(&g);
  load(C.x, rfx);
  load(C.y, rfy);
  load(C.z, rfz);
  MGPCG m;
  m.max_levels = levels;
  m.pre = pre;
  m.post = post;
  m.coarse_sweeps = coarse_sweeps;
  m.build(&g);
  m.coefficients(C);
  std::copy(r, r + m.L[0].r.size(), m.L[0].r.begin());
  m.vcycle(0);
  std::copy(m.L[0].e.begin(), m.L[0].e.end(), e);
}
void orc_glibc_rand(int n, int* out) {
  GlibcRand r;
  r.seed(1);
  for (int i = 0; i < n; i++) out[i] = r.next();
}

// ---- whole grid (box) ---------------------------------------------------------
void* orc_grid_create(const char* params_path) {
  Params p;
  if (!p.parse(params_path)) return nullptr;
  Grid* g = new Grid();
  g->setup(p);
  return g;
}
// The same with the reference's run files (touch.check and the stat files
// now, info.txt at initialisation), relative to the working directory.
void* orc_grid_create_ex(const char* params_path, int outputs) {
  Params p;
  if (!p.parse(params_path)) return nullptr;
  Grid* g = new Grid();
  g->setup(p);
  g->outputs = outputs != 0;
  if (g->outputs) g->open_outputs();
  return g;
}
void orc_grid_destroy(void* h) { delete (Grid*)h; }
// grid::Statistics, grid::Store (returns 0 on success, -1 when a file could
// not be written, e.g. a particle file that exists), grid::Touch (returns
// the stop flag), grid::Write_info
void orc_grid_statistics(void* h) { ((Grid*)h)->Statistics(); }
int orc_grid_store(void* h) { return ((Grid*)h)->Store() ? 0 : -1; }
int orc_grid_touch(void* h) { return ((Grid*)h)->Touch() ? 1 : 0; }
void orc_grid_write_info(void* h) { ((Grid*)h)->Write_info(); }
int orc_grid_num_timestep_set(void* h, int n) {
  ((Grid*)h)->num_timestep = n;
  return 0;
}
void orc_grid_initialize(void* h, const double* U, const double* V,
                         const double* W) {
  ((Grid*)h)->Initialize(U, V, W);
}
void orc_grid_finish_initialize(void* h) { ((Grid*)h)->Finish_Initialize(); }
void orc_grid_step(void* h) {
  Grid* g = (Grid*)h;
  if (g->par.IsBox)
    g->box_step();
  else
    g->duct_step();
}
// The coupled box `b` and duct `d` (simulation.cpp:19-73): initialise both
// (each from its own U, V, W blocks, or nullptr where Initial != 0), one step
// of both, and the shared touch.check stop (returns the flag).
void orc_pair_initialize(void* b, void* d, const double* Ub, const double* Vb,
                         const double* Wb, const double* Ud, const double* Vd,
                         const double* Wd) {
  const double* const ib[3] = {Ub, Vb, Wb};
  const double* const id[3] = {Ud, Vd, Wd};
  pair_initialize(*(Grid*)b, *(Grid*)d, ib, id);
}
void orc_pair_step(void* b, void* d) { pair_step(*(Grid*)b, *(Grid*)d); }
int orc_pair_touch(void* b, void* d) { return pair_touch(*(Grid*)b, *(Grid*)d) ? 1 : 0; }
// Duct inflow: planes = 8 ghosted [Nz][Ny] planes in the order rho_new,
// rho_np1, ru_new x,y,z, ru_np1 x,y,z; pbuf = n_in particles x 29 values
// (appended to any not yet consumed).
void orc_grid_set_inflow(void* h, const double* planes, int n_in, const double* pbuf) {
  Grid* g = (Grid*)h;
  const size_t m = (size_t)g->g.Ny * g->g.Nz;
  std::vector<double>* dst[8] = {&g->in.rho_new,   &g->in.rho_np1,   &g->in.ru_new[0],
                                 &g->in.ru_new[1], &g->in.ru_new[2], &g->in.ru_np1[0],
                                 &g->in.ru_np1[1], &g->in.ru_np1[2]};
  if (planes)
    for (int q = 0; q < 8; q++) dst[q]->assign(planes + q * m, planes + (q + 1) * m);
  if (pbuf && n_in > 0) g->in.pbuf.insert(g->in.pbuf.end(), pbuf, pbuf + (size_t)n_in * 29);
}
// Run one named grid method (the names of grid.h:93-114).
int orc_grid_call(void* h, const char* name, int rk4_count) {
  Grid* g = (Grid*)h;
  std::string n(name);
  g->RK4_count = rk4_count;
  const bool duct = !g->par.IsBox;
  if (n == "update") g->par.update(g->T_cur);
  else if (n == "Update_Rho") duct ? g->Update_Rho_duct() : g->Update_Rho();
  else if (n == "Update_P0") g->Update_P0();
  else if (n == "Update_Particle") duct ? g->Update_Particle_duct() : g->Update_Particle();
  else if (n == "Update_RU_WOP") duct ? g->Update_RU_WOP_duct() : g->Update_RU_WOP();
  else if (n == "Correct_OutFlow" && duct) g->Correct_OutFlow();
  else if (n == "Finish_Initialize") g->Finish_Initialize();
  else if (n == "Solve_Poisson_IO" && duct) g->Solve_Poisson_IO();
  else if (n == "Compute_Div_U_new") g->Compute_Div_U_new();
  else if (n == "Compute_RHS_Pois") g->Compute_RHS_Pois();
  else if (n == "Solve_Poisson") g->Solve_Poisson();
  else if (n == "Update_RU_WP") g->Update_RU_WP();
  else if (n == "TimeAdvance_RK4") g->TimeAdvance_RK4();
  else if (n == "Correct_mean_RU") g->Correct_mean_RU();
  else if (n == "TimeAdvance") g->TimeAdvance();
  else if (n == "Send_Recv") duct ? g->part.Send_Recv_duct(g->in.pbuf) : g->part.Send_Recv();
  else return -1;
  return 0;
}
// Field access: comp 0..2 selects a tensor1 component, -1 a tensor0.
double* orc_grid_field(void* h, const char* name, int comp) {
  Grid* g = (Grid*)h;
  if (comp < 0) {
    F0* f = g->field0(name);
    return f ? f->v.data() : nullptr;
  }
  F1* f = g->field1(name);
  if (!f) return nullptr;
  return comp == 0 ? f->x.v.data() : comp == 1 ? f->y.v.data() : f->z.v.data();
}
// scalars: P0, P0_int, P0_new, P0_np1, T_cur, C_BC, U_avg_in, U_avg_out
double* orc_grid_scalar(void* h, const char* name) {
  Grid* g = (Grid*)h;
  std::string n(name);
  if (n == "P0") return &g->P0;
  if (n == "P0_int") return &g->P0_int;
  if (n == "P0_new") return &g->P0_new;
  if (n == "P0_np1") return &g->P0_np1;
  if (n == "T_cur") return &g->T_cur;
  if (n == "C_BC") return &g->C_BC;
  if (n == "U_avg_in") return &g->U_avg_in;
  if (n == "U_avg_out") return &g->U_avg_out;
  return nullptr;
}
int orc_grid_np(void* h) { return ((Grid*)h)->part.Np; }
// The '>hi' list of the last migration's x pass (29 values per particle):
// the box's Rbuf1 on one rank (particle.cpp:821, 836-838); *n its length.
const double* orc_grid_sbuf0(void* h, int* n) {
  Grid* g = (Grid*)h;
  *n = (int)(g->part.Sbuf0.size() / 29);
  return g->part.Sbuf0.data();
}
void orc_grid_set_np(void* h, int np) {
  Grid* g = (Grid*)h;
  g->part.Np = np;
  g->part.RESIZE();
}
// Particle stage arrays: q 0..6 = x,y,z,u,v,w,T; s 0..3 = base,_int,_new,_np1;
// q = 7 returns RHS_T (s ignored).
double* orc_grid_part(void* h, int q, int s) {
  Grid* g = (Grid*)h;
  if (q == 7) return g->part.RHS_T.data();
  return g->part.stage(q, s)->data();
}
int* orc_grid_flag(void* h) { return ((Grid*)h)->part.flag.data(); }
void orc_grid_diagnostics(void* h, double* out) { ((Grid*)h)->diagnostics(out); }
int orc_grid_num_timestep(void* h) { return ((Grid*)h)->num_timestep; }
void orc_grid_geom(void* h, orc_geom* out) {
  Grid* g = (Grid*)h;
  out->Nx = g->g.Nx;
  out->Ny = g->g.Ny;
  out->Nz = g->g.Nz;
  out->bs = g->g.bs;
  out->dx = g->g.dx;
  out->dy = g->g.dy;
  out->dz = g->g.dz;
  out->Lx = g->g.Lx;
  out->Ly = g->g.Ly;
  out->Lz = g->g.Lz;
  out->is_box = g->g.box;
}

// Particle physics on the grid's particle set, same op codes as the partial
// reference build's harness (oracle/ref_harness.cpp ref_part_op).
void orc_grid_part_op(void* h, int op, int k, double pre, double post,
                      double* f0, double* f1, double* f2) {
  Grid* gr = (Grid*)h;
  Particles& p = gr->part;
  const Geom* g = &gr->g;
  F1 U;
  F0 A;
  U.init(g);
  A.init(g);
  switch (op) {
    case 0:
      load(U.x, f0); load(U.y, f1); load(U.z, f2);
      p.gas2part_velocity(U);
      break;
    case 1: p.update_position(k, pre, post); break;
    case 2: p.update_velocity(k, pre, post); break;
    case 3: p.update_Temp(k, pre, post); break;
    case 4: load(A, f0); p.gas2part_Temp(A, true); break;
    case 5: p.Compute_RHS_Temp(true); break;
    case 6: p.part2gas_concentration(A); store(A, f0); break;
    case 7:
      p.part2gas_velocity(U);
      store(U.x, f0); store(U.y, f1); store(U.z, f2);
      break;
    case 8: p.part2gas_Temp(A, true); store(A, f0); break;
    case 9: p.Send_Recv(); break;
    case 10: load(A, f0); p.gas2part_Temp(A, false); break;
    case 11: p.Compute_RHS_Temp(false); break;
    case 12: p.part2gas_Temp(A, false); store(A, f0); break;
    case 13: p.load_random(); break;
  }
}
// Gathered values and momentum RHS: which 0..3 = ug,vg,wg,Tg; 4..6 = RHS_u,v,w.
double* orc_grid_part_aux(void* h, int which) {
  Particles& p = ((Grid*)h)->part;
  std::vector<double>* t[7] = {&p.ug, &p.vg, &p.wg, &p.Tg, &p.RHS_u, &p.RHS_v, &p.RHS_w};
  return t[which]->data();
}
}


// ---- post-process operators of grid::Statistics (same op codes as the
// partial reference build's ref_post; RU, U and R reduced first for 7-11)
extern "C" void orc_post(const orc_geom* gg, int op, double* o0, double* o1, double* o2,
                         double* red, const double* const* U3, const double* const* RU3,
                         const double* a, const double* b, double s) {
  Geom g = mk(gg);
  F1 U, RU, O;
  F0 A, B;
  U.init(&g);
  RU.init(&g);
  O.init(&g);
  A.init(&g);
  B.init(&g);
  const int nx = g.Nx;
  if (U3) {
    load(U.x, U3[0]);
    load(U.y, U3[1]);
    load(U.z, U3[2]);
  }
  if (RU3) {
    load(RU.x, RU3[0]);
    load(RU.y, RU3[1]);
    load(RU.z, RU3[2]);
  }
  if (a) load(A, a);
  if (b) load(B, b);
  load(O.x, o0);
  if (op >= 7) {
    load(O.y, o1);
    load(O.z, o2);
  }
  switch (op) {
    case 0: O.x.Equal_SijSij(U); break;
    case 1: O.x.Equal_WijWij(U); break;
    case 2: O.x.Equal_eta(A, B); break;
    case 3: O.x.Equal_ST(s, A); break;
    case 4: O.x.Equal_Reduced_sd(A); break;
    case 5: O.x.Equal_nu(s, A); break;
    case 6: O.x.Equal_dissipation(A, B); break;
    default:
      A.reduceX();
      if (op <= 8) RU.reduceX(); else U.reduceX();
      if (op == 7) O.Equal_TKE2(RU, U, A);
      if (op == 8) O.Equal_UiUi2(RU, U, A);
      if (op == 9) O.Equal_Cor_Rho_U(U, A);
      if (op == 10) O.Equal_TKE(U, A);
      if (op == 11) O.Equal_UiUi(U, A);
  }
  store(O.x, o0);
  if (red) std::memcpy(red, O.x.red.data(), sizeof(double) * nx);
  if (op >= 7) {
    store(O.y, o1);
    store(O.z, o2);
    if (red) {
      std::memcpy(red + nx, O.y.red.data(), sizeof(double) * nx);
      std::memcpy(red + 2 * nx, O.z.red.data(), sizeof(double) * nx);
    }
  }
}

extern "C" void orc_post_scalars(const orc_geom* gg, const double* const* U3, double dt,
                                 double* out3) {
  Geom g = mk(gg);
  F1 U;
  U.init(&g);
  load(U.x, U3[0]);
  load(U.y, U3[1]);
  load(U.z, U3[2]);
  out3[0] = U.max();
  out3[1] = U.min();
  out3[2] = U.max_cfl(dt);
}
