// Harness around the reference's own tensor0/tensor1/communicator/particle
// translation units (compiled from /root/reference by oracle/build_ref.sh,
// never copied).  It runs them on one MPI rank so the oracle's restatement can
// be checked against the real code.  TEST INFRASTRUCTURE.
#include <mpi.h>

#include <cstring>
#include <vector>

#include "communicator.h"
#include "gridsize.h"
#include "params.h"
#include "particle.h"
#include "proc.h"
#include "tensor0.h"
#include "tensor1.h"

struct RefCtx {
  params* par;
  proc* pc;
  gridsize* gs;
  communicator* com;
  particle* part;
};

static void ensure_mpi() {
  int inited = 0;
  MPI_Initialized(&inited);
  if (!inited) MPI_Init(nullptr, nullptr);
}

static void load(tensor0& t, const double* p) {
  std::memcpy(t.pointer(), p, sizeof(double) * t.size());
}
static void store(tensor0& t, double* p) {
  std::memcpy(p, t.pointer(), sizeof(double) * t.size());
}

extern "C" {

void* ref_create(const char* params_path) {
  ensure_mpi();
  RefCtx* c = new RefCtx();
  c->par = new params((char*)params_path);
  c->pc = new proc();
  c->gs = new gridsize(c->par, c->pc);
  c->com = new communicator(c->gs, c->par, c->pc);
  c->part = new particle(c->par, c->pc, c->gs);
  return c;
}

int ref_size(void* h) { return ((RefCtx*)h)->gs->size(); }

// Tensor operators: op 0 Equal_Div_F2C(a,b,c), 1 Equal_Del2(a), 2
// Equal_Grad_C2F(a) -> out,o2,o3, 3..8 Equal_{Ix,Iy,Iz}_C2F/_F2C(a), 9
// Equal_I_F2C(a,b,c), 10 Update_Ghosts, 11 Update_Ghosts_CUM.
void ref_op(void* h, int op, double* out, double* o2, double* o3,
            const double* a, const double* b, const double* cc) {
  RefCtx* c = (RefCtx*)h;
  tensor0 O(c->gs, c->com), A(c->gs, c->com);
  tensor1 T(c->gs, c->com), G(c->gs, c->com);
  load(O, out);
  if (a) load(A, a);
  switch (op) {
    case 0:
    case 9:
      load(T.x, a);
      load(T.y, b);
      load(T.z, cc);
      if (op == 0) O.Equal_Div_F2C(T); else O.Equal_I_F2C(T);
      break;
    case 1: O.Equal_Del2(A); break;
    case 2:
      load(G.x, out);
      load(G.y, o2);
      load(G.z, o3);
      G.Equal_Grad_C2F(A);
      store(G.x, out);
      store(G.y, o2);
      store(G.z, o3);
      return;
    case 3: O.Equal_Ix_C2F(A); break;
    case 4: O.Equal_Iy_C2F(A); break;
    case 5: O.Equal_Iz_C2F(A); break;
    case 6: O.Equal_Ix_F2C(A); break;
    case 7: O.Equal_Iy_F2C(A); break;
    case 8: O.Equal_Iz_F2C(A); break;
    case 10: O.Update_Ghosts(); break;
    case 11: O.Update_Ghosts_CUM(); break;
  }
  store(O, out);
}

// Reductions: 0 sum, 1 mean, 2 max, 3 min, 4 max_abs, 5 sum_squares.
double ref_reduce(void* h, const double* f, int op) {
  RefCtx* c = (RefCtx*)h;
  tensor0 A(c->gs, c->com);
  load(A, f);
  switch (op) {
    case 0: return A.sum();
    case 1: return A.mean();
    case 2: return A.max();
    case 3: return A.min();
    case 4: return A.max_abs();
    case 5: return A.sum_squares();
  }
  return 0;
}

void ref_reducex(void* h, const double* f, double* out) {
  RefCtx* c = (RefCtx*)h;
  tensor0 A(c->gs, c->com);
  load(A, f);
  A.reduceX();
  std::memcpy(out, A.ptr_reduce_, sizeof(double) * c->gs->Nx());
}

// ---- particles -------------------------------------------------------------
static std::vector<double>* stage(particle* p, int q, int s) {
  std::vector<double>* tab[7][4] = {
      {&p->x, &p->x_int, &p->x_new, &p->x_np1}, {&p->y, &p->y_int, &p->y_new, &p->y_np1},
      {&p->z, &p->z_int, &p->z_new, &p->z_np1}, {&p->u, &p->u_int, &p->u_new, &p->u_np1},
      {&p->v, &p->v_int, &p->v_new, &p->v_np1}, {&p->w, &p->w_int, &p->w_new, &p->w_np1},
      {&p->T, &p->T_int, &p->T_new, &p->T_np1}};
  return tab[q][s];
}

// Set Np particles from 28 stage arrays laid out [q*4+s][Np] plus flags,
// RHS_T, and make every vector Np long.
void ref_part_set(void* h, int np, const double* stages, const int* flags,
                  const double* rhsT) {
  particle* p = ((RefCtx*)h)->part;
  p->Np = np;
  for (int q = 0; q < 7; q++)
    for (int s = 0; s < 4; s++)
      stage(p, q, s)->assign(stages + (size_t)(q * 4 + s) * np,
                             stages + (size_t)(q * 4 + s + 1) * np);
  p->flag.assign(flags, flags + np);
  p->RHS_u.assign(np, 0.0);
  p->RHS_v.assign(np, 0.0);
  p->RHS_w.assign(np, 0.0);
  p->RHS_T.assign(rhsT, rhsT + np);
  p->ug.assign(np, 0.0);
  p->vg.assign(np, 0.0);
  p->wg.assign(np, 0.0);
  p->Tg.assign(np, 0.0);
}
int ref_part_np(void* h) { return ((RefCtx*)h)->part->Np; }
void ref_part_get(void* h, double* stages, int* flags, double* rhsT,
                  double* rhs_uvw, double* g_uvwT) {
  particle* p = ((RefCtx*)h)->part;
  int np = p->Np;
  for (int q = 0; q < 7; q++)
    for (int s = 0; s < 4; s++)
      std::memcpy(stages + (size_t)(q * 4 + s) * np, stage(p, q, s)->data(), 8 * np);
  std::memcpy(flags, p->flag.data(), 4 * np);
  std::memcpy(rhsT, p->RHS_T.data(), 8 * np);
  if (rhs_uvw) {
    std::memcpy(rhs_uvw, p->RHS_u.data(), 8 * np);
    std::memcpy(rhs_uvw + np, p->RHS_v.data(), 8 * np);
    std::memcpy(rhs_uvw + 2 * np, p->RHS_w.data(), 8 * np);
  }
  if (g_uvwT) {
    std::memcpy(g_uvwT, p->ug.data(), 8 * np);
    std::memcpy(g_uvwT + np, p->vg.data(), 8 * np);
    std::memcpy(g_uvwT + 2 * np, p->wg.data(), 8 * np);
    std::memcpy(g_uvwT + 3 * np, p->Tg.data(), 8 * np);
  }
}
// Particle physics on the set state.  op: 0 gas2part_velocity(U), 1
// update_position, 2 update_velocity, 3 update_Temp, 4 gas2part_Temp_new(T),
// 5 Compute_RHS_Temp_new, 6 part2gas_concentration -> f0, 7
// part2gas_velocity -> f0,f1,f2, 8 part2gas_Temp_new -> f0, 9 Send_Recv,
// 10 gas2part_Temp_int(T), 11 Compute_RHS_Temp_int, 12 part2gas_Temp_int.
void ref_part_op(void* h, int op, int k, double pre, double post, double* f0,
                 double* f1, double* f2) {
  RefCtx* c = (RefCtx*)h;
  particle* p = c->part;
  tensor1 U(c->gs, c->com);
  tensor0 A(c->gs, c->com);
  switch (op) {
    case 0:
      load(U.x, f0);
      load(U.y, f1);
      load(U.z, f2);
      p->gas2part_velocity(U);
      break;
    case 1: p->update_position(k, pre, post); break;
    case 2: p->update_velocity(k, pre, post); break;
    case 3: p->update_Temp(k, pre, post); break;
    case 4:
      load(A, f0);
      p->gas2part_Temp_new(A);
      break;
    case 5: p->Compute_RHS_Temp_new(); break;
    case 6:
      p->part2gas_concentration(A);
      store(A, f0);
      break;
    case 7:
      p->part2gas_velocity(U);
      store(U.x, f0);
      store(U.y, f1);
      store(U.z, f2);
      break;
    case 8:
      p->part2gas_Temp_new(A);
      store(A, f0);
      break;
    case 9: p->Send_Recv(); break;
    case 10:
      load(A, f0);
      p->gas2part_Temp_int(A);
      break;
    case 11: p->Compute_RHS_Temp_int(); break;
    case 12:
      p->part2gas_Temp_int(A);
      store(A, f0);
      break;
  }
}
// particle::load_random on one rank (glibc rand, default seed).
void ref_part_load_random(void* h) { ((RefCtx*)h)->part->load_random(); }

// particle::Send_Recv(particle& pbox) (particle.cpp:860-865, 814-851) on a
// duct rank that is both inlet and outlet.  A second particle object stands
// in for the box grid's set: its Rbuf1/Rnum1 hold the n_in particles (29
// values each, ToBuf order) that entered through the box's left face, which
// the duct appends after its own z, y and x sorting.
void ref_part_send_recv_duct(void* h, int n_in, const double* inbuf) {
  RefCtx* c = (RefCtx*)h;
  particle pbox(c->par, c->pc, c->gs);
  pbox.Rnum1 = 29 * n_in;
  if (n_in > 0) std::memcpy(pbox.Rbuf1, inbuf, sizeof(double) * 29 * (size_t)n_in);
  c->part->Send_Recv(pbox);
}

// The particle exchange of one coupled substep: GRID.Update_Particle's
// part.Send_Recv() (grid.cpp:672), then GRID_IO.Update_Particle(GRID)'s
// part.Send_Recv(g.part) (grid.cpp:696), which reads the box's Rbuf1/Rnum1.
void ref_part_send_recv_pair(void* hbox, void* hduct) {
  RefCtx* b = (RefCtx*)hbox;
  RefCtx* d = (RefCtx*)hduct;
  b->part->Send_Recv();
  d->part->Send_Recv(*b->part);
}
}

// Element-wise tensor0 operators (tensor0.cpp:93-241) under the op codes of
// fr_t0_elementwise: 0 Equal_LinComb(a,A,b,B,c,C) / Equal_LinComb(a,A,b,B) /
// Equal_Mult(a,A) as B and C are given, 1 PlusEqual_Mult(a,A), 2
// Equal_LinComb(A,B,b,C), 3 Equal_Divide(a,A), 4 Equal_Divide(A,B), 5 =a,
// 6 +=a, 7 *=a, 8 =A, 9 *=A, 10 -=A, 11 +=A.
extern "C" void ref_elementwise(void* h, int op, double* out, double a, const double* A,
                                double b, const double* B, double c, const double* Cc) {
  RefCtx* ctx = (RefCtx*)h;
  tensor0 O(ctx->gs, ctx->com), TA(ctx->gs, ctx->com), TB(ctx->gs, ctx->com),
      TC(ctx->gs, ctx->com);
  load(O, out);
  if (A) load(TA, A);
  if (B) load(TB, B);
  if (Cc) load(TC, Cc);
  switch (op) {
    case 0:
      if (B && Cc) O.Equal_LinComb(a, TA, b, TB, c, TC);
      else if (B) O.Equal_LinComb(a, TA, b, TB);
      else O.Equal_Mult(a, TA);
      break;
    case 1: O.PlusEqual_Mult(a, TA); break;
    case 2: O.Equal_LinComb(TA, TB, b, TC); break;
    case 3: O.Equal_Divide(a, TA); break;
    case 4: O.Equal_Divide(TA, TB); break;
    case 5: O = a; break;
    case 6: O += a; break;
    case 7: O *= a; break;
    case 8: O = TA; break;
    case 9: O *= TA; break;
    case 10: O -= TA; break;
    case 11: O += TA; break;
  }
  store(O, out);
}

// Post-process operators of grid::Statistics (tensor0.cpp:575-663,
// tensor1.cpp:318-439).  tensor0 forms (out holds the prior content, red
// receives ptr_reduce_): 0 Equal_SijSij(U), 1 Equal_WijWij(U), 2
// Equal_eta(a, b), 3 Equal_ST(s, a), 4 Equal_Reduced_sd(a), 5 Equal_nu(s, a),
// 6 Equal_dissipation(a, b).  tensor1 forms (three outputs and their
// ptr_reduce_): 7 Equal_TKE2(RU, U, R), 8 Equal_UiUi2(RU, U, R), 9
// Equal_Cor_Rho_U(U, R), 10 Equal_TKE(U, R), 11 Equal_UiUi(U, R); RU, U and R
// are reduced first, as grid::Statistics has them.  ref_post_scalars returns
// tensor1::max, min and max_cfl(dt) of U.
extern "C" void ref_post(void* h, int op, double* o0, double* o1, double* o2, double* red,
                         const double* const* U3, const double* const* RU3, const double* a,
                         const double* b, double s) {
  RefCtx* c = (RefCtx*)h;
  tensor1 U(c->gs, c->com), RU(c->gs, c->com), O(c->gs, c->com);
  tensor0 A(c->gs, c->com), B(c->gs, c->com);
  const int nx = c->gs->Nx();
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
  if (red) std::memcpy(red, O.x.ptr_reduce_, sizeof(double) * nx);
  if (op >= 7) {
    store(O.y, o1);
    store(O.z, o2);
    if (red) {
      std::memcpy(red + nx, O.y.ptr_reduce_, sizeof(double) * nx);
      std::memcpy(red + 2 * nx, O.z.ptr_reduce_, sizeof(double) * nx);
    }
  }
}

extern "C" void ref_post_scalars(void* h, const double* const* U3, double dt, double* out3) {
  RefCtx* c = (RefCtx*)h;
  tensor1 U(c->gs, c->com);
  load(U.x, U3[0]);
  load(U.y, U3[1]);
  load(U.z, U3[2]);
  out3[0] = U.max();
  out3[1] = U.min();
  out3[2] = U.max_cfl(dt);
}
