// Oracle grid: the box time-step methods and the step diagnostics.
// TEST INFRASTRUCTURE.  Restates grid.cpp (box calls) and simulation.cpp.
#include <cstring>

#include "orc.h"

namespace orc {

void Grid::setup(const Params& p) {
  par = p;
  g.setup(par.Nx_tot, par.Ny_tot, par.Nz_tot, par.bs, par.Lx, par.Ly, par.Lz,
          par.IsBox != 0);
  // grid.cpp:16-23
  RK4_pre[0] = 0.5;
  RK4_pre[1] = 0.5;
  RK4_pre[2] = 1.;
  RK4_pre[3] = 1. / 6.;
  RK4_post[0] = 1. / 6.;
  RK4_post[1] = 1. / 3.;
  RK4_post[2] = 1. / 3.;
  RK4_post[3] = 1. / 6.;
  part.init(&par, &g);
  for (F1* f : {&RU, &RU_int, &RU_new, &RU_np1, &RHS_RU, &U, &Rho_face, &dummy,
                &dummy2, &dummy3})
    f->init(&g);
  for (F0* f : {&P, &RHS_Pois, &C, &Rho, &Rho_int, &Rho_new, &Rho_np1, &RHS_Rho,
                &T, &divergence, &RHS_Part_Temp})
    f->init(&g);
  PS.init(&par, &g);
  PS_IO.init(&par, &g);
}

// grid.cpp:184-196 (Initial=0; the three blocks are the global interior
// [Nz][Ny][Nx] layout of U.bin/V.bin/W.bin) and grid.cpp:112-122 (Initial=2).
void Grid::Initialize(const double* Ub, const double* Vb, const double* Wb) {
  const int bs = g.bs;
  if (par.Initial == 0 && Ub && Vb && Wb) {
    const double* blk[3] = {Ub, Vb, Wb};
    F0* dst[3] = {&RU.x, &RU.y, &RU.z};
    for (int c = 0; c < 3; c++) {
      size_t n = 0;
      for (int k = 0; k < g.Nz_tot; k++)
        for (int j = 0; j < g.Ny_tot; j++)
          for (int i = 0; i < g.Nx_tot; i++) (*dst[c])(i + bs, j + bs, k + bs) = blk[c][n++];
      dst[c]->Update_Ghosts();
    }
    RU.x.add(par.Rho0 * par.U0);
  } else if (par.Initial == 1) {  // grid.cpp:124-182
    Restart();
    Finish_Initialize();
    if (outputs) Write_info();
    return;
  } else {
    RU.set(0);
    RU.x.set(par.Rho0 * par.U0);
  }
  Rho.set(par.Rho0);
  P.set(0);
  if (par.IsBox) part.load_random();
  P0 = par.P0;
  T_cur = 0;
  num_timestep = 0;
  Finish_Initialize();
  if (outputs) Write_info();
}

// grid.cpp:197-229
void Grid::Finish_Initialize() {
  RU_int.copy(RU);
  RU_np1.copy(RU);
  Rho_int.copy(Rho);
  Rho_np1.copy(Rho);
  P0_int = P0;
  P0_np1 = P0;
  Rho_face.Equal_I_C2F(Rho);
  part.copy_stage(1, 0);
  part.copy_stage(3, 0);
  T.Equal_Divide(P0_int / par.R, Rho_int);
  part.gas2part_Temp(T, false);
  part.Compute_RHS_Temp(false);
  part.part2gas_Temp(RHS_Part_Temp, false);
  U.Equal_Divide(RU, Rho_face);
  C_BC = U.x.max();
}

void Grid::Update_Rho() {  // grid.cpp:327-333
  RHS_Rho.Equal_Div_F2C(RU_int);
  Rho_np1.PlusEqual_Mult(-(par.dt * RK4_post[RK4_count]), RHS_Rho);
  if (RK4_count != 3)
    Rho_new.Equal_LinComb(1, Rho, -par.dt * RK4_pre[RK4_count], RHS_Rho);
  else
    Rho_new.copy(Rho_np1);
}

void Grid::Update_P0() {  // grid.cpp:539-554
  mean_energy_transferred = RHS_Part_Temp.mean();
  if (par.IsBox) {
    if (par.cooling)
      dP0_dt = 0;
    else
      dP0_dt = -par.R / par.Cv * mean_energy_transferred;
    P0_np1 += par.dt * RK4_post[RK4_count] * dP0_dt;
    if (RK4_count != 3)
      P0_new = P0 + par.dt * RK4_pre[RK4_count] * dP0_dt;
    else
      P0_new = P0_np1;
  } else {
    P0_new = P0;
    P0_np1 = P0;
  }
}

void Grid::Update_Particle() {  // grid.cpp:659-680
  U.Equal_Divide(RU_int, Rho_face);
  part.gas2part_velocity(U);
  part.update_position(RK4_count, RK4_pre[RK4_count], RK4_post[RK4_count]);
  part.update_velocity(RK4_count, RK4_pre[RK4_count], RK4_post[RK4_count]);
  part.update_Temp(RK4_count, RK4_pre[RK4_count], RK4_post[RK4_count]);
  part.part2gas_concentration(C);
  if (par.TWC)
    part.part2gas_velocity(RHS_RU);
  else
    RHS_RU.set(0);
  part.Send_Recv();
  T.Equal_Divide(P0_new / par.R, Rho_new);
  part.gas2part_Temp(T, true);
  part.Compute_RHS_Temp(true);
  part.part2gas_Temp(RHS_Part_Temp, true);
}

void Grid::Update_RU_WOP() {  // grid.cpp:387-425
  divergence.Equal_Div_F2C(U);
  dummy2.Equal_Grad_C2F(divergence);
  dummy.Equal_Del2(U);
  RHS_RU.Equal_LinComb(-par.mp / g.Vcell(), RHS_RU, par.Mu0 / 3., dummy2, par.Mu0, dummy);
  dummy.Equal_I_C2F(U.x);
  dummy2.Equal_Ix_C2F(RU_int);
  dummy.mul(dummy2);
  dummy2.x.Equal_Div_F2C(dummy);
  RHS_RU.x.sub(dummy2.x);
  dummy.Equal_I_C2F(U.y);
  dummy2.Equal_Iy_C2F(RU_int);
  dummy.mul(dummy2);
  dummy2.y.Equal_Div_F2C(dummy);
  RHS_RU.y.sub(dummy2.y);
  dummy.Equal_I_C2F(U.z);
  dummy2.Equal_Iz_C2F(RU_int);
  dummy.mul(dummy2);
  dummy2.z.Equal_Div_F2C(dummy);
  RHS_RU.z.sub(dummy2.z);
  RHS_RU.x.PlusEqual_Mult(par.A, RU_int.x);
  RHS_RU.y.PlusEqual_Mult(par.A, RU_int.y);
  RHS_RU.z.PlusEqual_Mult(par.A, RU_int.z);
  RHS_RU.x.sub(par.A * par.Rho0 * par.U0);
  Rho_face.x.sub(par.Rho0);
  Rho_face.y.sub(par.Rho0);
  Rho_face.z.sub(par.Rho0);
  dummy.x.Equal_Mult(Rho_face.x, par.gx);
  dummy.y.Equal_Mult(Rho_face.y, par.gy);
  dummy.z.Equal_Mult(Rho_face.z, par.gz);
  RHS_RU.add(dummy);
  RU_np1.PlusEqual_Mult(par.dt * RK4_post[RK4_count], RHS_RU);
  if (RK4_count != 3)
    RU_new.Equal_LinComb(1, RU, par.dt * RK4_pre[RK4_count], RHS_RU);
  else
    RU_new.copy(RU_np1);
}

void Grid::Compute_Div_U_new() {  // grid.cpp:556-563
  dummy.x.Equal_Divide(1., Rho_new);
  dummy.y.Equal_Del2(dummy.x);
  dummy.z.copy(RHS_Part_Temp);
  if (par.cooling) dummy.z.make_mean_zero();
  divergence.Equal_LinComb(par.k / par.Cp, dummy.y,
                           -par.R / (par.Cp * P0_new * g.Vcell()), dummy.z);
}

void Grid::Compute_RHS_Pois() {  // grid.cpp:629-637
  if (par.IsBox) Rho_face.Equal_I_C2F(Rho_new);
  U.Equal_Divide(RU_new, Rho_face);
  RHS_Pois.Equal_Div_F2C(U);
  RHS_Pois.sub(divergence);
  RHS_Pois.mul(1. / (par.dt * RK4_pre[RK4_count]));
  RHS_Pois.make_mean_zero();
}

void Grid::Solve_Poisson() {  // grid.cpp:639-643
  if (par.Solver == 2) {
    last_iterations = mg.solve(Rho_face, P, RHS_Pois, par.Iteration, par.epsilon, &last_relres);
    return;
  }
  dummy.Equal_Divide(1., Rho_face);
  PS.Solve(dummy, P, RHS_Pois);
  last_iterations = PS.num_iteration;
}

void Grid::Solve_Poisson_IO() {  // grid.cpp:645-649
  if (par.Solver == 2) {
    last_iterations = mg.solve(Rho_face, P, RHS_Pois, par.Iteration, par.epsilon, &last_relres);
    return;
  }
  dummy.Equal_Divide(1., Rho_face);
  PS_IO.Solve(dummy, P, RHS_Pois);
  last_iterations = PS_IO.num_iteration1;
}

void Grid::Update_RU_WP() {  // grid.cpp:651-657
  dummy.Equal_Grad_C2F(P);
  RU_np1.PlusEqual_Mult(-(par.dt * RK4_post[RK4_count]), dummy);
  if (RK4_count != 3)
    RU_new.PlusEqual_Mult(-(par.dt * RK4_pre[RK4_count]), dummy);
  else
    RU_new.copy(RU_np1);
}

void Grid::TimeAdvance_RK4() {  // grid.cpp:705-712
  Rho_int.copy(Rho_new);
  RU_int.copy(RU_new);
  P0_int = P0_new;
  part.copy_stage(1, 2);
}

void Grid::Correct_mean_RU() {  // grid.cpp:565-577
  if (par.IsBox) {
    RU_np1.make_mean_zero();
    RU_np1.x.add(par.Rho0 * par.U0);
  } else {
    RU_np1.y.make_mean_zero();
    RU_np1.z.make_mean_zero();
  }
}

void Grid::TimeAdvance() {  // grid.cpp:315-325
  Rho.copy(Rho_np1);
  RU.copy(RU_np1);
  P0 = P0_np1;
  part.copy_stage(0, 3);
  T_cur += par.dt;
  num_timestep++;
}

// simulation.cpp:36-70, the box calls.
void Grid::box_step() {
  par.update(T_cur);
  for (RK4_count = 0; RK4_count < 4; RK4_count++) {
    Update_Rho();
    Update_P0();
    Update_Particle();
    Update_RU_WOP();
    Compute_Div_U_new();
    Compute_RHS_Pois();
    Solve_Poisson();
    Update_RU_WP();
    TimeAdvance_RK4();
  }
  Correct_mean_RU();
  TimeAdvance();
}

// ---- duct (IsBox = 0): the box grid g of the reference is the Inflow `in`

// grid.cpp:335-385 for a rank that is inlet and outlet.
void Grid::Update_Rho_duct() {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  RHS_Rho.Equal_Div_F2C(RU_int);
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++) RHS_Rho(Nx - 1, j, k) = 0;
  Rho_np1.PlusEqual_Mult(-(par.dt * RK4_post[RK4_count]), RHS_Rho);
  if (RK4_count != 3)
    Rho_new.Equal_LinComb(1, Rho, -par.dt * RK4_pre[RK4_count], RHS_Rho);
  else
    Rho_new.copy(Rho_np1);
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++) Rho_new(0, j, k) = in.rho_new[(size_t)k * Ny + j];
  if (RK4_count == 3)
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) Rho_np1(0, j, k) = in.rho_np1[(size_t)k * Ny + j];
  const double idx = 1. / g.dx;
  const int i = Nx - 1;
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++)
      Rho_np1(i, j, k) -=
          (par.dt * RK4_post[RK4_count]) * C_BC * (Rho_int(i, j, k) - Rho_int(i - 1, j, k)) * idx;
  if (RK4_count != 3)
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++)
        Rho_new(i, j, k) = Rho(i, j, k) - par.dt * RK4_pre[RK4_count] * C_BC *
                                              (Rho_int(i, j, k) - Rho_int(i - 1, j, k)) * idx;
  else
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) Rho_new(i, j, k) = Rho_np1(i, j, k);
}

// grid.cpp:427-537: no forcing or gravity; RHS zero on the outlet slice;
// inflow copies; convective outflow.
void Grid::Update_RU_WOP_duct() {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  divergence.Equal_Div_F2C(U);
  dummy2.Equal_Grad_C2F(divergence);
  dummy.Equal_Del2(U);
  RHS_RU.Equal_LinComb(-par.mp / g.Vcell(), RHS_RU, par.Mu0 / 3., dummy2, par.Mu0, dummy);
  dummy.Equal_I_C2F(U.x);
  dummy2.Equal_Ix_C2F(RU_int);
  dummy.mul(dummy2);
  dummy2.x.Equal_Div_F2C(dummy);
  RHS_RU.x.sub(dummy2.x);
  dummy.Equal_I_C2F(U.y);
  dummy2.Equal_Iy_C2F(RU_int);
  dummy.mul(dummy2);
  dummy2.y.Equal_Div_F2C(dummy);
  RHS_RU.y.sub(dummy2.y);
  dummy.Equal_I_C2F(U.z);
  dummy2.Equal_Iz_C2F(RU_int);
  dummy.mul(dummy2);
  dummy2.z.Equal_Div_F2C(dummy);
  RHS_RU.z.sub(dummy2.z);
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++) {
      RHS_RU.x(Nx - 1, j, k) = 0;
      RHS_RU.y(Nx - 1, j, k) = 0;
      RHS_RU.z(Nx - 1, j, k) = 0;
    }
  RU_np1.PlusEqual_Mult(par.dt * RK4_post[RK4_count], RHS_RU);
  if (RK4_count != 3)
    RU_new.Equal_LinComb(1, RU, par.dt * RK4_pre[RK4_count], RHS_RU);
  else
    RU_new.copy(RU_np1);
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++) {
      const size_t q = (size_t)k * Ny + j;
      RU_new.x(1, j, k) = in.ru_new[0][q];
      RU_new.y(0, j, k) = in.ru_new[1][q];
      RU_new.z(0, j, k) = in.ru_new[2][q];
    }
  if (RK4_count == 3)
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) {
        const size_t q = (size_t)k * Ny + j;
        RU_np1.x(1, j, k) = in.ru_np1[0][q];
        RU_np1.y(0, j, k) = in.ru_np1[1][q];
        RU_np1.z(0, j, k) = in.ru_np1[2][q];
      }
  const double idx = 1. / g.dx;
  const int i = Nx - 1;
  F0* n1[3] = {&RU_np1.x, &RU_np1.y, &RU_np1.z};
  F0* nw[3] = {&RU_new.x, &RU_new.y, &RU_new.z};
  const F0* ri[3] = {&RU_int.x, &RU_int.y, &RU_int.z};
  const F0* r0[3] = {&RU.x, &RU.y, &RU.z};
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++)
      for (int c = 0; c < 3; c++)
        (*n1[c])(i, j, k) -= (par.dt * RK4_post[RK4_count]) * C_BC *
                             ((*ri[c])(i, j, k) - (*ri[c])(i - 1, j, k)) * idx;
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++)
      for (int c = 0; c < 3; c++)
        (*nw[c])(i, j, k) = RK4_count != 3
                                ? (*r0[c])(i, j, k) - par.dt * RK4_pre[RK4_count] * C_BC *
                                                          ((*ri[c])(i, j, k) - (*ri[c])(i - 1, j, k)) * idx
                                : (*n1[c])(i, j, k);
}

// grid.cpp:682-703
void Grid::Update_Particle_duct() {
  U.Equal_Divide(RU_int, Rho_face);
  part.gas2part_velocity(U);
  part.update_position(RK4_count, RK4_pre[RK4_count], RK4_post[RK4_count]);
  part.update_velocity(RK4_count, RK4_pre[RK4_count], RK4_post[RK4_count]);
  part.update_Temp(RK4_count, RK4_pre[RK4_count], RK4_post[RK4_count]);
  part.part2gas_concentration(dummy.x);
  C.copy(dummy.x);
  if (par.TWC)
    part.part2gas_velocity(RHS_RU);
  else
    RHS_RU.set(0);
  part.Send_Recv_duct(in.pbuf);
  T.Equal_Divide(P0_new / par.R, Rho_new);
  part.gas2part_Temp(T, true);
  part.Compute_RHS_Temp(true);
  part.part2gas_Temp(RHS_Part_Temp, true);
}

// grid.cpp:579-627 for a rank that is inlet and outlet (the Reduce/Bcast
// pairs are identities).
void Grid::Correct_OutFlow() {
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  double sum_div = divergence.sum() * g.dx;
  Rho_face.Equal_I_C2F(Rho_new);
  U.Equal_Divide(RU_new, Rho_face);
  double sum_out = 0;
  for (int j = 1; j < Ny - 1; j++)
    for (int k = 1; k < Nz - 1; k++) sum_out += U.x(Nx - 1, j, k);
  double sum_in = 0;
  for (int j = 1; j < Ny - 1; j++)
    for (int k = 1; k < Nz - 1; k++) sum_in += U.x(1, j, k);
  double ratio = (sum_div + sum_in) / sum_out;
  U_avg_in = sum_in / (g.Ny_tot * g.Nz_tot);
  U_avg_out = sum_out / (g.Ny_tot * g.Nz_tot);
  C_BC = U_avg_out;
  for (int j = 0; j < Ny; j++)
    for (int k = 0; k < Nz; k++) RU_new.x(Nx - 1, j, k) *= ratio;
  if (RK4_count == 3)
    for (int j = 0; j < Ny; j++)
      for (int k = 0; k < Nz; k++) RU_np1.x(Nx - 1, j, k) *= ratio;
}

// simulation.cpp:36-70, the GRID_IO calls, with the box replaced by `in`.
void Grid::duct_step() {
  par.update(T_cur);
  for (RK4_count = 0; RK4_count < 4; RK4_count++) {
    Update_Rho_duct();
    Update_P0();
    Update_Particle_duct();
    Update_RU_WOP_duct();
    Compute_Div_U_new();
    Correct_OutFlow();
    Compute_RHS_Pois();
    Solve_Poisson_IO();
    Update_RU_WP();
    TimeAdvance_RK4();
  }
  Correct_mean_RU();
  TimeAdvance();
}

// ---- the coupled box and duct (simulation.cpp:19-73) -----------------------

// grid.cpp:352-363 and 485-503 (bs = 1): the duct's inlet reads the box's
// first interior slice: Rho_new and Rho_np1 at i = 1, RU_new/RU_np1 x at the
// face i = 2 and y, z at i = 1, over the whole ghosted (j, k) plane.  At
// initialisation the restart reads the box's RU.y, RU.z at i = 1
// (grid.cpp:155-164), which is what the ru_np1 y, z planes carry then.
void pair_inflow(Grid& b, Grid& d, bool init) {
  const int Ny = b.g.Ny, Nz = b.g.Nz;
  const size_t m = (size_t)Ny * Nz;
  std::vector<double>* dst[8] = {&d.in.rho_new,   &d.in.rho_np1,   &d.in.ru_new[0],
                                 &d.in.ru_new[1], &d.in.ru_new[2], &d.in.ru_np1[0],
                                 &d.in.ru_np1[1], &d.in.ru_np1[2]};
  const F0* src[8] = {&b.Rho_new,  &b.Rho_np1,  &b.RU_new.x, &b.RU_new.y,
                      &b.RU_new.z, &b.RU_np1.x, &b.RU_np1.y, &b.RU_np1.z};
  if (init) {
    const F0* s0[8] = {&b.Rho, &b.Rho, &b.RU.x, &b.RU.y, &b.RU.z, &b.RU.x, &b.RU.y, &b.RU.z};
    for (int q = 0; q < 8; q++) src[q] = s0[q];
  }
  for (int q = 0; q < 8; q++) {
    const int i = (q == 2 || q == 5) ? 2 : 1;
    dst[q]->assign(m, 0.0);
    for (int k = 0; k < Nz; k++)
      for (int j = 0; j < Ny; j++) (*dst[q])[(size_t)k * Ny + j] = (*src[q])(i, j, k);
  }
}

// simulation.cpp:27-30: GRID.Initialize(GRID), then GRID_IO.Initialize(GRID).
// Each grid reads its own initial fields (the reference reads U.bin, V.bin,
// W.bin from one directory for both).
void pair_initialize(Grid& b, Grid& d, const double* const ic_b[3],
                     const double* const ic_d[3]) {
  b.Initialize(ic_b[0], ic_b[1], ic_b[2]);
  pair_inflow(b, d, true);
  d.Initialize(ic_d[0], ic_d[1], ic_d[2]);
}

// simulation.cpp:34-62: per substep the box's eight methods, then the duct's
// with the box's planes and '>hi' list, then the box's TimeAdvance_RK4 and
// the duct's last two; the mean corrections and time advances at the end.
void pair_step(Grid& b, Grid& d) {
  b.par.update(b.T_cur);
  d.par.update(d.T_cur);
  for (int k = 0; k < 4; k++) {
    b.RK4_count = k;
    b.Update_Rho();
    b.Update_P0();
    b.Update_Particle();
    b.Update_RU_WOP();
    b.Compute_Div_U_new();
    b.Compute_RHS_Pois();
    b.Solve_Poisson();
    b.Update_RU_WP();
    d.RK4_count = k;
    pair_inflow(b, d, false);
    // particle.cpp:821 and 836-838: the duct's Rbuf1 is the box's Rbuf1,
    // which on one rank is the box's own '>hi' list of its x pass
    d.in.pbuf = b.part.Sbuf0;
    d.Update_Rho_duct();
    d.Update_P0();
    d.Update_Particle_duct();
    d.Update_RU_WOP_duct();
    d.Compute_Div_U_new();
    d.Correct_OutFlow();
    d.Compute_RHS_Pois();
    d.Solve_Poisson_IO();
    b.TimeAdvance_RK4();
    d.Update_RU_WP();
    d.TimeAdvance_RK4();
  }
  b.Correct_mean_RU();
  d.Correct_mean_RU();
  b.TimeAdvance();
  d.TimeAdvance();
}

// grid.cpp:77-98: one read of touch.check stops both grids and stores both.
bool pair_touch(Grid& b, Grid& d) {
  const bool ans = b.Touch();
  d.Is_touch = ans;
  if (ans) d.Store();
  return ans;
}

// grid.cpp:714-811 subset: out = {TKE, Rho_Mean, Rho_Max, Rho_Min, C_Max,
// C_Min, C_Mean, mean dissipation (x-average of the DISSIPATION profile),
// mean particle temperature (extension: sum of T in vector order / Np),
// Poisson iterations of the last solve}.  Uses private scratch so the step
// state is untouched.
void Grid::diagnostics(double* out) {
  const int bs = g.bs;
  F0 rho;
  rho.init(&g);
  rho.copy(Rho);
  out[1] = rho.mean();
  out[2] = rho.max();
  out[3] = rho.min();
  out[4] = C.max();
  out[5] = C.min();
  out[6] = C.mean();
  // TKE: grid.cpp:731-775
  rho.reduceX();
  F1 u2, t3;
  u2.init(&g);
  t3.init(&g);
  u2.Equal_Divide(RU, Rho);
  u2.reduceX();
  t3.Equal_TKE(u2, rho);
  out[0] = t3.x.mean() + t3.y.mean() + t3.z.mean();
  // dissipation: grid.cpp:723, 794-811
  F1 uf;
  uf.init(&g);
  uf.Equal_Divide(RU_np1, Rho_face);
  F0 sij, nu, dis;
  sij.init(&g);
  nu.init(&g);
  dis.init(&g);
  sij.copy(u2.x);  // ghosts of the reference's dummy2.x hold RU.x/Rho
  sij.Equal_SijSij(uf);
  nu.Equal_nu(par.Mu0, Rho);
  dis.Equal_dissipation(nu, sij);
  double s = 0;
  for (int i = bs; i < g.Nx - bs; i++) s += dis.red[i];
  out[7] = s / g.Nx_tot;
  double ts = 0;
  for (int i = 0; i < part.Np; i++) ts += part.T[i];
  out[8] = part.Np > 0 ? ts / part.Np : 0.0;
  out[9] = last_iterations;
}

F0* Grid::field0(const std::string& n) {
  if (n == "P") return &P;
  if (n == "RHS_Pois") return &RHS_Pois;
  if (n == "C") return &C;
  if (n == "Rho") return &Rho;
  if (n == "Rho_int") return &Rho_int;
  if (n == "Rho_new") return &Rho_new;
  if (n == "Rho_np1") return &Rho_np1;
  if (n == "T") return &T;
  if (n == "divergence") return &divergence;
  if (n == "RHS_Part_Temp") return &RHS_Part_Temp;
  return nullptr;
}

F1* Grid::field1(const std::string& n) {
  if (n == "RU") return &RU;
  if (n == "RU_int") return &RU_int;
  if (n == "RU_new") return &RU_new;
  if (n == "RU_np1") return &RU_np1;
  if (n == "RHS_RU") return &RHS_RU;
  if (n == "U") return &U;
  if (n == "Rho_face") return &Rho_face;
  return nullptr;
}

}  // namespace orc
