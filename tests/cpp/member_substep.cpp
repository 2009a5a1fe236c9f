// Drives the time step of the box through the reference's member surface of
// tensor0, tensor1, particle and poisson (the calls of grid.cpp's box
// methods, 315-333, 387-425, 539-577, 629-680, 705-712, written out here as
// a caller of those classes would) on the device classes of libfr_amd.so,
// and the duct's pressure solve through poisson_IO.  Writes the resulting
// state for the parity test to compare with the oracle.  TEST PROGRAM.
//
//   member_substep box  <params.in> <ic_dir> <out_dir> <steps>
//   member_substep duct <params.in> <ic_dir> <out_dir>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../first_repo_amd/host/fr_host.h"

static std::string out_dir;

static int put(const std::string& name, const void* p, size_t bytes) {
  FILE* f = fopen((out_dir + "/" + name).c_str(), "wb");
  if (!f) return 1;
  const size_t w = fwrite(p, 1, bytes, f);
  fclose(f);
  return w == bytes ? 0 : 1;
}

static int put_field(const std::string& name, tensor0& t) {
  std::vector<double> h((size_t)t.size());
  t.download(h.data());
  return put(name + ".bin", h.data(), sizeof(double) * h.size());
}

#define CHECK(x)                                                      \
  do {                                                                \
    int rc_ = (x);                                                    \
    if (rc_) {                                                        \
      fprintf(stderr, "member_substep: %s -> %d (line %d)\n", #x, rc_, __LINE__); \
      return 2;                                                       \
    }                                                                 \
  } while (0)

static int box(const char* params_path, const char* ic, int steps) {
  grid g;
  CHECK(g.setup(params_path, ic));
  CHECK(g.Initialize());
  params& par = g.param;
  gridsize& size = g.size;
  tensor1 dummy2;
  dummy2.init(&size, g.st);
  tensor0 RHS_Rho(&size, g.st), T(&size, g.st);
  tensor1& RU = g.RU;
  tensor1& RU_int = g.RU_int;
  tensor1& RU_new = g.RU_new;
  tensor1& RU_np1 = g.RU_np1;
  tensor1& RHS_RU = g.RHS_RU;
  tensor1& U = g.U;
  tensor1& Rho_face = g.Rho_face;
  tensor1& dummy = g.dummy;
  tensor0& Rho = g.Rho;
  tensor0& Rho_int = g.Rho_int;
  tensor0& Rho_new = g.Rho_new;
  tensor0& Rho_np1 = g.Rho_np1;
  tensor0& P = g.P;
  tensor0& RHS_Pois = g.RHS_Pois;
  tensor0& C = g.C;
  tensor0& RHS_Part_Temp = g.RHS_Part_Temp;
  tensor0& divergence = g.divergence;
  particle& part = g.part;
  const double* pre = g.RK4_preCoeff;
  const double* post = g.RK4_postCoeff;
  double P0 = g.P0, P0_int = g.P0_int, P0_new = g.P0_new, P0_np1 = g.P0_np1, T_cur = g.T_cur;
  for (int step = 0; step < steps; step++) {
    par.update(T_cur);
    const double dt = par.dt();
    for (int k = 0; k < 4; k++) {
      // Update_Rho
      RHS_Rho.Equal_Div_F2C(RU_int);
      Rho_np1.PlusEqual_Mult(-(dt * post[k]), RHS_Rho);
      if (k != 3)
        Rho_new.Equal_LinComb(1, Rho, -dt * pre[k], RHS_Rho);
      else
        Rho_new = Rho_np1;
      // Update_P0
      const double q = RHS_Part_Temp.mean();
      const double dP0_dt = par.Is_Cooling() ? 0 : -par.R() / par.Cv() * q;
      P0_np1 += dt * post[k] * dP0_dt;
      P0_new = (k != 3) ? P0 + dt * pre[k] * dP0_dt : P0_np1;
      // Update_Particle
      U.Equal_Divide(RU_int, Rho_face);
      CHECK(part.gas2part_velocity(U));
      CHECK(part.update_position(k, pre[k], post[k]));
      CHECK(part.update_velocity(k, pre[k], post[k]));
      CHECK(part.update_Temp(k, pre[k], post[k]));
      CHECK(part.part2gas_concentration(C));
      if (par.TWC())
        CHECK(part.part2gas_velocity(RHS_RU));
      else
        RHS_RU = 0;
      CHECK(part.Send_Recv());
      T.Equal_Divide(P0_new / par.R(), Rho_new);
      CHECK(part.gas2part_Temp_new(T));
      CHECK(part.Compute_RHS_Temp_new());
      CHECK(part.part2gas_Temp_new(RHS_Part_Temp));
      // Update_RU_WOP
      divergence.Equal_Div_F2C(U);
      dummy2.Equal_Grad_C2F(divergence);
      dummy.Equal_Del2(U);
      RHS_RU.Equal_LinComb(-par.mp() / size.Vcell(), RHS_RU, par.Mu0() / 3., dummy2, par.Mu0(),
                           dummy);
      tensor0* rc[3] = {&RHS_RU.x, &RHS_RU.y, &RHS_RU.z};
      tensor0* uc[3] = {&U.x, &U.y, &U.z};
      tensor0* d2c[3] = {&dummy2.x, &dummy2.y, &dummy2.z};
      for (int a = 0; a < 3; a++) {
        dummy.Equal_I_C2F(*uc[a]);
        if (a == 0) dummy2.Equal_Ix_C2F(RU_int);
        if (a == 1) dummy2.Equal_Iy_C2F(RU_int);
        if (a == 2) dummy2.Equal_Iz_C2F(RU_int);
        dummy *= dummy2;
        d2c[a]->Equal_Div_F2C(dummy);
        *rc[a] -= *d2c[a];
      }
      RHS_RU.PlusEqual_Mult(par.A(), RU_int);
      RHS_RU.x -= par.A() * par.Rho0() * par.U0();
      Rho_face -= par.Rho0();
      dummy.x.Equal_Mult(Rho_face.x, par.gx());
      dummy.y.Equal_Mult(Rho_face.y, par.gy());
      dummy.z.Equal_Mult(Rho_face.z, par.gz());
      RHS_RU += dummy;
      RU_np1.PlusEqual_Mult(dt * post[k], RHS_RU);
      if (k != 3)
        RU_new.Equal_LinComb(1, RU, dt * pre[k], RHS_RU);
      else
        RU_new = RU_np1;
      // Compute_Div_U_new
      dummy.x.Equal_Divide(1., Rho_new);
      dummy.y.Equal_Del2(dummy.x);
      dummy.z = RHS_Part_Temp;
      if (par.Is_Cooling()) dummy.z.make_mean_zero();
      divergence.Equal_LinComb(par.k() / par.Cp(), dummy.y,
                               -par.R() / (par.Cp() * P0_new * size.Vcell()), dummy.z);
      // Compute_RHS_Pois
      Rho_face.Equal_I_C2F(Rho_new);
      U.Equal_Divide(RU_new, Rho_face);
      RHS_Pois.Equal_Div_F2C(U);
      RHS_Pois -= divergence;
      RHS_Pois *= (1. / (dt * pre[k]));
      RHS_Pois.make_mean_zero();
      // Solve_Poisson
      dummy.Equal_Divide(1., Rho_face);
      CHECK(g.PS_.Solve(dummy, P, RHS_Pois));
      // Update_RU_WP
      dummy.Equal_Grad_C2F(P);
      RU_np1.PlusEqual_Mult(-(dt * post[k]), dummy);
      if (k != 3)
        RU_new.PlusEqual_Mult(-(dt * pre[k]), dummy);
      else
        RU_new = RU_np1;
      // TimeAdvance_RK4
      Rho_int = Rho_new;
      RU_int = RU_new;
      P0_int = P0_new;
      CHECK(part.dev.copy_stage(1, 2, g.st));
    }
    // Correct_mean_RU, TimeAdvance
    RU_np1.make_mean_zero();
    RU_np1.x += par.Rho0() * par.U0();
    Rho = Rho_np1;
    RU = RU_np1;
    P0 = P0_np1;
    CHECK(part.dev.copy_stage(0, 3, g.st));
    T_cur += dt;
  }
  CHECK(fr::host_status());
  CHECK((int)hipStreamSynchronize(g.st));
  const char* f1n[4] = {"RU", "RU_int", "RU_np1", "Rho_face"};
  tensor1* f1[4] = {&RU, &RU_int, &RU_np1, &Rho_face};
  for (int i = 0; i < 4; i++)
    for (int c = 0; c < 3; c++) CHECK(put_field(std::string(f1n[i]) + "." + std::to_string(c), f1[i]->c(c)));
  const char* f0n[7] = {"P", "RHS_Pois", "C", "Rho", "Rho_int", "Rho_np1", "RHS_Part_Temp"};
  tensor0* f0[7] = {&P, &RHS_Pois, &C, &Rho, &Rho_int, &Rho_np1, &RHS_Part_Temp};
  for (int i = 0; i < 7; i++) CHECK(put_field(f0n[i], *f0[i]));
  const int np = part.dev.np;
  std::vector<double> h((size_t)np);
  for (int q = 0; q < 7; q++)
    for (int s : {0, 1, 3}) {
      CHECK(part.dev.get(h.data(), part.dev.s[q][s], g.st));
      CHECK(put("part" + std::to_string(q) + "." + std::to_string(s) + ".bin", h.data(),
                sizeof(double) * np));
    }
  std::vector<int> fl((size_t)np);
  CHECK(part.dev.get_flags(fl.data(), g.st));
  CHECK(put("flag.bin", fl.data(), sizeof(int) * np));
  const double sc[5] = {P0, P0_int, P0_new, P0_np1, T_cur};
  CHECK(put("scalars.bin", sc, sizeof(sc)));
  return 0;
}

// poisson_IO::Solve on the duct: C = 1/Rho_face from the density and P and
// RHS_Pois as the grid holds them (the test sets them first from files).
static int duct(const char* params_path, const char* ic) {
  grid g;
  CHECK(g.setup(params_path, ic));
  CHECK(g.Initialize());
  gridsize& size = g.size;
  std::vector<double> h((size_t)size.size());
  auto load = [&](const char* f, tensor0& t) {
    FILE* fp = fopen((out_dir + "/" + f).c_str(), "rb");
    if (!fp) return 1;
    const size_t r = fread(h.data(), sizeof(double), h.size(), fp);
    fclose(fp);
    if (r != h.size()) return 1;
    t.upload(h.data());
    return 0;
  };
  tensor0 rho(&size, g.st);
  CHECK(load("in_rho.bin", rho));
  CHECK(load("in_P.bin", g.P));
  CHECK(load("in_RHS.bin", g.RHS_Pois));
  tensor1 rf, Cf;
  rf.init(&size, g.st);
  Cf.init(&size, g.st);
  rf = 1.;  // the inlet x-face ghost keeps 1 (the solve does not read it)
  rf.Equal_I_C2F(rho);
  Cf.Equal_Divide(1., rf);
  poisson_IO PS;
  CHECK(PS.init(&g.param, &size, g.st));
  CHECK(PS.Solve(Cf, g.P, g.RHS_Pois));
  CHECK(fr::host_status());
  CHECK(put_field("P", g.P));
  const double info[4] = {(double)PS.num_iteration1(), PS.num_iteration2(), PS.FFT_TIME,
                          PS.SOLVER_TIME};
  CHECK(put("info.bin", info, sizeof(info)));
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: member_substep box|duct params ic_dir out_dir [steps]\n");
    return 1;
  }
  out_dir = argv[4];
  if (!strcmp(argv[1], "box")) return box(argv[2], argv[3], argc > 5 ? atoi(argv[5]) : 1);
  if (!strcmp(argv[1], "duct")) return duct(argv[2], argv[3]);
  return 1;
}
