// Oracle run files: statistics, snapshots, restart, info.txt and the
// touch.check stop, for one rank.  TEST INFRASTRUCTURE.  Restates
// grid.cpp:25-98 (stat files, touch.check), 124-182 (restart), 233-313
// (Store), 714-866 (Statistics), 869-973 (Write_info);
// communicator.cpp:77-172 (MPI-IO layouts on one rank: the global interior,
// [Nz_tot][Ny_tot][Nx_tot] per component, x fastest; the outlet slice
// [Nz_tot][Ny_tot] from local i = Nx-1); particle.cpp:867-1025 (Store_All,
// Load_All).  Paths are relative to the working directory, as in the
// reference.
#include <fcntl.h>
#include <unistd.h>

#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "orc.h"

namespace orc {

// ---- raw blocks ------------------------------------------------------------
static bool write_all(const std::string& path, const void* p, size_t n, bool excl,
                      bool append = false) {
  int flags = O_WRONLY | O_CREAT | (excl ? O_EXCL : 0) | (append ? O_APPEND : 0);
  int fd = open(path.c_str(), flags, 0644);
  if (fd < 0) return false;
  const char* c = (const char*)p;
  size_t done = 0;
  while (done < n) {
    ssize_t w = write(fd, c + done, n - done);
    if (w <= 0) {
      close(fd);
      return false;
    }
    done += (size_t)w;
  }
  close(fd);
  return true;
}

static bool read_all(const std::string& path, void* p, size_t n, size_t offset = 0) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  if (offset) fseek(f, (long)offset, SEEK_SET);
  size_t got = fread(p, 1, n, f);
  fclose(f);
  return got == n;
}

// interior block of a ghosted field, x fastest (communicator.cpp:29-39)
static void interior(const F0& f, std::vector<double>& out) {
  const Geom& g = *f.g;
  out.clear();
  for (int k = 0; k < g.Nz_tot; k++)
    for (int j = 0; j < g.Ny_tot; j++)
      for (int i = 0; i < g.Nx_tot; i++) out.push_back(f(i + g.bs, j + g.bs, k + g.bs));
}
static void set_interior(F0& f, const double* blk) {
  const Geom& g = *f.g;
  size_t n = 0;
  for (int k = 0; k < g.Nz_tot; k++)
    for (int j = 0; j < g.Ny_tot; j++)
      for (int i = 0; i < g.Nx_tot; i++) f(i + g.bs, j + g.bs, k + g.bs) = blk[n++];
}
// outlet slice, local i = Nx-1 (communicator.cpp:41-51)
static void slice(const F0& f, std::vector<double>& out) {
  const Geom& g = *f.g;
  out.clear();
  for (int k = 0; k < g.Nz_tot; k++)
    for (int j = 0; j < g.Ny_tot; j++) out.push_back(f(g.Nx - 1, j + g.bs, k + g.bs));
}
static void set_slice(F0& f, const double* blk) {
  const Geom& g = *f.g;
  size_t n = 0;
  for (int k = 0; k < g.Nz_tot; k++)
    for (int j = 0; j < g.Ny_tot; j++) f(g.Nx - 1, j + g.bs, k + g.bs) = blk[n++];
}

bool write_field(const std::string& path, const F0& f) {  // communicator.cpp:129-135
  std::vector<double> b;
  interior(f, b);
  return write_all(path, b.data(), b.size() * 8, false);
}
bool write_field(const std::string& path, const F1& f) {  // communicator.cpp:137-147
  std::vector<double> b, all;
  for (const F0* c : {&f.x, &f.y, &f.z}) {
    interior(*c, b);
    all.insert(all.end(), b.begin(), b.end());
  }
  return write_all(path, all.data(), all.size() * 8, false);
}
static bool write_slice(const std::string& path, const F0* const* c, int n) {
  std::vector<double> b, all;
  for (int m = 0; m < n; m++) {
    slice(*c[m], b);
    all.insert(all.end(), b.begin(), b.end());
  }
  return write_all(path, all.data(), all.size() * 8, false);
}
static bool read_field(const std::string& path, F0& f, int comp = 0) {
  const Geom& g = *f.g;
  std::vector<double> b((size_t)g.size_tot());
  if (!read_all(path, b.data(), b.size() * 8, (size_t)comp * b.size() * 8)) return false;
  set_interior(f, b.data());
  return true;
}
static bool read_slice(const std::string& path, F0& f, int comp = 0) {
  const Geom& g = *f.g;
  std::vector<double> b((size_t)g.Ny_tot * g.Nz_tot);
  if (!read_all(path, b.data(), b.size() * 8, (size_t)comp * b.size() * 8)) return false;
  set_slice(f, b.data());
  return true;
}

// ---- constructor part: touch.check and the stat files (grid.cpp:25-49) -----
static const char* kStatNames[15] = {"TKE",     "C_Max",         "C_Min",    "C_Mean",
                                     "Rho_Max", "Rho_Min",       "Rho_Mean", "Max_CFL_Vp",
                                     "Max_CFL_U", "Num_Iteration", "Balance_Index", "Uin",
                                     "Uout",    "dx_o_eta",      "ST"};

void Grid::open_outputs() {
  {
    std::ofstream t("touch.check", std::ios::out | std::ios::trunc);
    t << 0;
  }
  if (par.Statistics)
    for (int q = 0; q < 15; q++) {
      stat_files[q] = par.Stat_dir + kStatNames[q] + ".dat";
      std::ofstream f(stat_files[q], std::ios::out | std::ios::app);
    }
}

// ---- grid::Statistics (grid.cpp:714-866) --------------------------------------
static void store_stat_binary(const Params& par, const char* name, const F0& T) {
  // communicator.cpp:164-172: interior x-profile appended
  const Geom& g = *T.g;
  write_all(par.Stat_dir + name + ".bin", T.red.data() + g.bs, sizeof(double) * g.Nx_tot, false,
            true);
}

static void stat_line(const std::string& path, double t, double v) {
  std::ofstream f(path, std::ios::out | std::ios::app);
  f << t << " " << v << std::endl;
}
static void stat_line(const std::string& path, double t, int v) {
  std::ofstream f(path, std::ios::out | std::ios::app);
  f << t << " " << v << std::endl;
}

void Grid::Statistics() {
  if (!par.Statistics) return;
  double Rho_mean = Rho.mean();
  double Rho_max = Rho.max();
  double Rho_min = Rho.min();
  double C_max = C.max();
  double C_min = C.min();
  double C_mean = C.mean();
  U.Equal_Divide(RU_np1, Rho_face);
  double Gas_CFL_Max = U.max_cfl(par.dt);
  const double lnp = part.Np;  // particle::Balance_Index on one rank: Np / Np
  double Load_Balance = lnp / lnp;
  Rho.reduceX();
  store_stat_binary(par, "RHO", Rho);
  dummy.x.Equal_Reduced_sd(Rho);
  store_stat_binary(par, "RHO_SD", dummy.x);
  dummy.Equal_I_F2C(RU);
  dummy2.Equal_Divide(RU, Rho);
  dummy.reduceX();
  store_stat_binary(par, "RU_x", dummy.x);
  store_stat_binary(par, "RU_y", dummy.x);  // the reference stores x for y and z
  store_stat_binary(par, "RU_z", dummy.x);
  dummy2.reduceX();
  store_stat_binary(par, "U_x", dummy2.x);
  store_stat_binary(par, "U_y", dummy2.y);
  store_stat_binary(par, "U_z", dummy2.z);
  Rho_face.reduceX();
  dummy3.Equal_TKE2(dummy, dummy2, Rho);
  store_stat_binary(par, "TKE2x", dummy3.x);
  store_stat_binary(par, "TKE2y", dummy3.y);
  store_stat_binary(par, "TKE2z", dummy3.z);
  dummy3.Equal_UiUi2(dummy, dummy2, Rho);
  store_stat_binary(par, "UiUi2x", dummy3.x);
  store_stat_binary(par, "UiUi2y", dummy3.y);
  store_stat_binary(par, "UiUi2z", dummy3.z);
  dummy3.Equal_Cor_Rho_U(dummy2, Rho);
  store_stat_binary(par, "Cor_R_Ux", dummy3.x);
  store_stat_binary(par, "Cor_R_Uy", dummy3.y);
  store_stat_binary(par, "Cor_R_Uz", dummy3.z);
  dummy3.Equal_TKE(dummy2, Rho);
  store_stat_binary(par, "TKEx", dummy3.x);
  store_stat_binary(par, "TKEy", dummy3.y);
  store_stat_binary(par, "TKEz", dummy3.z);
  double TKE = dummy3.x.mean() + dummy3.y.mean() + dummy3.z.mean();
  dummy3.Equal_UiUi(dummy2, Rho);
  store_stat_binary(par, "UiUix", dummy3.x);
  store_stat_binary(par, "UiUiy", dummy3.y);
  store_stat_binary(par, "UiUiz", dummy3.z);
  C.reduceX();
  store_stat_binary(par, "C_avg", C);
  P.reduceX();
  store_stat_binary(par, "P_avg", P);
  dummy2.x.Equal_SijSij(U);
  dummy2.y.Equal_WijWij(U);
  dummy2.x.reduceX();
  store_stat_binary(par, "SijSij", dummy2.x);
  dummy2.y.reduceX();
  store_stat_binary(par, "WijWij", dummy2.y);
  dummy2.z.Equal_ST(par.Tp, dummy2.x);
  store_stat_binary(par, "ST", dummy2.z);
  double avg_ST = dummy2.z.mean();
  dummy2.z.Equal_nu(par.Mu0, Rho);
  store_stat_binary(par, "NU", dummy2.z);
  dummy2.y.Equal_dissipation(dummy2.z, dummy2.x);
  store_stat_binary(par, "DISSIPATION", dummy2.y);
  dummy2.x.Equal_eta(dummy2.z, dummy2.y);
  store_stat_binary(par, "ETA", dummy2.x);
  double dx_over_eta_min = g.dx / dummy2.x.min();
  RHS_Part_Temp.reduceX();
  store_stat_binary(par, "HT", RHS_Part_Temp);
  const double v[15] = {TKE,     C_max,       C_min,        C_mean, Rho_max,
                        Rho_min, Rho_mean,    0,            Gas_CFL_Max, 0,
                        Load_Balance, U_avg_in, U_avg_out, dx_over_eta_min, avg_ST};
  for (int q = 0; q < 15; q++) {
    if (q == 7) continue;  // Max_CFL_Vp.dat is opened but never written
    if (q == 9)
      stat_line(stat_files[q], T_cur, last_iterations);
    else
      stat_line(stat_files[q], T_cur, v[q]);
  }
  write_all(par.Stat_dir + "TIME.bin", &T_cur, sizeof(double), false, true);
}

// ---- grid::Store (grid.cpp:233-313) and particle::Store_All (867-917) -----
static std::string snap(const Params& par, const char* name, int n, const char* ext) {
  std::ostringstream s;
  s << par.Data_dir << name << "_" << n << ext;
  return s.str();
}

bool Grid::Store() {
  bool ok = true;
  if (num_timestep == 1) {
    F1 xyz;
    xyz.init(&g);
    for (int k = 0; k < g.Nz_tot; k++)
      for (int j = 0; j < g.Ny_tot; j++)
        for (int i = 0; i < g.Nx_tot; i++) {
          xyz.x(i + g.bs, j + g.bs, k + g.bs) = g.dx / 2. + i * g.dx;
          xyz.y(i + g.bs, j + g.bs, k + g.bs) = g.dy / 2. + j * g.dy;
          xyz.z(i + g.bs, j + g.bs, k + g.bs) = g.dz / 2. + k * g.dz;
        }
    ok &= write_field(par.Data_dir + "XYZ.bin", xyz);
  }
  if ((num_timestep % par.Data_freq_fast == 0) || Is_touch) {
    ok &= write_field(snap(par, "RU", num_timestep, ".bin"), RU);
    if (!par.IsBox) {
      const F0* c[3] = {&RU.x, &RU.y, &RU.z};
      ok &= write_slice(snap(par, "RU_S", num_timestep, ".bin"), c, 3);
    }
    std::ofstream numbers(snap(par, "numbers", num_timestep, ".dat"));
    numbers << P0 << std::endl;
    numbers << T_cur << std::endl;
    numbers << num_timestep << std::endl;
    numbers << part.Np << std::endl;
  }
  if ((num_timestep % par.Data_freq_slow == 0) || Is_touch) {
    const char* names[7] = {"Part_x", "Part_y", "Part_z", "Part_u", "Part_v", "Part_w",
                            "Part_T"};
    for (int q = 0; q < 7; q++) {
      const std::vector<double>& v = *part.stage(q, 0);
      ok &= write_all(snap(par, names[q], num_timestep, ".bin"), v.data(),
                      sizeof(double) * part.Np, true);
    }
    ok &= write_all(snap(par, "Part_flag", num_timestep, ".bin"), part.flag.data(),
                    sizeof(int) * part.Np, true);
    ok &= write_field(snap(par, "C", num_timestep, ".bin"), C);
    ok &= write_field(snap(par, "Rho", num_timestep, ".bin"), Rho);
    if (!par.IsBox) {
      const F0* c[1] = {&Rho};
      ok &= write_slice(snap(par, "Rho_S", num_timestep, ".bin"), c, 1);
    }
  }
  return ok;
}

// ---- grid::Touch (grid.cpp:77-98) ------------------------------------------------
bool Grid::Touch() {
  bool ans = false;
  std::ifstream t("touch.check");
  t >> ans;
  Is_touch = ans;
  if (ans) Store();
  return ans;
}

// ---- restart (grid.cpp:124-182; particle.cpp:961-1025) --------------------
bool Grid::Restart() {
  const bool box = par.IsBox != 0;
  const std::string sfx = box ? "" : "_IO";
  for (int m = 0; m < 3; m++) {
    F0* c[3] = {&RU.x, &RU.y, &RU.z};
    if (!read_field("Restart_RU" + sfx + ".bin", *c[m], m)) return false;
  }
  if (!read_field("Restart_Rho" + sfx + ".bin", Rho)) return false;
  int np = par.N0;
  if (box) {
    RU.Update_Ghosts();
    Rho.Update_Ghosts();
  } else {
    F0* c[3] = {&RU.x, &RU.y, &RU.z};
    for (int m = 0; m < 3; m++)
      if (!read_slice("Restart_RU_IO_S.bin", *c[m], m)) return false;
    if (!read_slice("Restart_Rho_IO_S.bin", Rho)) return false;
    // inlet ghosts: Rho0, and the box's first-interior v, w (grid.cpp:150-160),
    // taken here from the inflow planes of the box's RU_np1
    for (int j = 0; j < g.Ny; j++)
      for (int k = 0; k < g.Nz; k++) {
        Rho(0, j, k) = par.Rho0;
        if (!in.ru_np1[1].empty()) {
          RU.y(0, j, k) = in.ru_np1[1][(size_t)k * g.Ny + j];
          RU.z(0, j, k) = in.ru_np1[2][(size_t)k * g.Ny + j];
        }
      }
    RU.Update_Ghosts();
    Rho.Update_Ghosts();
  }
  P.set(0);
  {
    std::ifstream numbers(box ? "Restart_numbers.dat" : "Restart_numbers_IO.dat");
    numbers >> P0;
    numbers >> T_cur;
    numbers >> num_timestep;
    if (!box) {
      numbers >> np;
      par.N0 = np;  // params::Set_Np
    }
    if (!numbers) return false;
  }
  // particle::Load_All
  part.Np = np;
  part.RESIZE();
  const char* names[7] = {"Part_x", "Part_y", "Part_z", "Part_u", "Part_v", "Part_w", "Part_T"};
  for (int q = 0; q < 7; q++)
    if (!read_all("Restart_" + std::string(names[q]) + sfx + ".bin", part.stage(q, 0)->data(),
                  sizeof(double) * np))
      return false;
  if (!read_all("Restart_Part_flag" + sfx + ".bin", part.flag.data(), sizeof(int) * np))
    return false;
  part.x_new = part.x;
  part.y_new = part.y;
  part.z_new = part.z;
  part.Send_Recv();
  return true;
}

// ---- grid::Write_info (grid.cpp:869-973), one rank -----------------------------
void Grid::Write_info() {
  std::ofstream info("info.txt");
  const Params& p = par;
  info << ":::Global grid is " << g.Nx_tot << "x" << g.Ny_tot << "x" << g.Nz_tot
       << " with bordersize (number of ghost cells around local grid) =" << g.bs << std::endl;
  info << std::endl;
  info << ":::Code started at t=" << T_cur << " with dt=" << p.dt
       << " ,and is supposed to finish at t=" << p.T_final << std::endl;
  info << std::endl;
  if (p.Statistics)
    info << ":::Some statistics are stored each timestep at .dat files. If files already exist "
            "code append new data to them. At each line there are two numbers: time and "
            "quantity of interest, which are separated by a space"
         << std::endl;
  else
    info << ":::No statistics will be saved" << std::endl;
  info << std::endl;
  info << ":::Gas velocity field is storing each " << p.Data_freq_fast << " timesteps (=each "
       << p.Data_freq_fast * p.dt
       << " time units) at RU.bin and gas density, hydrodynamic pressure, particle "
          "concentration, and particle velocity field are storing each "
       << p.Data_freq_slow << " timesteps (=each " << p.Data_freq_slow * p.dt
       << " time units) at Rho.bin, P.bin, C.bin, and Vp.bin respectively. Moreover "
          "thermodynamic pressure (P0), current time, and number of ellapsed timesteps are "
          "being stored each "
       << p.Data_freq_fast
       << " timesteps at numbers.dat. All files have a name extension according to the "
          "snapshot number in which they have been saved. You have to add the prefix "
          "\"Restart_\" to files in order to inform the code to retart code using those data."
       << std::endl;
  info << std::endl;
  info << ":::Initial condition for this run is: ";
  if (p.Initial == 0) info << "ARTIFICIAL TURBULENT READ FROM U.bin, V.bin, and W.bin" << std::endl;
  if (p.Initial == 1)
    info << "RESTART FILES READ FROM Restart_RU.bin, Restart_Rho.bin, Restart_P.bin, "
            "Restart_C.bin, Restart_Vp.bin,and  Restart_numbers.dat"
         << std::endl;
  if (p.Initial == 2)
    info << "USER-DESIGNED INITIAL CONDITION DEFINED IN grid::Initialize()@grid.cpp" << std::endl;
  info << std::endl;
  if (p.cooling)
    info << ":::We are assuming zero net external energy from radiation, by introducing "
            "cooling term in the energy equation."
         << std::endl;
  else
    info << ":::There is a finite net external energy to the system from radiation." << std::endl;
  info << std::endl;
  info << ":::Here is the list of dimensional quantities read from the input file:" << std::endl;
  info << std::endl;
  info << "  DIMENSIONAL VARIABLE          BEFORE THRESHOLD                  AFTER THRESHOLD"
       << "         UNIT" << std::endl;
  info << "          ---------------------------------------------------------------------"
          "-------------"
       << std::endl;
  struct Row {
    const char* name;
    double a, b;
    const char* unit;
    bool dash;
  };
  const Row rows[] = {
      {"          Lx                      ", p.Lx, p.Lx, "           m", false},
      {"          Ly----------------------", p.Ly, p.Ly, "-----------m", true},
      {"          Lz                      ", p.Lz, p.Lz, "           m", false},
      {"          A-----------------------", p.A1, p.A2, "-----------1/s", true},
      {"         Rho0                     ", p.Rho0, p.Rho0, "           kg/m^3", false},
      {"          U0----------------------", p.U0, p.U0, "-----------m/s", true},
      {"          T0                      ", p.T0, p.T0, "           K", false},
      {"         Mu0----------------------", p.Mu0, p.Mu0, "-----------kg/(m.s)", true},
      {"          k                       ", p.k, p.k, "           kg.m/(K.s^3)", false},
      {"          gx----------------------", p.gx1, p.gx2, "-----------m/s^2", true},
      {"          gy                      ", p.gy1, p.gy2, "           m/s^2", false},
      {"          gz----------------------", p.gz1, p.gz2, "-----------m/s^2", true},
      {"          Cp                      ", p.Cp, p.Cp, "           J/(kg.K)", false},
      {"          Cv----------------------", p.Cv, p.Cv, "-----------J/(kg.K)", true},
      {"          n0                      ", p.np0, p.np0, "           1/m^3", false},
      {"          Dp----------------------", p.Dp, p.Dp, "-----------m", true},
      {"        Rho_p                     ", p.Rhop, p.Rhop, "           kg/m^3", false},
      {"         Cv_p---------------------", p.Cvp, p.Cvp, "-----------J/(kg.K)", true},
      {"       epsilon                    ", p.epsilonp, p.epsilon, "           ", false},
      {"          I0----------------------", p.I01, p.I02, "-----------J/(s.m^2)", true},
  };
  for (const Row& r : rows)
    info << r.name << std::scientific << r.a
         << (r.dash ? "----------------------" : "                      ") << std::scientific
         << r.b << r.unit << std::endl;
  info << std::endl;
  info << ":::Following dimensional quantities are resulted from the above:" << std::endl;
  info << std::endl;
  info << "  DIMENSIONAL VARIABLE                VALUE" << "                            UNIT"
       << std::endl;
  const double Vcell = g.dx * g.dy * g.dz;
  info << "          R-----------------------" << std::scientific << p.R
       << "------------------------J/(kg.K)" << std::endl;
  info << "          P0----------------------" << std::scientific << p.P0
       << "------------------------Pa" << std::endl;
  info << "         Vol----------------------" << std::scientific << p.Vol()
       << "------------------------m^3" << std::endl;
  info << "        Vcell---------------------" << std::scientific << Vcell
       << "------------------------m^3" << std::endl;
  info << "        Tau_p---------------------" << std::scientific << p.Tp
       << "------------------------s" << std::endl;
  info << "          mp----------------------" << std::scientific << p.mp
       << "------------------------kg" << std::endl;
  info << std::endl;
  info << ":::Following non-dimensional quantities are resulted from the above:" << std::endl;
  info << std::endl;
  info << "NONDIMENSIONAL VARIABLE              VALUE" << std::endl;
  info << "         Np0----------------------" << std::scientific << p.N0 << std::endl;
  info << "          Re----------------------" << std::scientific
       << p.Rho0 * p.U0 * p.W / p.Mu0 << std::endl;
  info << "          Fr----------------------" << std::scientific << p.U0 / sqrt(p.gz * p.W)
       << std::endl;
  info << "          Pr----------------------" << std::scientific << p.Mu0 * p.Cp / p.k
       << std::endl;
  info << "        gamma---------------------" << std::scientific << p.Cp / p.Cv << std::endl;
  info << "        GAMMA---------------------" << std::scientific << p.Cvp / p.Cv << std::endl;
  info << "          Nu----------------------" << std::scientific << p.Nu << std::endl;
  info << std::endl;
  info << ":::The poisson solver: Preconditioner: ";
  if (p.PreCond == 0) info << "NONE";
  if (p.PreCond == 1) info << "AMG";
  if (p.PreCond == 2) info << "Euclid";
  info << "   ,   Solver: ";
  const char* solvers[7] = {"FFT Based", "AMG", "PCG", "GMRES", "FlexGMRES", "LGMRES", "BiCGSTAB"};
  if (p.Solver >= 0 && p.Solver <= 6) info << solvers[p.Solver] << std::endl;
  info << std::endl;
  info << ":::In this run " << 1 << " processors have been exploited,  which are distributed in "
       << "3D as " << 1 << "x" << 1 << "x" << 1 << " ." << std::endl;
  info << std::endl;
  info << ":::What does each processor do:" << std::endl << std::endl;
  info.close();
  // the per-rank section (one rank: every neighbour is rank 0)
  std::ofstream a("info.txt", std::ios::out | std::ios::app);
  a << "  ***Processor number: " << 0 << std::endl;
  a << "  Location in CPU grid: (" << 0 << "," << 0 << "," << 0 << ") , Neighbor processes: TOP:"
    << 0 << " BOT:" << 0 << " RIGHT:" << 0 << " LEFT:" << 0 << " FRONT:" << 0 << " REAR:" << 0
    << std::endl;
  a << "  Local grid size: " << g.Nx << "x" << g.Ny << "x" << g.Nz << ", Global grid index: i:"
    << 0 << "->" << g.Nx_tot - 1 << " j:" << 0 << "->" << g.Ny_tot - 1 << " k:" << 0 << "->"
    << g.Nz_tot - 1 << "  ,  Global coordinate: x:" << g.xl << "->" << g.xh << " y:" << g.yl
    << "->" << g.yh << " z:" << g.zl << "->" << g.zh << std::endl;
  a << " I have initially " << part.Np << " particles." << std::endl << std::endl;
}

}  // namespace orc
