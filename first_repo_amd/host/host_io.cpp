// Run files of the solver, in the reference's layouts and with its paths
// (relative to the working directory): grid::Statistics (grid.cpp:714-866)
// computed on the device, grid::Store (233-313) and particle::Store_All
// (particle.cpp:867-917), the Initial = 1 restart (grid.cpp:124-182,
// particle.cpp:961-1025), grid::Write_info (869-973), touch.check
// (25-29, 77-98).  Field files are the global interior of each component,
// [Nz_tot][Ny_tot][Nx_tot] with x fastest (the MPI-IO subarray of
// communicator.cpp:29-39 on one rank); the duct's outlet files are the
// [Nz_tot][Ny_tot] slice at local i = Nx-1 (communicator.cpp:41-51).
#include <fcntl.h>
#include <unistd.h>

#include <cmath>
#include <fstream>
#include <sstream>

#include "fr_host.h"

using namespace fr;

#define FR_TRY(x)                \
  do {                           \
    int rc_ = (x);               \
    if (rc_) {                   \
      if (!err) err = rc_;       \
      return rc_;                \
    }                            \
  } while (0)

namespace {

bool put(const std::string& path, const void* p, size_t n, bool excl, bool append) {
  const int flags = O_WRONLY | O_CREAT | (excl ? O_EXCL : 0) | (append ? O_APPEND : 0);
  const int fd = open(path.c_str(), flags, 0644);
  if (fd < 0) return false;
  size_t done = 0;
  while (done < n) {
    const ssize_t w = write(fd, (const char*)p + done, n - done);
    if (w <= 0) break;
    done += (size_t)w;
  }
  close(fd);
  return done == n;
}

bool get(const std::string& path, void* p, size_t n, size_t offset) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  bool ok = fseek(f, (long)offset, SEEK_SET) == 0 && fread(p, 1, n, f) == n;
  fclose(f);
  return ok;
}

// interior of a ghosted host array, x fastest
void interior(const gridsize& s, const std::vector<double>& h, std::vector<double>& out) {
  for (int k = 1; k <= s.Nz_tot_; k++)
    for (int j = 1; j <= s.Ny_tot_; j++) {
      const double* row = h.data() + ((size_t)k * s.Ny() + j) * s.Nx() + 1;
      out.insert(out.end(), row, row + s.Nx_tot_);
    }
}
void outlet(const gridsize& s, const std::vector<double>& h, std::vector<double>& out) {
  for (int k = 1; k <= s.Nz_tot_; k++)
    for (int j = 1; j <= s.Ny_tot_; j++)
      out.push_back(h[((size_t)k * s.Ny() + j) * s.Nx() + s.Nx() - 1]);
}

std::string snap(const std::string& dir, const char* name, int n, const char* ext) {
  std::ostringstream o;
  o << dir << name << "_" << n << ext;
  return o.str();
}

const char* kStat[15] = {"TKE",           "C_Max", "C_Min", "C_Mean",   "Rho_Max",
                         "Rho_Min",       "Rho_Mean", "Max_CFL_Vp", "Max_CFL_U", "Num_Iteration",
                         "Balance_Index", "Uin",   "Uout",  "dx_o_eta", "ST"};

}  // namespace

// grid.cpp:25-49: touch.check holds 0; the stat files are opened for append
// (created when absent).
int grid::open_outputs() {
  {
    std::ofstream t("touch.check", std::ios::out | std::ios::trunc);
    t << 0;
    if (!t) return FR_EINVAL;
  }
  if (param.Statistics_)
    for (int q = 0; q < 15; q++) {
      std::ofstream f(param.Stat_dir_ + kStat[q] + ".dat", std::ios::out | std::ios::app);
      if (!f) return FR_EINVAL;
    }
  return FR_OK;
}

// grid.cpp:77-98: a 1 in touch.check stops the run after a store of both
// snapshot kinds.
int grid::Touch(bool* stop) {
  bool ans = false;
  std::ifstream t("touch.check");
  t >> ans;
  Is_touch = ans;
  *stop = ans;
  if (ans) return Store();
  return FR_OK;
}

// Every field the run files read, widened to FP64 in the precision form.
const double* grid::wide(tensor0& f, int slot) {
  if (!f.is_f32()) return f.pointer();
  tensor0& w = io_wide[slot];
  if (!w.pointer()) w.init(&size, st);
  to_double(f.p<float>(), w.pointer(), size.size(), st);
  return w.pointer();
}

int grid::download(tensor0& f, std::vector<double>& h, int slot) {
  h.resize((size_t)size.size());
  const double* p = wide(f, slot);
  FR_CHECK(hipMemcpyAsync(h.data(), p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

// grid::Store (grid.cpp:233-313).  Particle files are created exclusively
// (MPI_MODE_EXCL, particle.cpp:900, 913): a store over an existing snapshot
// fails with FR_EINVAL after writing what it could.
int grid::Store() {
  FR_TRY(settle_side());
  const std::string& dir = param.Data_dir_;
  bool ok = true;
  std::vector<double> h, blk;
  if (num_timestep == 1) {  // XYZ.bin: cell centres of every component
    const gridsize& s = size;
    for (int m = 0; m < 3; m++)
      for (int k = 0; k < s.Nz_tot_; k++)
        for (int j = 0; j < s.Ny_tot_; j++)
          for (int i = 0; i < s.Nx_tot_; i++)
            blk.push_back(m == 0 ? s.dx() / 2. + i * s.dx()
                                 : (m == 1 ? s.dy() / 2. + j * s.dy() : s.dz() / 2. + k * s.dz()));
    ok &= put(dir + "XYZ.bin", blk.data(), blk.size() * 8, false, false);
  }
  if ((num_timestep % param.Data_freq_fast_ == 0) || Is_touch) {
    blk.clear();
    std::vector<double> sl;
    for (int m = 0; m < 3; m++) {
      FR_TRY(download(RU.c(m), h, 0));
      interior(size, h, blk);
      outlet(size, h, sl);
    }
    ok &= put(snap(dir, "RU", num_timestep, ".bin"), blk.data(), blk.size() * 8, false, false);
    if (!param.IsBox())
      ok &= put(snap(dir, "RU_S", num_timestep, ".bin"), sl.data(), sl.size() * 8, false, false);
    std::ofstream numbers(snap(dir, "numbers", num_timestep, ".dat"));
    numbers << P0 << std::endl;
    numbers << T_cur << std::endl;
    numbers << num_timestep << std::endl;
    numbers << part.dev.np << std::endl;
    ok &= (bool)numbers;
  }
  if ((num_timestep % param.Data_freq_slow_ == 0) || Is_touch) {
    const int np = part.dev.np;
    const char* names[7] = {"Part_x", "Part_y", "Part_z", "Part_u", "Part_v", "Part_w",
                            "Part_T"};
    std::vector<double> v((size_t)np);
    for (int q = 0; q < 7; q++) {
      if (np) FR_TRY(part.dev.get(v.data(), part.dev.s[q][0], st));
      ok &= put(snap(dir, names[q], num_timestep, ".bin"), v.data(), 8 * (size_t)np, true,
                false);
    }
    std::vector<int> fl((size_t)np);
    if (np) FR_TRY(part.dev.get_flags(fl.data(), st));
    ok &= put(snap(dir, "Part_flag", num_timestep, ".bin"), fl.data(), 4 * (size_t)np, true,
              false);
    blk.clear();
    FR_TRY(download(C, h, 0));
    interior(size, h, blk);
    ok &= put(snap(dir, "C", num_timestep, ".bin"), blk.data(), blk.size() * 8, false, false);
    blk.clear();
    std::vector<double> sl;
    FR_TRY(download(Rho, h, 0));
    interior(size, h, blk);
    outlet(size, h, sl);
    ok &= put(snap(dir, "Rho", num_timestep, ".bin"), blk.data(), blk.size() * 8, false, false);
    if (!param.IsBox())
      ok &= put(snap(dir, "Rho_S", num_timestep, ".bin"), sl.data(), sl.size() * 8, false, false);
  }
  return ok ? FR_OK : FR_EINVAL;
}

// The Initial = 1 restart of grid::Initialize (grid.cpp:124-182) and
// particle::Load_All (particle.cpp:961-1025) on one rank: fields from
// Restart_RU[_IO].bin and Restart_Rho[_IO].bin (the duct also its outlet
// slices, Rho0 and the box's first-interior v, w in its inlet ghosts -- taken
// from the RU_np1 inflow planes handed over last), P = 0, the numbers file,
// the particles' base stage, and a particle exchange on x_new = x.
int grid::Restart() {
  const bool box = param.IsBox();
  const std::string sfx = box ? "" : "_IO";
  const gridsize& s = size;
  const size_t nint = (size_t)s.Nx_tot_ * s.Ny_tot_ * s.Nz_tot_;
  const size_t nsl = (size_t)s.Ny_tot_ * s.Nz_tot_;
  std::vector<double> blk(nint), sl(nsl), h((size_t)s.size());
  auto fill = [&](std::vector<double>& g, const double* b) {
    size_t n = 0;
    for (int k = 1; k <= s.Nz_tot_; k++)
      for (int j = 1; j <= s.Ny_tot_; j++)
        for (int i = 1; i <= s.Nx_tot_; i++) g[((size_t)k * s.Ny() + j) * s.Nx() + i] = b[n++];
  };
  auto fill_outlet = [&](std::vector<double>& g, const double* b) {
    size_t n = 0;
    for (int k = 1; k <= s.Nz_tot_; k++)
      for (int j = 1; j <= s.Ny_tot_; j++) g[((size_t)k * s.Ny() + j) * s.Nx() + s.Nx() - 1] = b[n++];
  };
  for (int m = 0; m < 4; m++) {  // RU x, y, z, then Rho
    std::fill(h.begin(), h.end(), 0.0);
    const std::string f = m < 3 ? "Restart_RU" + sfx + ".bin" : "Restart_Rho" + sfx + ".bin";
    if (!get(f, blk.data(), nint * 8, (m < 3 ? m : 0) * nint * 8)) return FR_EINVAL;
    fill(h, blk.data());
    if (!box) {
      const std::string fs = m < 3 ? "Restart_RU_IO_S.bin" : "Restart_Rho_IO_S.bin";
      if (!get(fs, sl.data(), nsl * 8, (m < 3 ? m : 0) * nsl * 8)) return FR_EINVAL;
      fill_outlet(h, sl.data());
      for (int k = 0; k < s.Nz(); k++)
        for (int j = 0; j < s.Ny(); j++) {
          double& g0 = h[((size_t)k * s.Ny() + j) * s.Nx()];
          if (m == 3) g0 = param.Rho0();
          if ((m == 1 || m == 2) && !inflow_h.empty())
            g0 = inflow_h[(size_t)(5 + m) * s.Ny() * s.Nz() + (size_t)k * s.Ny() + j];
        }
    }
    tensor0& t = m < 3 ? RU.c(m) : Rho;
    t.upload(h.data());
    t.Update_Ghosts();
  }
  P.assign(0);
  int np = param.N0();
  {
    std::ifstream numbers(box ? "Restart_numbers.dat" : "Restart_numbers_IO.dat");
    numbers >> P0;
    numbers >> T_cur;
    numbers >> num_timestep;
    if (!box) {
      numbers >> np;
      param.N0_ = np;  // params::Set_Np
    }
    if (!numbers) return FR_EINVAL;
  }
  if (np > part.dev.cap) return FR_EOVERFLOW;
  PartDev& p = part.dev;
  p.np = part.Np = np;
  FR_TRY(p.reset_order(st));
  std::vector<double> v((size_t)np);
  const char* names[7] = {"Part_x", "Part_y", "Part_z", "Part_u", "Part_v", "Part_w", "Part_T"};
  for (int q = 0; q < 7; q++) {
    if (!get("Restart_" + std::string(names[q]) + sfx + ".bin", v.data(), 8 * (size_t)np, 0))
      return FR_EINVAL;
    if (np) FR_TRY(p.set(v.data(), p.s[q][0], st));
  }
  std::vector<int> fl((size_t)np);
  if (!get("Restart_Part_flag" + sfx + ".bin", fl.data(), 4 * (size_t)np, 0)) return FR_EINVAL;
  if (np) FR_TRY(p.set_flags(fl.data(), st));
  for (int q = 0; q < 3; q++)  // x_new = x (particle.cpp:993)
    if (np)
      FR_CHECK(hipMemcpyAsync(p.s[q][2], p.s[q][0], sizeof(double) * np,
                              hipMemcpyDeviceToDevice, st));
  if (np) FR_TRY(part.Send_Recv());
  FR_CHECK(hipStreamSynchronize(st));
  return FR_OK;
}

// grid::Statistics (grid.cpp:714-866) on the device: the x-profiles of the
// 32 .bin files appended with Nx_tot doubles each, the 14 .dat lines
// ("%g %g", the stream's default format) and TIME.bin.  The operators keep
// the reference's expressions (stats.hip); plane means are tree sums, so a
// profile agrees with the reference's loop to rounding.
int grid::Statistics() {
  if (!param.Statistics_) return FR_OK;
  FR_TRY(fill_np1());
  FR_TRY(fill_faces());
  FR_TRY(settle_side());
  const Dev d = size.dev();
  const int Nx = size.Nx();
  for (tensor1* t : {&st_a, &st_b, &st_c})
    if (!t->x.pointer()) t->init(&size, st);
  if (!dprof) FR_CHECK(hipMalloc(&dprof, sizeof(double) * 10 * (size_t)Nx));
  double* pr = dprof;             // profiles of the field at hand (up to three)
  double* pa = dprof + 3 * Nx;    // dummy's three profiles (RU at cell centres)
  double* pb = dprof + 6 * Nx;    // dummy2's three (RU/Rho)
  double* prho = dprof + 9 * Nx;  // Rho's
  std::vector<double> hp((size_t)Nx * 3);
  auto fetch = [&](const double* dp, int n) -> int {
    FR_CHECK(hipMemcpyAsync(hp.data(), dp, sizeof(double) * Nx * n, hipMemcpyDeviceToHost, st));
    FR_CHECK(hipStreamSynchronize(st));
    return FR_OK;
  };
  auto bin = [&](const char* name, const double* prof) {
    put(param.Stat_dir_ + name + ".bin", prof + 1, sizeof(double) * size.Nx_tot_, false, true);
  };
  auto red1 = [&](const double* f, double* out) { return reducex(d, f, out, st); };
  auto mean = [&](const double* f) {
    double r = 0;
    reduce_sync(d, f, 1, &r, st);
    return r;
  };
  auto red3 = [&](const double* f, int op) {
    double r = 0;
    reduce_sync(d, f, op, &r, st);
    return r;
  };
  const double* rho = wide(Rho, 0);
  const double* c = wide(C, 1);
  const double* q = wide(RHS_Part_Temp, 2);
  const double* ru[3] = {wide(RU.x, 3), wide(RU.y, 4), wide(RU.z, 5)};
  const double* rn[3] = {wide(RU_np1.x, 6), wide(RU_np1.y, 7), wide(RU_np1.z, 8)};
  const double* rf[3] = {wide(Rho_face.x, 9), wide(Rho_face.y, 10), wide(Rho_face.z, 11)};
  const double Rho_mean = mean(rho), Rho_max = red3(rho, 2), Rho_min = red3(rho, 3);
  const double C_max = red3(c, 2), C_min = red3(c, 3), C_mean = mean(c);
  double* A[3] = {st_a.x.pointer(), st_a.y.pointer(), st_a.z.pointer()};  // U
  double* B[3] = {st_b.x.pointer(), st_b.y.pointer(), st_b.z.pointer()};  // dummy, dummy3
  double* D2[3] = {st_c.x.pointer(), st_c.y.pointer(), st_c.z.pointer()}; // dummy2
  const double* cA[3] = {A[0], A[1], A[2]};
  const double* cB[3] = {B[0], B[1], B[2]};
  const double* cD2[3] = {D2[0], D2[1], D2[2]};
  const double* none[3] = {nullptr, nullptr, nullptr};
  FR_TRY(stat_point(d, ST_DIV1, A, rn, rf, nullptr, nullptr, nullptr, 0, st));  // U
  double cfl = 0;
  const double dd[3] = {size.dx(), size.dy(), size.dz()};
  for (int m = 0; m < 3; m++) {
    const double v = red3(cA[m], 2) * param.dt() / dd[m];
    cfl = m == 0 ? v : (v > cfl ? v : cfl);
  }
  const double lnp = part.dev.np;
  const double Load_Balance = lnp / lnp;  // particle::Balance_Index on one rank
  FR_TRY(red1(rho, prho));
  FR_TRY(fetch(prho, 1));
  bin("RHO", hp.data());
  double* b0[3] = {B[0], nullptr, nullptr};
  FR_TRY(stat_point(d, ST_RSD, b0, none, none, rho, nullptr, prho, 0, st));
  FR_TRY(red1(B[0], pr));
  FR_TRY(fetch(pr, 1));
  for (int i = 0; i < Nx; i++) hp[i] = sqrt(hp[i]);
  bin("RHO_SD", hp.data());
  for (int m = 0; m < 3; m++) FR_TRY(fr_t0_interp(&size.g, B[m], ru[m], m, 1, st));  // dummy
  FR_TRY(stat_point(d, ST_DIV0, D2, none, ru, rho, nullptr, nullptr, 0, st));      // dummy2
  for (int m = 0; m < 3; m++) FR_TRY(red1(B[m], pa + m * Nx));
  FR_TRY(fetch(pa, 3));
  bin("RU_x", hp.data());
  bin("RU_y", hp.data());  // the reference stores the x component for y and z
  bin("RU_z", hp.data());
  for (int m = 0; m < 3; m++) FR_TRY(red1(D2[m], pb + m * Nx));
  FR_TRY(fetch(pb, 3));
  bin("U_x", hp.data());
  bin("U_y", hp.data() + Nx);
  bin("U_z", hp.data() + 2 * Nx);
  // dummy3 = B (dummy's values are no longer needed once its profiles are in pa)
  const char* names[5][3] = {{"TKE2x", "TKE2y", "TKE2z"},
                             {"UiUi2x", "UiUi2y", "UiUi2z"},
                             {"Cor_R_Ux", "Cor_R_Uy", "Cor_R_Uz"},
                             {"TKEx", "TKEy", "TKEz"},
                             {"UiUix", "UiUiy", "UiUiz"}};
  const int ops[5] = {ST_TKE2, ST_UIUI2, ST_CORRU, ST_TKE, ST_UIUI};
  double TKE = 0;
  for (int o = 0; o < 5; o++) {
    const double* ra = o < 2 ? pa : pb;  // RU reduced for TKE2/UiUi2, U reduced after
    FR_TRY(stat_point(d, ops[o], B, cD2, none, rho, ra, prho, 0, st));
    for (int m = 0; m < 3; m++) FR_TRY(red1(B[m], pr + m * Nx));
    FR_TRY(fetch(pr, 3));
    for (int m = 0; m < 3; m++) bin(names[o][m], hp.data() + m * Nx);
    if (ops[o] == ST_TKE) TKE = mean(cB[0]) + mean(cB[1]) + mean(cB[2]);
  }
  FR_TRY(red1(c, pr));
  FR_TRY(fetch(pr, 1));
  bin("C_avg", hp.data());
  FR_TRY(red1(P.pointer(), pr));
  FR_TRY(fetch(pr, 1));
  bin("P_avg", hp.data());
  FR_TRY(stat_strain(d, false, cA, D2[0], st));  // SijSij into dummy2.x's interior
  FR_TRY(stat_strain(d, true, cA, D2[1], st));   // WijWij added to dummy2.y
  FR_TRY(red1(D2[0], pr));
  FR_TRY(fetch(pr, 1));
  bin("SijSij", hp.data());
  FR_TRY(red1(D2[1], pr));
  FR_TRY(fetch(pr, 1));
  bin("WijWij", hp.data());
  const double* x0[3] = {D2[0], nullptr, nullptr};
  double* z0[3] = {D2[2], nullptr, nullptr};
  FR_TRY(stat_point(d, ST_ST, z0, x0, none, nullptr, nullptr, nullptr,
                    2 * param.Tp() * param.Tp(), st));
  FR_TRY(red1(D2[2], pr));
  FR_TRY(fetch(pr, 1));
  bin("ST", hp.data());
  const double avg_ST = mean(cD2[2]);
  FR_TRY(stat_point(d, ST_NU, z0, none, none, rho, nullptr, nullptr, param.Mu0(), st));
  FR_TRY(red1(D2[2], pr));
  FR_TRY(fetch(pr, 1));
  bin("NU", hp.data());
  double* y0[3] = {D2[1], nullptr, nullptr};
  const double* zc[3] = {D2[2], nullptr, nullptr};
  FR_TRY(stat_point(d, ST_DISS, y0, zc, x0, nullptr, nullptr, nullptr, 0, st));
  FR_TRY(red1(D2[1], pr));
  FR_TRY(fetch(pr, 1));
  bin("DISSIPATION", hp.data());
  double* x0w[3] = {D2[0], nullptr, nullptr};
  const double* yc[3] = {D2[1], nullptr, nullptr};
  FR_TRY(stat_point(d, ST_ETA, x0w, zc, yc, nullptr, nullptr, nullptr, 0, st));
  FR_TRY(red1(D2[0], pr));
  FR_TRY(fetch(pr, 1));
  bin("ETA", hp.data());
  const double dx_over_eta_min = size.dx() / red3(cD2[0], 3);
  FR_TRY(red1(q, pr));
  FR_TRY(fetch(pr, 1));
  bin("HT", hp.data());
  const double v[15] = {TKE,     C_max, C_min, C_mean,       Rho_max,   Rho_min,
                        Rho_mean, 0,    cfl,   0,            Load_Balance, U_avg_in,
                        U_avg_out, dx_over_eta_min, avg_ST};
  for (int k = 0; k < 15; k++) {
    if (k == 7) continue;  // Max_CFL_Vp.dat is opened but never written
    std::ofstream f(param.Stat_dir_ + kStat[k] + ".dat", std::ios::out | std::ios::app);
    if (k == 9)
      f << T_cur << " " << PS_.num_iteration() << std::endl;
    else
      f << T_cur << " " << v[k] << std::endl;
  }
  put(param.Stat_dir_ + "TIME.bin", &T_cur, sizeof(double), false, true);
  return (int)hipGetLastError();
}

// grid::Write_info (grid.cpp:869-973) for one rank.
int grid::Write_info() {
  std::ofstream info("info.txt");
  const params& p = param;
  const gridsize& s = size;
  info << ":::Global grid is " << s.Nx_tot_ << "x" << s.Ny_tot_ << "x" << s.Nz_tot_
       << " with bordersize (number of ghost cells around local grid) =" << s.bs() << std::endl;
  info << std::endl;
  info << ":::Code started at t=" << T_cur << " with dt=" << p.dt()
       << " ,and is supposed to finish at t=" << p.T_final() << std::endl;
  info << std::endl;
  if (p.Statistics_)
    info << ":::Some statistics are stored each timestep at .dat files. If files already exist "
            "code append new data to them. At each line there are two numbers: time and "
            "quantity of interest, which are separated by a space"
         << std::endl;
  else
    info << ":::No statistics will be saved" << std::endl;
  info << std::endl;
  info << ":::Gas velocity field is storing each " << p.Data_freq_fast_ << " timesteps (=each "
       << p.Data_freq_fast_ * p.dt()
       << " time units) at RU.bin and gas density, hydrodynamic pressure, particle "
          "concentration, and particle velocity field are storing each "
       << p.Data_freq_slow_ << " timesteps (=each " << p.Data_freq_slow_ * p.dt()
       << " time units) at Rho.bin, P.bin, C.bin, and Vp.bin respectively. Moreover "
          "thermodynamic pressure (P0), current time, and number of ellapsed timesteps are "
          "being stored each "
       << p.Data_freq_fast_
       << " timesteps at numbers.dat. All files have a name extension according to the "
          "snapshot number in which they have been saved. You have to add the prefix "
          "\"Restart_\" to files in order to inform the code to retart code using those data."
       << std::endl;
  info << std::endl;
  info << ":::Initial condition for this run is: ";
  if (p.Initial() == 0) info << "ARTIFICIAL TURBULENT READ FROM U.bin, V.bin, and W.bin" << std::endl;
  if (p.Initial() == 1)
    info << "RESTART FILES READ FROM Restart_RU.bin, Restart_Rho.bin, Restart_P.bin, "
            "Restart_C.bin, Restart_Vp.bin,and  Restart_numbers.dat"
         << std::endl;
  if (p.Initial() == 2)
    info << "USER-DESIGNED INITIAL CONDITION DEFINED IN grid::Initialize()@grid.cpp" << std::endl;
  info << std::endl;
  if (p.Is_Cooling())
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
      {"          Lx                      ", p.Lx(), p.Lx(), "           m", false},
      {"          Ly----------------------", p.Ly(), p.Ly(), "-----------m", true},
      {"          Lz                      ", p.Lz(), p.Lz(), "           m", false},
      {"          A-----------------------", p.A1_, p.A2_, "-----------1/s", true},
      {"         Rho0                     ", p.Rho0(), p.Rho0(), "           kg/m^3", false},
      {"          U0----------------------", p.U0(), p.U0(), "-----------m/s", true},
      {"          T0                      ", p.T0(), p.T0(), "           K", false},
      {"         Mu0----------------------", p.Mu0(), p.Mu0(), "-----------kg/(m.s)", true},
      {"          k                       ", p.k(), p.k(), "           kg.m/(K.s^3)", false},
      {"          gx----------------------", p.gx1_, p.gx2_, "-----------m/s^2", true},
      {"          gy                      ", p.gy1_, p.gy2_, "           m/s^2", false},
      {"          gz----------------------", p.gz1_, p.gz2_, "-----------m/s^2", true},
      {"          Cp                      ", p.Cp(), p.Cp(), "           J/(kg.K)", false},
      {"          Cv----------------------", p.Cv(), p.Cv(), "-----------J/(kg.K)", true},
      {"          n0                      ", p.np0_, p.np0_, "           1/m^3", false},
      {"          Dp----------------------", p.Dp(), p.Dp(), "-----------m", true},
      {"        Rho_p                     ", p.Rhop_, p.Rhop_, "           kg/m^3", false},
      {"         Cv_p---------------------", p.Cvp(), p.Cvp(), "-----------J/(kg.K)", true},
      {"       epsilon                    ", p.epsilonp(), p.epsilon(), "           ", false},
      {"          I0----------------------", p.I01_, p.I02_, "-----------J/(s.m^2)", true},
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
  info << "          R-----------------------" << std::scientific << p.R()
       << "------------------------J/(kg.K)" << std::endl;
  info << "          P0----------------------" << std::scientific << p.P0()
       << "------------------------Pa" << std::endl;
  info << "         Vol----------------------" << std::scientific << p.Vol()
       << "------------------------m^3" << std::endl;
  info << "        Vcell---------------------" << std::scientific << s.Vcell()
       << "------------------------m^3" << std::endl;
  info << "        Tau_p---------------------" << std::scientific << p.Tp()
       << "------------------------s" << std::endl;
  info << "          mp----------------------" << std::scientific << p.mp()
       << "------------------------kg" << std::endl;
  info << std::endl;
  info << ":::Following non-dimensional quantities are resulted from the above:" << std::endl;
  info << std::endl;
  info << "NONDIMENSIONAL VARIABLE              VALUE" << std::endl;
  info << "         Np0----------------------" << std::scientific << p.N0() << std::endl;
  info << "          Re----------------------" << std::scientific
       << p.Rho0() * p.U0() * p.W_ / p.Mu0() << std::endl;
  info << "          Fr----------------------" << std::scientific << p.U0() / sqrt(p.gz() * p.W_)
       << std::endl;
  info << "          Pr----------------------" << std::scientific << p.Mu0() * p.Cp() / p.k()
       << std::endl;
  info << "        gamma---------------------" << std::scientific << p.Cp() / p.Cv() << std::endl;
  info << "        GAMMA---------------------" << std::scientific << p.Cvp() / p.Cv() << std::endl;
  info << "          Nu----------------------" << std::scientific << p.Nu() << std::endl;
  info << std::endl;
  info << ":::The poisson solver: Preconditioner: ";
  if (p.PreCond_ == 0) info << "NONE";
  if (p.PreCond_ == 1) info << "AMG";
  if (p.PreCond_ == 2) info << "Euclid";
  info << "   ,   Solver: ";
  const char* solvers[7] = {"FFT Based", "AMG", "PCG", "GMRES", "FlexGMRES", "LGMRES", "BiCGSTAB"};
  if (p.Solver() >= 0 && p.Solver() <= 6) info << solvers[p.Solver()] << std::endl;
  info << std::endl;
  info << ":::In this run " << 1 << " processors have been exploited,  which are distributed in "
       << "3D as " << 1 << "x" << 1 << "x" << 1 << " ." << std::endl;
  info << std::endl;
  info << ":::What does each processor do:" << std::endl << std::endl;
  info.close();
  std::ofstream a("info.txt", std::ios::out | std::ios::app);
  a << "  ***Processor number: " << 0 << std::endl;
  a << "  Location in CPU grid: (" << 0 << "," << 0 << "," << 0 << ") , Neighbor processes: TOP:"
    << 0 << " BOT:" << 0 << " RIGHT:" << 0 << " LEFT:" << 0 << " FRONT:" << 0 << " REAR:" << 0
    << std::endl;
  a << "  Local grid size: " << s.Nx() << "x" << s.Ny() << "x" << s.Nz()
    << ", Global grid index: i:" << 0 << "->" << s.Nx_tot_ - 1 << " j:" << 0 << "->"
    << s.Ny_tot_ - 1 << " k:" << 0 << "->" << s.Nz_tot_ - 1 << "  ,  Global coordinate: x:"
    << s.xl_ << "->" << s.xh_ << " y:" << s.yl_ << "->" << s.yh_ << " z:" << s.zl_ << "->"
    << s.zh_ << std::endl;
  a << " I have initially " << part.dev.np << " particles." << std::endl << std::endl;
  return a ? FR_OK : FR_EINVAL;
}
