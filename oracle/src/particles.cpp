// Oracle particles: gather, ODE updates, deposits, one-rank migration order,
// random placement.  TEST INFRASTRUCTURE.  Restates particle.cpp.
#include <cstdio>

#include "orc.h"

namespace orc {

// glibc random_r TYPE_3: r[0]=seed; r[i]=16807*r[i-1] mod (2^31-1) for
// i<31 (Schrage); r[i]=r[i-31] for 31<=i<34; r[i]=r[i-31]+r[i-3] (mod 2^32)
// thereafter; the first 310 outputs are discarded; rand() = r[i] >> 1.
void GlibcRand::seed(unsigned s) {
  int32_t r[34];
  r[0] = (int32_t)s;
  for (int i = 1; i < 31; i++) {
    long hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
    long word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r[i] = (int32_t)word;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int i = 0; i < 34; i++) ring[i] = (uint32_t)r[i];
  head = 0;
  for (int i = 34; i < 344; i++) next();
}
int GlibcRand::next() {
  // ring[head] = r[i-34]; r[i-31] = ring[head+3]; r[i-3] = ring[head+31]
  uint32_t val = ring[(head + 3) % 34] + ring[(head + 31) % 34];
  ring[head] = val;
  head = (head + 1) % 34;
  return (int)(val >> 1);
}

void Particles::init(const Params* p, const Geom* gg) {
  par = p;
  g = gg;
  idx = 1. / g->dx;  // particle.cpp:19-21
  idy = 1. / g->dy;
  idz = 1. / g->dz;
  Np = 0;
}

std::vector<double>* Particles::stage(int q, int s) {
  std::vector<double>* tab[7][4] = {
      {&x, &x_int, &x_new, &x_np1}, {&y, &y_int, &y_new, &y_np1},
      {&z, &z_int, &z_new, &z_np1}, {&u, &u_int, &u_new, &u_np1},
      {&v, &v_int, &v_new, &v_np1}, {&w, &w_int, &w_new, &w_np1},
      {&T, &T_int, &T_new, &T_np1}};
  return tab[q][s];
}

void Particles::copy_stage(int dst, int src) {
  for (int q = 0; q < 7; q++) *stage(q, dst) = *stage(q, src);
}

void Particles::RESIZE() {  // particle.cpp:136-148 (all vectors resized)
  for (int q = 0; q < 7; q++)
    for (int s = 0; s < 4; s++) stage(q, s)->resize(Np);
  flag.resize(Np);
  RHS_u.resize(Np);
  RHS_v.resize(Np);
  RHS_w.resize(Np);
  RHS_T.resize(Np);
  ug.resize(Np);
  vg.resize(Np);
  wg.resize(Np);
  Tg.resize(Np);
}

// particle.cpp:314-338
void Particles::update_position(int k, double pre, double post) {
  double dt = par->dt;
  for (int i = 0; i < Np; i++) {
    x_np1[i] += dt * post * u_int[i];
    y_np1[i] += dt * post * v_int[i];
    z_np1[i] += dt * post * w_int[i];
  }
  if (k != 3)
    for (int i = 0; i < Np; i++) {
      x_new[i] = x[i] + dt * pre * u_int[i];
      y_new[i] = y[i] + dt * pre * v_int[i];
      z_new[i] = z[i] + dt * pre * w_int[i];
    }
  else
    for (int i = 0; i < Np; i++) {
      x_new[i] = x_np1[i];
      y_new[i] = y_np1[i];
      z_new[i] = z_np1[i];
    }
}

// particle.cpp:340-360 with x2i/y2j/z2k (106-119): truncating cell index,
// 1-D linear interpolation of each face component along its own axis.
void Particles::gas2part_velocity(const F1& U) {
  for (int i = 0; i < Np; i++) {
    int ii = (int)(x_int[i] * idx) - g->il + g->bs;
    int jj = (int)(y_int[i] * idy) - g->jl + g->bs;
    int kk = (int)(z_int[i] * idz) - g->kl + g->bs;
    double x_l = (x_int[i] * idx - ((int)(x_int[i] * idx)));
    double y_l = (y_int[i] * idy - ((int)(y_int[i] * idy)));
    double z_l = (z_int[i] * idz - ((int)(z_int[i] * idz)));
    ug[i] = x_l * U.x(ii + 1, jj, kk) + (1 - x_l) * U.x(ii, jj, kk);
    vg[i] = y_l * U.y(ii, jj + 1, kk) + (1 - y_l) * U.y(ii, jj, kk);
    wg[i] = z_l * U.z(ii, jj, kk + 1) + (1 - z_l) * U.z(ii, jj, kk);
  }
}

// particle.cpp:362-399
void Particles::update_velocity(int k, double pre, double post) {
  double iTp = 0;
  if (par->Tp != 0) iTp = 1. / par->Tp;
  const bool pp = par->per_particle_Tp();
  const double span = par->Tp_hi - par->Tp_lo;
  for (int i = 0; i < Np; i++) {
    // per-particle extension: 1/Tp of this particle's flag
    const double it = pp ? 1. / TpOfFlag(flag[i], par->Tp_lo, span) : iTp;
    RHS_u[i] = (ug[i] - u_int[i]) * it;
    RHS_v[i] = (vg[i] - v_int[i]) * it;
    RHS_w[i] = (wg[i] - w_int[i]) * it;
  }
  double dt = par->dt;
  double gx = par->ParticleGravity * par->gx;
  double gy = par->ParticleGravity * par->gy;
  double gz = par->ParticleGravity * par->gz;
  for (int i = 0; i < Np; i++) {
    u_np1[i] += dt * post * (RHS_u[i] + gx);
    v_np1[i] += dt * post * (RHS_v[i] + gy);
    w_np1[i] += dt * post * (RHS_w[i] + gz);
  }
  if (k != 3)
    for (int i = 0; i < Np; i++) {
      u_new[i] = u[i] + dt * pre * (RHS_u[i] + gx);
      v_new[i] = v[i] + dt * pre * (RHS_v[i] + gy);
      w_new[i] = w[i] + dt * pre * (RHS_w[i] + gz);
    }
  else
    for (int i = 0; i < Np; i++) {
      u_new[i] = u_np1[i];
      v_new[i] = v_np1[i];
      w_new[i] = w_np1[i];
    }
}

// particle.cpp:402-456 with x2i_round (121-134): rounded index of the
// upper cell centre, trilinear weights on cell centres.
void Particles::gas2part_Temp(const F0& Tf, bool use_new) {
  const std::vector<double>& X = use_new ? x_new : x_int;
  const std::vector<double>& Y = use_new ? y_new : y_int;
  const std::vector<double>& Z = use_new ? z_new : z_int;
  for (int i = 0; i < Np; i++) {
    int ii = (int)(round(X[i] * idx));
    int jj = (int)(round(Y[i] * idy));
    int kk = (int)(round(Z[i] * idz));
    double x_l = X[i] * idx - ii + 0.5;
    double y_l = Y[i] * idy - jj + 0.5;
    double z_l = Z[i] * idz - kk + 0.5;
    ii += g->bs - g->il;
    jj += g->bs - g->jl;
    kk += g->bs - g->kl;
    Tg[i] = x_l * y_l * z_l * Tf(ii, jj, kk) +
            (1 - x_l) * y_l * z_l * Tf(ii - 1, jj, kk) + x_l * (1 - y_l) * z_l * Tf(ii, jj - 1, kk) +
            x_l * y_l * (1 - z_l) * Tf(ii, jj, kk - 1) +
            (1 - x_l) * (1 - y_l) * z_l * Tf(ii - 1, jj - 1, kk) +
            (1 - x_l) * y_l * (1 - z_l) * Tf(ii - 1, jj, kk - 1) +
            x_l * (y_l - 1) * (z_l - 1) * Tf(ii, jj - 1, kk - 1) +
            (1 - x_l) * (1 - y_l) * (1 - z_l) * Tf(ii - 1, jj - 1, kk - 1);
  }
}

// particle.cpp:460-478
void Particles::Compute_RHS_Temp(bool use_new) {
  double iTth = par->Nu * PI * par->Dp * par->k;
  const std::vector<double>& Tp_ = use_new ? T_new : T_int;
  for (int i = 0; i < Np; i++) RHS_T[i] = (Tg[i] - Tp_[i]) * iTth;
}

// particle.cpp:480-496
void Particles::update_Temp(int k, double pre, double post) {
  double alpha = 0.25 * par->epsilonp * PI * par->Dp * par->Dp * par->I0;
  double dt = par->dt;
  double iS = 1. / (par->mp * par->Cvp);
  for (int i = 0; i < Np; i++) T_np1[i] += dt * post * iS * (RHS_T[i] + alpha);
  if (k != 3)
    for (int i = 0; i < Np; i++) T_new[i] = T[i] + dt * pre * iS * (RHS_T[i] + alpha);
  else
    for (int i = 0; i < Np; i++) T_new[i] = T_np1[i];
}

// particle.cpp:498-524
void Particles::part2gas_velocity(F1& A) {
  A.set(0);
  for (int i = 0; i < Np; i++) {
    int ii = (int)(x_int[i] * idx) - g->il + g->bs;
    int jj = (int)(y_int[i] * idy) - g->jl + g->bs;
    int kk = (int)(z_int[i] * idz) - g->kl + g->bs;
    double x_l = (x_int[i] * idx - ((int)(x_int[i] * idx)));
    double y_l = (y_int[i] * idy - ((int)(y_int[i] * idy)));
    double z_l = (z_int[i] * idz - ((int)(z_int[i] * idz)));
    A.x(ii + 1, jj, kk) += x_l * RHS_u[i];
    A.x(ii, jj, kk) += (1 - x_l) * RHS_u[i];
    A.y(ii, jj + 1, kk) += y_l * RHS_v[i];
    A.y(ii, jj, kk) += (1 - y_l) * RHS_v[i];
    A.z(ii, jj, kk + 1) += z_l * RHS_w[i];
    A.z(ii, jj, kk) += (1 - z_l) * RHS_w[i];
  }
  A.Update_Ghosts_CUM();
  A.Update_Ghosts();
}

// particle.cpp:526-559 (no halo after the cumulative pass)
void Particles::part2gas_concentration(F0& A) {
  A.set(0);
  for (int i = 0; i < Np; i++) {
    int ii = (int)(round(x_int[i] * idx));
    int jj = (int)(round(y_int[i] * idy));
    int kk = (int)(round(z_int[i] * idz));
    double x_l = x_int[i] * idx - ii + 0.5;
    double y_l = y_int[i] * idy - jj + 0.5;
    double z_l = z_int[i] * idz - kk + 0.5;
    ii += (g->bs - g->il);
    jj += (g->bs - g->jl);
    kk += (g->bs - g->kl);
    A(ii, jj, kk) += (x_l * y_l * z_l);
    A(ii - 1, jj, kk) += (1 - x_l) * y_l * z_l;
    A(ii, jj - 1, kk) += x_l * (1 - y_l) * z_l;
    A(ii, jj, kk - 1) += x_l * y_l * (1 - z_l);
    A(ii - 1, jj - 1, kk) += (1 - x_l) * (1 - y_l) * z_l;
    A(ii - 1, jj, kk - 1) += (1 - x_l) * y_l * (1 - z_l);
    A(ii, jj - 1, kk - 1) += x_l * (1 - y_l) * (1 - z_l);
    A(ii - 1, jj - 1, kk - 1) += (1 - x_l) * (1 - y_l) * (1 - z_l);
  }
  A.Update_Ghosts_CUM();
}

// particle.cpp:586-620 (int) and 624-658 (new)
void Particles::part2gas_Temp(F0& A, bool use_new) {
  const std::vector<double>& X = use_new ? x_new : x_int;
  const std::vector<double>& Y = use_new ? y_new : y_int;
  const std::vector<double>& Z = use_new ? z_new : z_int;
  A.set(0);
  for (int i = 0; i < Np; i++) {
    int ii = (int)(round(X[i] * idx));
    int jj = (int)(round(Y[i] * idy));
    int kk = (int)(round(Z[i] * idz));
    double x_l = X[i] * idx - ii + 0.5;
    double y_l = Y[i] * idy - jj + 0.5;
    double z_l = Z[i] * idz - kk + 0.5;
    ii += g->bs - g->il;
    jj += g->bs - g->jl;
    kk += g->bs - g->kl;
    A(ii, jj, kk) += x_l * y_l * z_l * RHS_T[i];
    A(ii - 1, jj, kk) += (1 - x_l) * y_l * z_l * RHS_T[i];
    A(ii, jj - 1, kk) += x_l * (1 - y_l) * z_l * RHS_T[i];
    A(ii, jj, kk - 1) += x_l * y_l * (1 - z_l) * RHS_T[i];
    A(ii - 1, jj - 1, kk) += (1 - x_l) * (1 - y_l) * z_l * RHS_T[i];
    A(ii - 1, jj, kk - 1) += (1 - x_l) * y_l * (1 - z_l) * RHS_T[i];
    A(ii, jj - 1, kk - 1) += x_l * (1 - y_l) * (1 - z_l) * RHS_T[i];
    A(ii - 1, jj - 1, kk - 1) += (1 - x_l) * (1 - y_l) * (1 - z_l) * RHS_T[i];
  }
  A.Update_Ghosts_CUM();
  A.Update_Ghosts();
}

// particle.cpp:58-104 (MODE_L*), 150-250 (SortValid*), 252-311 (ToBuf,
// FromBuf, SWAP) and 745-812 (Send_Recv_*) for one rank: leavers are packed
// while scanning upward with swap-with-last removal; the '<lo' list
// (received as Rbuf0) then the '>hi' list (Rbuf1) are appended.
void Particles::SortValid(int axis, bool keep) {
  std::vector<double>* pos_new = axis == 0 ? &x_new : axis == 1 ? &y_new : &z_new;
  double lo = axis == 0 ? g->xl : axis == 1 ? g->yl : g->zl;
  double hi = axis == 0 ? g->xh : axis == 1 ? g->yh : g->zh;
  double L = axis == 0 ? g->Lx : axis == 1 ? g->Ly : g->Lz;
  std::vector<double>& P0 = *stage(axis, 0);
  std::vector<double>& Pn = *stage(axis, 2);
  std::vector<double>& P1 = *stage(axis, 3);
  Sbuf0.clear();
  Sbuf1.clear();
  auto mode = [&](int i) {
    if (Pn[i] > L) {
      Pn[i] -= L;
      P0[i] -= L;
      P1[i] -= L;
    } else if (Pn[i] < 0) {
      Pn[i] += L;
      P0[i] += L;
      P1[i] += L;
    }
  };
  auto tobuf = [&](int i, std::vector<double>& B) {
    for (int q = 0; q < 7; q++)
      for (int s = 0; s < 4; s++) B.push_back((*stage(q, s))[i]);
    B.push_back((double)flag[i]);
  };
  auto swap = [&](int i, int j) {
    for (int q = 0; q < 7; q++)
      for (int s = 0; s < 4; s++) std::swap((*stage(q, s))[i], (*stage(q, s))[j]);
    std::swap(flag[i], flag[j]);
  };
  int i = 0;
  while (i < Np) {
    if ((*pos_new)[i] > hi) {
      mode(i);
      tobuf(i, Sbuf0);
      swap(i, Np - 1);
      Np--;
      i--;
    } else if ((*pos_new)[i] < lo) {
      mode(i);
      tobuf(i, Sbuf1);
      swap(i, Np - 1);
      Np--;
      i--;
    }
    i++;
  }
  RESIZE();
  if (!keep) return;  // duct x: nothing is received back
  // one rank: Rbuf0 <- Sbuf1 (from the FRONT/TOP/RIGHT neighbour's '<lo'
  // list), Rbuf1 <- Sbuf0; FromBuf appends Rbuf0 then Rbuf1.
  const std::vector<double>* R[2] = {&Sbuf1, &Sbuf0};
  for (int r = 0; r < 2; r++) {
    const std::vector<double>& B = *R[r];
    for (size_t o = 0; o + 29 <= B.size(); o += 29) {
      size_t c = o;
      for (int q = 0; q < 7; q++)
        for (int s = 0; s < 4; s++) stage(q, s)->push_back(B[c++]);
      flag.push_back((int)B[c++]);
      Np++;
    }
  }
  RESIZE();
}

void Particles::Send_Recv() {  // particle.cpp:853-858: Z, then Y, then X
  SortValid(2);
  SortValid(1);
  SortValid(0);
}

// particle.cpp:860-865 with Send_Recv_X(particle& pbox), 814-851, on a rank
// that is both inlet and outlet: Rnum0 = 0, Rbuf1 = the box's Rbuf1.
void Particles::Send_Recv_duct(std::vector<double>& inbuf) {
  SortValid(2);
  SortValid(1);
  SortValid(0, false);
  for (size_t o = 0; o + 29 <= inbuf.size(); o += 29) {
    size_t c = o;
    for (int q = 0; q < 7; q++)
      for (int s = 0; s < 4; s++) stage(q, s)->push_back(inbuf[c++]);
    flag.push_back((int)inbuf[c++]);
    Np++;
  }
  RESIZE();
  inbuf.clear();
}

// particle.cpp:919-958 on one rank (flag = running index).
void Particles::load_random() {
  GlibcRand rng;
  rng.seed(1);
  const double RM = 2147483647.0;  // RAND_MAX
  Np = par->N0;
  for (int q = 0; q < 7; q++)
    for (int s = 0; s < 4; s++) stage(q, s)->clear();
  flag.clear();
  for (int i = 0; i < Np; i++) {
    double r = ((double)rng.next() / RM) * (g->xh - g->xl) + g->xl;
    x.push_back(r);
    u.push_back(par->U0);
    r = ((double)rng.next() / RM) * (g->yh - g->yl) + g->yl;
    y.push_back(r);
    v.push_back(0);
    r = ((double)rng.next() / RM) * (g->zh - g->zl) + g->zl;
    z.push_back(r);
    w.push_back(0);
    T.push_back(par->T0);
    flag.push_back(i);
  }
  RESIZE();
}

}  // namespace orc
