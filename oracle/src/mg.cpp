// Oracle of the preconditioned-CG pressure solve with an FP32 geometric
// multigrid V-cycle as preconditioner.  TEST INFRASTRUCTURE.
//
// Counterpart of the reference's Hypre branch poisson::Solve with
// Solver = 2 (PCG, poisson.cpp:115-172, 219-248): same operator (the 7-point
// div(C grad) with C = 1/Rho_face, values C*inv_d2), two-norm relative
// residual test at `epsilon`, at most Iteration iterations.  BoomerAMG is
// replaced by the geometric V-cycle below, and the singular row-0 pin of
// poisson.cpp:138-145 by working on the zero-mean subspace (the right-hand
// side is zero-mean, every preconditioned residual is made zero-mean).
//
// Arithmetic is fixed so the device solver reproduces it bit for bit: the
// V-cycle runs in float with the expression orders written here, the outer
// iteration in double, and every sum uses the blocked order of dsum() (1024
// chunks, 256 strided partial sums per chunk, pairwise tree, then the chunk
// sums by the same pairwise tree).
#include <cmath>
#include <cstring>

#include "orc.h"

namespace orc {

namespace {
constexpr int NB = 1024, NT = 256;
}

// Sum of f(t), t = 0..n-1, in the device order described above.
template <typename F>
static double dsum(long n, F f) {
  const long chunk = (n + NB - 1) / NB;
  double bsum[NB];
  for (int b = 0; b < NB; b++) {
    const long lo = (long)b * chunk;
    double part[NT];
    for (int th = 0; th < NT; th++) {
      double a = 0.0;
      for (long t = lo + th; t < lo + chunk && t < n; t += NT) a += f(t);
      part[th] = a;
    }
    for (int s = NT / 2; s > 0; s >>= 1)
      for (int th = 0; th < s; th++) part[th] += part[th + s];
    bsum[b] = part[0];
  }
  for (int s = NB / 2; s > 0; s >>= 1)
    for (int b = 0; b < s; b++) bsum[b] += bsum[b + s];
  return bsum[0];
}

void MGPCG::build(const Geom* gg) {
  g = gg;
  duct = !g->box;
  L.clear();
  int nx = g->Nx_tot, ny = g->Ny_tot, nz = g->Nz_tot;
  while (true) {
    MGLevel lv;
    lv.nx = nx;
    lv.ny = ny;
    lv.nz = nz;
    const size_t n = (size_t)nx * ny * nz;
    lv.cx.assign(n, 0.f);
    lv.cy.assign(n, 0.f);
    lv.cz.assign(n, 0.f);
    lv.e.assign(n, 0.f);
    lv.r.assign(n, 0.f);
    lv.t.assign(n, 0.f);
    L.push_back(lv);
    if (max_levels && (int)L.size() >= max_levels) break;
    if (nx % 2 || ny % 2 || nz % 2 || nx < 8 || ny < 8 || nz < 8) break;
    nx /= 2;
    ny /= 2;
    nz /= 2;
  }
}

// Level-0 face coefficients C*inv_d2 with C = 1/Rho_face (the Hypre matrix
// values of poisson.cpp:131-136), then each coarse face the mean of the four
// fine faces it covers times 1/4 (spacing doubled): sum/16.
void MGPCG::coefficients(const F1& Rho_face) {
  const int bs = g->bs;
  const double ix2 = 1. / (g->dx * g->dx), iy2 = 1. / (g->dy * g->dy), iz2 = 1. / (g->dz * g->dz);
  MGLevel& f = L[0];
  for (int k = 0; k < f.nz; k++)
    for (int j = 0; j < f.ny; j++)
      for (int i = 0; i < f.nx; i++) {
        const size_t q = f.at(i, j, k);
        f.cx[q] = (duct && i == 0) ? 0.f : (float)((1. / Rho_face.x(i + bs, j + bs, k + bs)) * ix2);
        f.cy[q] = (float)((1. / Rho_face.y(i + bs, j + bs, k + bs)) * iy2);
        f.cz[q] = (float)((1. / Rho_face.z(i + bs, j + bs, k + bs)) * iz2);
      }
  for (size_t l = 1; l < L.size(); l++) {
    const MGLevel& a = L[l - 1];
    MGLevel& c = L[l];
    for (int k = 0; k < c.nz; k++)
      for (int j = 0; j < c.ny; j++)
        for (int i = 0; i < c.nx; i++) {
          const size_t q = c.at(i, j, k);
          const int fi = 2 * i, fj = 2 * j, fk = 2 * k;
          c.cx[q] = (((a.cx[a.at(fi, fj, fk)] + a.cx[a.at(fi, fj + 1, fk)]) +
                      a.cx[a.at(fi, fj, fk + 1)]) + a.cx[a.at(fi, fj + 1, fk + 1)]) * 0.0625f;
          c.cy[q] = (((a.cy[a.at(fi, fj, fk)] + a.cy[a.at(fi + 1, fj, fk)]) +
                      a.cy[a.at(fi, fj, fk + 1)]) + a.cy[a.at(fi + 1, fj, fk + 1)]) * 0.0625f;
          c.cz[q] = (((a.cz[a.at(fi, fj, fk)] + a.cz[a.at(fi + 1, fj, fk)]) +
                      a.cz[a.at(fi, fj + 1, fk)]) + a.cz[a.at(fi + 1, fj + 1, fk)]) * 0.0625f;
        }
  }
}

// Neighbour data of cell (i,j,k) on a level: coefficients of its six faces
// (zero where the duct's x ends close the domain) and neighbour indices
// (periodic wrap).
static inline void nbrs(const MGLevel& v, bool duct, int i, int j, int k, float c[6],
                        size_t q[6]) {
  const int ip = i + 1 == v.nx ? 0 : i + 1, im = i == 0 ? v.nx - 1 : i - 1;
  const int jp = j + 1 == v.ny ? 0 : j + 1, jm = j == 0 ? v.ny - 1 : j - 1;
  const int kp = k + 1 == v.nz ? 0 : k + 1, km = k == 0 ? v.nz - 1 : k - 1;
  q[0] = v.at(im, j, k);
  q[1] = v.at(ip, j, k);
  q[2] = v.at(i, jm, k);
  q[3] = v.at(i, jp, k);
  q[4] = v.at(i, j, km);
  q[5] = v.at(i, j, kp);
  const size_t s = v.at(i, j, k);
  c[0] = v.cx[s];
  c[1] = (duct && i + 1 == v.nx) ? 0.f : v.cx[q[1]];
  c[2] = v.cy[s];
  c[3] = v.cy[q[3]];
  c[4] = v.cz[s];
  c[5] = v.cz[q[5]];
}

// One red-black half sweep of S e = r, S = -div(C grad):
// e_c = (r_c + sum_f C_f e_f) / sum_f C_f, faces in the order -x,+x,-y,+y,-z,+z.
void MGPCG::smooth(MGLevel& v, int color) {
  for (int k = 0; k < v.nz; k++)
    for (int j = 0; j < v.ny; j++)
      for (int i = 0; i < v.nx; i++) {
        if (((i + j + k) & 1) != color) continue;
        float c[6];
        size_t q[6];
        nbrs(v, duct, i, j, k, c, q);
        float num = v.r[v.at(i, j, k)];
        float den = 0.f;
        for (int f = 0; f < 6; f++) {
          num = num + c[f] * v.e[q[f]];
          den = den + c[f];
        }
        v.e[v.at(i, j, k)] = num / den;
      }
}

// t = r - S e
void MGPCG::residual(MGLevel& v) {
  for (int k = 0; k < v.nz; k++)
    for (int j = 0; j < v.ny; j++)
      for (int i = 0; i < v.nx; i++) {
        float c[6];
        size_t q[6];
        nbrs(v, duct, i, j, k, c, q);
        const size_t s = v.at(i, j, k);
        const float ec = v.e[s];
        float se = 0.f;
        for (int f = 0; f < 6; f++) se = se + c[f] * (ec - v.e[q[f]]);
        v.t[s] = v.r[s] - se;
      }
}

void MGPCG::vcycle(size_t l) {
  MGLevel& v = L[l];
  std::fill(v.e.begin(), v.e.end(), 0.f);
  if (l + 1 == L.size()) {
    for (int s = 0; s < coarse_sweeps; s++) {
      smooth(v, 0);
      smooth(v, 1);
    }
    for (int s = 0; s < coarse_sweeps; s++) {
      smooth(v, 1);
      smooth(v, 0);
    }
    return;
  }
  for (int s = 0; s < pre; s++) {
    smooth(v, 0);
    smooth(v, 1);
  }
  residual(v);
  MGLevel& c = L[l + 1];
  for (int k = 0; k < c.nz; k++)
    for (int j = 0; j < c.ny; j++)
      for (int i = 0; i < c.nx; i++) {
        const int fi = 2 * i, fj = 2 * j, fk = 2 * k;
        float a = 0.f;
        for (int dk = 0; dk < 2; dk++)
          for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) a = a + v.t[v.at(fi + di, fj + dj, fk + dk)];
        c.r[c.at(i, j, k)] = a * 0.125f;
      }
  vcycle(l + 1);
  for (int k = 0; k < v.nz; k++)
    for (int j = 0; j < v.ny; j++)
      for (int i = 0; i < v.nx; i++) {
        const size_t s = v.at(i, j, k);
        v.e[s] = v.e[s] + c.e[c.at(i / 2, j / 2, k / 2)];
      }
  for (int s = 0; s < post; s++) {
    smooth(v, 1);
    smooth(v, 0);
  }
}

// -D of the defect-correction residual (poisson.cpp:393-397): the fine
// operator S p = -div(C grad p) with C = 1./Rho_face, in double, at interior
// cell (i,j,k) of the ghosted field p (ghosts filled).
static inline double Sfine(const Geom* g, const F1& rf, const F0& p, int i, int j, int k) {
  const double idx = 1. / g->dx, idy = 1. / g->dy, idz = 1. / g->dz;
  const double Cx0 = 1. / rf.x(i, j, k), Cx1 = 1. / rf.x(i + 1, j, k);
  const double Cy0 = 1. / rf.y(i, j, k), Cy1 = 1. / rf.y(i, j + 1, k);
  const double Cz0 = 1. / rf.z(i, j, k), Cz1 = 1. / rf.z(i, j, k + 1);
  const double pc = p(i, j, k);
  const double D = (((p(i + 1, j, k) - pc) * idx) * Cx1 - ((pc - p(i - 1, j, k)) * idx) * Cx0) * idx +
                   (((p(i, j + 1, k) - pc) * idy) * Cy1 - ((pc - p(i, j - 1, k)) * idy) * Cy0) * idy +
                   (((p(i, j, k + 1) - pc) * idz) * Cz1 - ((pc - p(i, j, k - 1)) * idz) * Cz0) * idz;
  return -D;
}

static void fill_ghosts(F0& f, bool duct) {
  f.Update_Ghosts();
  if (duct)
    for (int k = 0; k < f.Nz; k++)
      for (int j = 0; j < f.Ny; j++) {
        f(0, j, k) = f(1, j, k);
        f(f.Nx - 1, j, k) = f(f.Nx - 2, j, k);
      }
}

int MGPCG::solve(const F1& rf, F0& P, const F0& RHS, int maxit, double tol, double* relres) {
  build(P.g);
  coefficients(rf);
  const int bs = g->bs, nx = g->Nx_tot, ny = g->Ny_tot;
  const long n = (long)nx * ny * g->Nz_tot;
  auto I = [&](long t, int& i, int& j, int& k) {
    i = (int)(t % nx) + bs;
    j = (int)((t / nx) % ny) + bs;
    k = (int)(t / ((long)nx * ny)) + bs;
  };
  F0 r, z, p;
  r.init(g);
  z.init(g);
  p.init(g);
  fill_ghosts(P, duct);
  // r = b' - S x with b' = -RHS
  for (long t = 0; t < n; t++) {
    int i, j, k;
    I(t, i, j, k);
    r(i, j, k) = -RHS(i, j, k) - Sfine(g, rf, P, i, j, k);
  }
  const double bn = std::sqrt(dsum(n, [&](long t) {
    int i, j, k;
    I(t, i, j, k);
    return RHS(i, j, k) * RHS(i, j, k);
  }));
  auto precond = [&]() {
    MGLevel& f = L[0];
    for (long t = 0; t < n; t++) {
      int i, j, k;
      I(t, i, j, k);
      f.r[t] = (float)r(i, j, k);
    }
    vcycle(0);
    const double m = dsum(n, [&](long t) { return (double)f.e[t]; }) / (double)n;
    for (long t = 0; t < n; t++) {
      int i, j, k;
      I(t, i, j, k);
      z(i, j, k) = (double)f.e[t] - m;
    }
  };
  auto dot = [&](const F0& a, const F0& b) {
    return dsum(n, [&](long t) {
      int i, j, k;
      I(t, i, j, k);
      return a(i, j, k) * b(i, j, k);
    });
  };
  double rr = dot(r, r);
  int it = 0;
  if (std::sqrt(rr) > tol * bn) {
    precond();
    for (long t = 0; t < n; t++) {
      int i, j, k;
      I(t, i, j, k);
      p(i, j, k) = z(i, j, k);
    }
    double rz = dot(r, z);
    while (it < maxit) {
      fill_ghosts(p, duct);
      // q = S p, kept in z's storage after p.q is formed
      F0& q = z;
      for (long t = 0; t < n; t++) {
        int i, j, k;
        I(t, i, j, k);
        q(i, j, k) = Sfine(g, rf, p, i, j, k);
      }
      const double pq = dot(p, q);
      const double alpha = rz / pq;
      for (long t = 0; t < n; t++) {
        int i, j, k;
        I(t, i, j, k);
        P(i, j, k) = P(i, j, k) + alpha * p(i, j, k);
        r(i, j, k) = r(i, j, k) - alpha * q(i, j, k);
      }
      it++;
      rr = dot(r, r);
      if (std::sqrt(rr) <= tol * bn) break;
      precond();
      const double rz2 = dot(r, z);
      const double beta = rz2 / rz;
      rz = rz2;
      for (long t = 0; t < n; t++) {
        int i, j, k;
        I(t, i, j, k);
        p(i, j, k) = z(i, j, k) + beta * p(i, j, k);
      }
    }
  }
  fill_ghosts(P, duct);
  if (relres) *relres = bn > 0 ? std::sqrt(rr) / bn : 0.0;
  return it;
}

}  // namespace orc
