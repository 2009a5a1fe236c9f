// Internal interfaces between the kernel files and the host classes of
// libfr_amd.so (not part of the C ABI).
#pragma once
#include "fr_common.hpp"

namespace fr {

// ---- fields (fields.hip) -------------------------------------------------------
// Templates instantiated for double (the FP64 path) and float (the FP32
// precision form); reductions accumulate in double for both.
template <typename T>
int ghost_fill(const Dev& d, T* f, hipStream_t s);
template <typename T>
int cum_fold(const Dev& d, T* f, hipStream_t s);
// the same for nf <= 6 independent fields in one launch per phase
template <typename T>
int ghost_fill_n(const Dev& d, T* const* f, int nf, hipStream_t s);
template <typename T>
int cum_fold_n(const Dev& d, T* const* f, int nf, hipStream_t s);
template <typename T>
int reduce_async(const Dev& d, const T* f, int op, double* dres, hipStream_t s);
template <typename T>
int reduce_sync(const Dev& d, const T* f, int op, double* result, hipStream_t s);
// reduce_async with a caller-owned workspace of RED_PARTS doubles
constexpr int RED_PARTS = 1024;
template <typename T>
int reduce_async_ws(const Dev& d, const T* f, int op, double* dres, double* ws, hipStream_t s);
// f -= mean(f) over n elements with the mean formed on the device (ws:
// RED_PARTS + 1 doubles); the same values as reduce_sync + elementwise add
template <typename T>
int make_mean_zero_async(const Dev& d, T* f, long n, double* ws, hipStream_t s);
// The same followed by f += c when has_c, with the result also written to g
// (one pass: the element values are those of the separate operations)
template <typename T>
int make_mean_zero_set_async(const Dev& d, T* f, T* g, long n, double* ws, bool has_c, double c,
                             hipStream_t s);
// out[0] = the sum of part[0..n) in a fixed order (one workgroup)
// sets > 1: `sets` arrays of n partials one after the other, into out[2g], out[2g+1]
int sum_parts(const double* part, long n, double* out, hipStream_t s, double denom = 0.0,
              int sets = 1);
// f[t] += (T)(-(dsum[0]/interior cells)) for the n elements of f: the shift
// of make_mean_zero_async with the sum already formed
template <typename T>
int add_neg_mean_async(const Dev& d, T* f, long n, const double* dsum, hipStream_t s);
int reducex(const Dev& d, const double* f, double* out, hipStream_t s);
// Device address, on the current device, of the file-scope scratch array of
// the allocation-free reduction (RED_PARTS + 8 doubles); looked up once per
// device.
int noalloc_scratch(double** out);
// The element-wise members of tensor0 (tensor0.cpp:93-241) over n values:
// o = out's value, A B C D fields, a b c scalars.  Products and two-term sums
// in the reference's own operand order; PlusEqual_Mult(T, s) is AR_PEM_ST and
// Equal_Mult(T, s) AR_MUL_ST (products commute).
enum ArithOp {
  AR_SET_S,      // o = a
  AR_ADD_S,      // o += a
  AR_SUB_S,      // o -= a
  AR_MUL_S,      // o *= a
  AR_DIV_S,      // o /= a
  AR_SET_T,      // o = A
  AR_ADD_T,      // o += A
  AR_SUB_T,      // o -= A
  AR_MUL_T,      // o *= A
  AR_DIV_T,      // o /= A
  AR_PEM_TT,     // o += A*B
  AR_PEM_ST,     // o += a*A
  AR_DIV_TT,     // o = A/B
  AR_DIV_TS,     // o = A/a
  AR_DIV_ST,     // o = a/A
  AR_MUL_TT,     // o = A*B
  AR_MUL_ST,     // o = a*A
  AR_LC_STST,    // o = a*A + b*C
  AR_LC_TTST,    // o = A*B + b*C
  AR_LC_TTTT,    // o = A*B + C*D
  AR_LC_STTT,    // o = a*A + C*D
  AR_LC_STSTST,  // o = a*A + b*C + c*D
  AR_NOPS
};
int t0_arith(long n, int op, double* out, double a, const double* A, const double* B, double b,
             const double* C, double c, const double* D, hipStream_t s);
template <typename T>
int elementwise(long n, int op, T* out, double a, const T* A, double b, const T* B, double c,
                const T* C, hipStream_t s);
template <typename T>
int interp_c2f(const fr_geom* g, T* out, const T* in, int axis, hipStream_t s);
int to_double(const float* src, double* dst, long n, hipStream_t s);

// ---- fused box substep kernels (box.hip) ------------------------------------------
struct BoxCoef {
  double cpost_neg;  // -(dt*post)
  double cpre_neg;   // -dt*pre
  double cpost;      // dt*post
  double cpre;       // dt*pre
  int last;          // RK4_count == 3
  // First stage only: Rho_np1 and RU_np1 hold the values of Rho and RU (a
  // full step ends with them equal) without having been written; the
  // kernels read Rho and RU in their place.
  int np1_base = 0;
  // Last stage of a full box step only: Rho_np1 and RU_np1 are not written
  // (the step's end makes them equal to Rho and RU); the momentum update
  // leaves the stage's RU_np1 in RU_new, which the pressure-gradient update
  // finishes in place.
  int lean = 0;
};

struct WopCoef {
  double a_s;      // -mp/Vcell
  double b_g;      // Mu0/3.
  double c_l;      // Mu0
  double A;        // forcing A
  double sx_forc;  // A*Rho0*U0 (x component only)
  double rho0;
  double gx, gy, gz;
  int gravity;
};

struct DivCoef {
  double kc;     // k/Cp
  double bq;     // -R/(Cp*P0_new*Vcell)
  double qmean;  // mean of RHS_Part_Temp (subtracted when cooling)
  int cooling;
  double s;      // 1./(dt*pre)
};

template <typename T>
int box_update_rho(const Dev& d, const BoxCoef& cf, const T* const ru_int[3], const T* rho,
                   T* rho_np1, T* rho_new, const T* const rf[3], T* const u[3], hipStream_t s,
                   T* rho_set = nullptr, const T* rho_int = nullptr);
// With rho_int: Rho_face = I_C2F(Rho_int) holds, and U = RU_int/Rho_face
// forms the face values from Rho_int (rf is not read).
// Rho_face = I_C2F(rn) at every ghosted cell, as box_div_rhs stores it.
template <typename T>
int box_faces_of(const Dev& d, const T* rn, T* const rf[3], hipStream_t s);
template <typename T>
int box_update_ru_wop(const Dev& d, const BoxCoef& cf, const WopCoef& w, const T* const u[3],
                      const T* const ru_int[3], const T* const S[3], const T* const ru[3],
                      T* const ru_np1[3], T* const ru_new[3], const T* const rf[3],
                      hipStream_t s);
template <typename T>
int box_div_rhs(const Dev& d, const DivCoef& q, const T* rho_new, const T* Q,
                const T* const ru_new[3], T* const rf[3], double* rhs, hipStream_t s,
                double* part = nullptr, double* dsum = nullptr);
// rf[0] null: the face values are used but not stored.
// With part (box_div_parts(d) doubles) and dsum: the interior sum of rhs is
// formed in the same pass (per-workgroup sums, then one fixed-order
// reduction into dsum[0]).
long box_div_parts(const Dev& d);
// With cf.lean: RU_new holds the stage's RU_np1 and is updated in place;
// part (box_wp_parts(d) doubles) takes per-workgroup sums of the three
// components, reduced into sums[2m] (the interior sum) and sums[2m+1] (minus
// the interior mean).
template <typename T>
int box_update_ru_wp(const Dev& d, const BoxCoef& cf, const double* P, T* const ru_np1[3],
                     T* const ru_new[3], hipStream_t s, double* part = nullptr,
                     double* sums = nullptr);
long box_wp_parts(const Dev& d);
// ru[m] = (src[m] + (T)sums[2m+1]) (+ c for m = 0): RU = RU_np1 after
// Correct_mean_RU, from the sums of the lean last stage.
template <typename T>
int box_set_mean_free(const Dev& d, T* const ru[3], const T* const src[3], const double* sums,
                      double c, hipStream_t s);

// ---- duct substep kernels (duct.hip) ----------------------------------------------
int duct_update_rho(const Dev& d, const BoxCoef& cf, double cbc, const double* const ru_int[3],
                    const double* rho, const double* rho_int, double* rho_np1, double* rho_new,
                    const double* const rf[3], double* const u[3], const double* in_new,
                    const double* in_np1, hipStream_t s);
int duct_update_ru_wop(const Dev& d, const BoxCoef& cf, const WopCoef& w, double cbc,
                       const double* const u[3], const double* const ru_int[3],
                       const double* const S[3], const double* const ru[3],
                       double* const ru_np1[3], double* const ru_new[3], const double* in_new,
                       const double* in_np1, hipStream_t s);
int duct_div(const Dev& d, const DivCoef& q, const double* rho_new, const double* Q,
             double* const rf[3], double* dvg, hipStream_t s);
int duct_planes(const Dev& d, const double* ru_new_x, const double* rfx, double* out2,
                hipStream_t s);
int duct_scale(const Dev& d, double ratio, int last, double* ru_new_x, double* ru_np1_x,
               hipStream_t s);
int duct_rhs(const Dev& d, double sc, const double* const ru_new[3], const double* const rf[3],
             const double* dvg, double* rhs, hipStream_t s);
int duct_update_ru_wp(const Dev& d, const BoxCoef& cf, const double* P, double* const ru_np1[3],
                      double* const ru_new[3], hipStream_t s);
// the duct's 8 inflow planes ([8][Nz][Ny]) from the box grid b: plane q is
// src[q] at x index xi[q] over the ghosted (j, k) plane
int pair_planes(const Dev& b, const double* const src[8], const int xi[8], double* out,
                hipStream_t s);

// ---- box pressure solve (fft_poisson.hip) -------------------------------------
// fr_box_poisson_solve with the coefficients in one of three forms (cf):
// 0: C itself; 1: Rho_face in place of C, C = 1./Rho_face formed inside the
// residual (grid.cpp:641); 2: the cell-centred density Rho_new in cx (cy, cz
// unused), the face values interpolated from it inside the residual as the
// divergence kernel forms Rho_face (one array read instead of three).
// TC: double, or float for the FP32 precision form (P and RHS stay double).
// rsum (device, 2 values): when RHS still holds its mean, its interior sum
// and minus its mean (sum_parts with a denominator); the residual then uses
// RHS + rsum[1], the value the field takes when the mean is removed
// (make_mean_zero).
template <typename TC>
int box_poisson_solve(fr_box_poisson* s, const TC* cx, const TC* cy, const TC* cz, int cf,
                      double* P, const double* RHS, int iterations, double epsilon, int* num_iter,
                      hipStream_t st, const double* rsum = nullptr);
// Summed duration (ms, HIP events) and count of the forward x-pass launches
// of the box solve since the last reset.
void box_poisson_xtimer(fr_box_poisson* s, double* ms, long* n, bool reset);

// Phase timers of the duct solve (poisson_IO's FFT_TIME and SOLVER_TIME):
// switched on by the first call, which sizes them for `iterations` per solve;
// returns the summed FFT-pass and Thomas-sweep milliseconds so far.
int duct_poisson_timing(fr_duct_poisson* s, int iterations, double* fft_ms, double* solver_ms);
// fr_duct_poisson_solve with the same Rho_face option (duct_poisson.hip).
int duct_poisson_solve(fr_duct_poisson* s, const double* cx, const double* cy, const double* cz,
                       bool inv, double* P, const double* RHS, int iterations, double epsilon,
                       int* num_iter, hipStream_t st);

// PCG + FP32 multigrid (mg_pcg.hip) on Rho_face (C = 1./Rho_face).
int pcg_solve(fr_pcg* s, const double* rfx, const double* rfy, const double* rfz, double* P,
              const double* RHS, int maxit, double tol, int* iters, double* relres,
              hipStream_t st);

// ---- diagnostics (diag.hip) ----------------------------------------------------------
int diag_tke(const Dev& d, const double* u, const double* rho, const double* red,
             double* out, hipStream_t s);
int diag_diss(const Dev& d, const double* const u[3], const double* rho, double mu,
              double* out, hipStream_t s);
int vsum(const double* v, int n, double* part257, double* host, hipStream_t s);

// ---- statistics (stats.hip) ------------------------------------------------------------
// Whole-array operators of grid::Statistics with per-x plane means (ra: three
// profiles of Nx doubles, rr: one): o[m] receives, per ghosted cell,
//   ST_RSD    (r - rr[i])^2                              (Equal_Reduced_sd)
//   ST_TKE2   0.5 r (a_m - ra_m[i]/rr[i])^2               (Equal_TKE2)
//   ST_UIUI2  (a_m - ra_m[i]/rr[i])^2                     (Equal_UiUi2)
//   ST_CORRU  (a_m - ra_m[i]) (r - rr[i])                 (Equal_Cor_Rho_U)
//   ST_TKE    0.5 r (a_m - ra_m[i])^2                     (Equal_TKE)
//   ST_UIUI   (a_m - ra_m[i])^2                           (Equal_UiUi)
//   ST_ST     sqrt(s a_0)                                 (Equal_ST, s = 2 Tp^2)
//   ST_NU     s / r                                       (Equal_nu, s = Mu0)
//   ST_DISS   2 a_0 b_0                                   (Equal_dissipation)
//   ST_ETA    (a_0^3 / b_0)^0.25                          (Equal_eta)
//   ST_DIV0   b_m / r                                     (tensor1 / tensor0)
//   ST_DIV1   a_m / b_m                                   (tensor1 / tensor1)
enum { ST_RSD, ST_TKE2, ST_UIUI2, ST_CORRU, ST_TKE, ST_UIUI, ST_ST, ST_NU, ST_DISS, ST_ETA,
       ST_DIV0, ST_DIV1 };
int stat_point(const Dev& d, int op, double* const o[3], const double* const a[3],
               const double* const b[3], const double* r, const double* ra, const double* rr,
               double s, hipStream_t st);
// Equal_SijSij (rotation false) or Equal_WijWij (true, added to out) at
// interior cells
int stat_strain(const Dev& d, bool rotation, const double* const u[3], double* out,
                hipStream_t st);

// ---- particles (particles.hip) -------------------------------------------------------
struct PartPtrs {
  double* s[7][4];  // x,y,z,u,v,w,T  x  base,_int,_new,_np1
  int* flag;
  double *rhs_u, *rhs_v, *rhs_w, *rhs_T;
  int np;
};

struct AdvCoef {
  double dt_post, dt_pre;      // dt*post, dt*pre
  double iTp;                  // 1/Tp
  // per-particle response times (extension): 1/Tp from the particle's flag
  int per_particle;
  double tp_lo, tp_span;
  double gx, gy, gz;           // ParticleGravity * g
  double dtpost_iS, dtpre_iS;  // dt*post*iS, dt*pre*iS
  double alpha;
  int last;
  // last stage of a full step: the end-of-step copy of stage 3 into stage 0
  // (grid::TimeAdvance, particle.cpp:1027-1040) made by the stage itself
  int set0 = 0;
};

// HIP events bound to the launches of one kernel (hipExtLaunchKernelGGL
// records them at the kernel's own start and end), summed after the stream
// has been synchronised.  Launches beyond NE between two collects are not
// timed.
struct KTimer {
  static constexpr int NE = 32;
  hipEvent_t e0[NE] = {}, e1[NE] = {};
  int used = 0;
  double ms = 0;
  long n = 0;
  int init() {
    for (int i = 0; i < NE; i++) {
      if (hipEventCreate(&e0[i]) != hipSuccess || hipEventCreate(&e1[i]) != hipSuccess)
        return FR_ENOMEM;
    }
    return FR_OK;
  }
  void release() {
    for (int i = 0; i < NE; i++) {
      if (e0[i]) (void)hipEventDestroy(e0[i]);
      if (e1[i]) (void)hipEventDestroy(e1[i]);
      e0[i] = e1[i] = nullptr;
    }
  }
  // the event pair of the next launch, or false when none is left
  bool next(hipEvent_t* a, hipEvent_t* b) {
    if (used >= NE || !e0[used]) return false;
    *a = e0[used];
    *b = e1[used];
    used++;
    return true;
  }
  void collect() {
    for (int i = 0; i < used; i++) {
      float t = 0;
      if (hipEventElapsedTime(&t, e0[i], e1[i]) == hipSuccess) {
        ms += t;
        n++;
      }
    }
    used = 0;
  }
  void reset() {
    ms = 0;
    n = 0;
  }
};

// Cell buckets of the particle set: particles grouped by cell index (rounded
// or truncating), each bucket in ascending particle order, with the deposit
// operands of each particle gathered into bucket order.
struct Buckets {
  int *cnt = nullptr, *start = nullptr, *list = nullptr, *key = nullptr, *slot = nullptr;
  int* list2 = nullptr;  // unsorted bucket lists while a build places them
  // scan scratch of this set's builds: the two sets are built on different
  // streams (the heat branch's R beside the momentum deposit's T), so they
  // share no scratch
  int* scan_tmp = nullptr;
  double* s[6] = {};
  // the same operands as one record per bucket position (split-form
  // deposits), written instead of s[] when rec_form is set at the build
  double* rec = nullptr;
  bool rec_form = false;
  bool valid = false;
};

// Device particle set (SoA, the 29 migrated values of particle.cpp:252-262
// plus the RHS arrays), its cell buckets and migration workspace.  Storage
// slot i holds the particle of reference index rank[i] (inv is the inverse);
// `alt` is a second storage set for re-sorting and host staging.
class PartDev {
 public:
  int np = 0, cap = 0;
  double* s[7][4] = {};
  int* flag = nullptr;
  double *rhs_u = nullptr, *rhs_v = nullptr, *rhs_w = nullptr, *rhs_T = nullptr;
  int *rank = nullptr, *inv = nullptr, *rank_alt = nullptr;
  PartPtrs alt = {};
  long ncell = 0;
  Buckets R;  // rounded index, operands x_l,y_l,z_l,RHS_T
  Buckets T;  // truncating index, operands x_l,y_l,z_l,RHS_u,RHS_v,RHS_w
  // the migration has its own scan scratch, and a copy of the reference
  // indices (rank_dep) lets a bucket build run beside a migration that
  // rewrites them
  int* mig_tmp = nullptr;
  int* rank_dep = nullptr;
  int *flg = nullptr, *nl = nullptr, *pNL = nullptr, *holepos = nullptr, *tailnl = nullptr;
  int *evt = nullptr, *islo = nullptr, *plo = nullptr, *dst = nullptr, *mslot = nullptr,
      *mlv = nullptr;
  int* hbuf = nullptr;
  int* dflag = nullptr;  // device flag of the re-sort's stage check
  // tiles of the trilinear deposits left to the dense kernel (count, ids)
  int *dense_c = nullptr, *dense_w = nullptr;
  double* cic_part = nullptr;  // per-workgroup sums of the weighted split deposit
  long cic_nparts = 0;
  bool heat_sum_field = false;  // FR_HEAT_SUM=field: never use them
  double* inj_buf = nullptr;
  int inj_cap = 0;
  bool flg_zero = false;  // the migration flags flg[0..cap] are all zero
  // deposit kernels: 0 chosen by the mean particles per cell (split form for
  // sparse sets, tiled for dense), 1 the pull form (one thread per target),
  // 2 the tiled LDS form, 3 pull for the momentum deposit and tiled for the
  // trilinear ones, 4 the split form (FR_DEPOSIT=pull|tile|mixed|split)
  int dep_form = 0;
  static constexpr int DEP_TILE = 0, DEP_PULL = 1, DEP_SPLIT = 2;
  int dep_kind(bool trilinear) const;
  // launch timers of the momentum and trilinear deposit kernels
  KTimer t_vel, t_cic;

  int alloc(int capacity, const Dev& d);
  void release();
  PartPtrs ptrs() const;
  template <typename F>
  // eq: stage slots 1 and 3 hold slot 0's values (the top of a full step)
  int advance(const Dev& d, const AdvCoef& c, const F* const u[3], hipStream_t st,
              bool eq = false);
  template <typename F>
  int temp(const Dev& d, int stage, double c0, const F* rho, double iTth, hipStream_t st);
  // buckets of the `stage` positions; mode 0 rounded (R), 1 truncating (T)
  // rk: reference indices to order the buckets by (default: rank)
  int build(const Dev& d, int stage, int mode, hipStream_t st, bool operands = true,
            const int* rk = nullptr);
  // build = build_head (count, scan, unsorted lists: the positions and the
  // reference indices only) + build_tail (placement by reference index and
  // the operands), which may follow on another stream behind an event;
  // nothing may change the positions, ranks or storage order in between
  int build_head(const Dev& d, int stage, int mode, hipStream_t st, const int* rk = nullptr);
  int build_tail(const Dev& d, int stage, int mode, hipStream_t st, bool operands = true,
                 const int* rk = nullptr);
  // trilinear deposit from R: weighted by RHS_T (heat) or not (concentration)
  // sum, summed: when the weighted deposit can form it in the same pass (the
  // split form), sum[0] takes the interior sum of out after the fold and
  // *summed is set; otherwise *summed is cleared and sum is not written
  template <typename F>
  int deposit_cic(const Dev& d, bool weighted, F* out, bool halo, hipStream_t st,
                  double* sum = nullptr, bool* summed = nullptr);
  // momentum deposit from T
  template <typename F>
  int deposit_vel(const Dev& d, F* const out[3], hipStream_t st, bool halo = true);
  // The particle members of the reference taken one at a time (FP64):
  // phase 1 gas2part_velocity (gathered velocity kept per storage slot in
  // gas[0..2]), 2 update_position, 4 update_velocity, 8 update_Temp
  // (particle.cpp:314-399, 480-496); temp_split mode 0 gas2part_Temp_*
  // (Tg from a temperature field into gas[3]), mode 1 Compute_RHS_Temp_*
  // (particle.cpp:402-478).  The kept values follow storage slots: no
  // migration or re-sort may come between a gather and its use.
  double* gas[4] = {};
  int gas_cap = 0;
  int ensure_gas();
  int advance_phase(const Dev& d, const AdvCoef& c, const double* const u[3], int phase,
                    hipStream_t st);
  int temp_split(const Dev& d, int stage, int mode, const double* T, double iTth, hipStream_t st);
  // periodic migration (box); with duct set, x leavers are dropped and the
  // n_in inlet particles of `inj` (29 values each; host memory, or device
  // memory read on st when inj_dev) appended
  int migrate(const double lo[3], const double hi[3], const double L[3], hipStream_t st,
              bool duct = false, const double* inj = nullptr, int n_in = 0,
              bool inj_dev = false);
  // after a periodic migrate on st: the x pass's '>hi' leavers (the
  // reference's Rbuf1) as 29-value rows into buf (at most cap rows), their
  // count into *n_out (device)
  int gather_hi(double* buf, int cap, int* n_out, hipStream_t st);
  void swap_stages(int a, int b);
  void swap_storage();  // exchange the storage set with `alt`
  int copy_stage(int dst_stage, int src_stage, hipStream_t st);
  // storage order: identity, or re-sorted by cell (values unchanged)
  int reset_order(hipStream_t st);
  int resort(const Dev& d, hipStream_t st);
  // host transfers of one array in reference order
  int get(double* host, const double* src, hipStream_t st);
  int set(const double* host, double* dst, hipStream_t st);
  int get_flags(int* host, hipStream_t st);
  int set_flags(const int* host, hipStream_t st);
};

// ---- learned preconditioner (unet.hip) --------------------------------------------
// Network output y (interior order, FP32) for the correction equation
// S dP = sign * r, and the scale that turns y into dP.
int unet_infer(fr_unet* u, const Dev& d, const double* r, double sign, double rrms,
               const double* rfx, const double* rfy, const double* rfz, const double* Pprev,
               const float** y, double* scale, hipStream_t st);
int unet_data_begin(fr_unet_data* D, const Dev& d, const double* r, double sign, double rrms,
                    const double* rfx, const double* rfy, const double* rfz, const double* Pprev,
                    double* scale, hipStream_t st);
bool unet_data_room(const fr_unet_data* D);
int unet_data_end(fr_unet_data* D, const Dev& d, const double* Pstart, const double* Pfinal,
                  double scale, hipStream_t st);

}  // namespace fr
