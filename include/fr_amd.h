/*
 * fr_amd.h - C ABI of the MI355X (gfx950) time-step library libfr_amd.so.
 *
 * Plain pointers and sizes only.  Every field pointer is a DEVICE pointer to
 * a ghosted double array laid out like the reference's tensor0
 * (tensor0.cpp:72-80): value(i,j,k) = ptr[(k*Ny + j)*Nx + i], i fastest,
 * Nx = Nx_tot + 2*bs.  A tensor1 is passed as its three components (x, y, z)
 * on the left faces of each cell.  Streams are hipStream_t passed as void*
 * (NULL = default stream).  Every entry point returns 0 on success or a
 * nonzero HIP error code / FR_E* code; nothing is allocated implicitly except
 * where an entry point says it owns a workspace.
 *
 * Each entry point names the reference interface it replaces (file:line).
 * The reference has no C ABI: its plug-in surface is the C++ classes tensor0,
 * tensor1, particle, poisson, poisson_IO and grid; the host library keeps
 * those classes (first_repo_amd/host/) and they call the functions below.
 */
#ifndef FR_AMD_H
#define FR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FR_OK 0
#define FR_EINVAL 1001   /* bad argument (size, pointer, unsupported shape) */
#define FR_ENOMEM 1002   /* device allocation failed */
#define FR_EOVERFLOW 1003 /* particle capacity exceeded */

/* Grid geometry of one rank's ghosted arrays (gridsize.h:9-58). */
typedef struct fr_geom {
  int Nx, Ny, Nz;  /* ghosted sizes, N_tot + 2*bs */
  int bs;          /* ghost width; kernels support bs == 1 (grid.cpp:366) */
  double dx, dy, dz;
  double Lx, Ly, Lz;
  int is_box;      /* 1: triply periodic box, 0: inflow/outflow duct */
} fr_geom;

/* ---------------------------------------------------------------- library */
/* Library version string and device check. */
const char* fr_version(void);
/* Number of visible HIP devices (0 when none: every compute entry point then
 * fails with the HIP error of its first launch). */
int fr_device_count(void);

/* -------------------------------------------------- tensor0/tensor1 stencils
 * Each applies the reference write range and then the reference halo
 * (Update_Ghosts: periodic wrap in the box; y,z wrap only in the duct). */
/* tensor0::Equal_Div_F2C (tensor0.cpp:303-310) */
int fr_t0_div_f2c(const fr_geom* g, double* out, const double* tx,
                  const double* ty, const double* tz, void* stream);
/* tensor0::Equal_Del2 (tensor0.cpp:312-319) */
int fr_t0_del2(const fr_geom* g, double* out, const double* in, void* stream);
/* tensor1::Equal_Grad_C2F (tensor1.cpp:242-260) */
int fr_t1_grad_c2f(const fr_geom* g, double* ox, double* oy, double* oz,
                   const double* in, void* stream);
/* tensor0::Equal_I{x,y,z}_{C2F,F2C} (tensor0.cpp:353-406):
 * axis 0/1/2, dir 0 = C2F, 1 = F2C */
int fr_t0_interp(const fr_geom* g, double* out, const double* in, int axis,
                 int dir, void* stream);
/* tensor0::Equal_I_F2C (tensor0.cpp:408-416) */
int fr_t0_i_f2c(const fr_geom* g, double* out, const double* tx,
                const double* ty, const double* tz, void* stream);
/* tensor0::Update_Ghosts / Update_Ghosts_CUM (tensor0.cpp:83-91 ->
 * communicator.cpp:177-396 / 422-626), single rank */
int fr_t0_update_ghosts(const fr_geom* g, double* f, int cumulative,
                        void* stream);
/* Element-wise arithmetic over the whole ghosted array (tensor0.cpp:93-241):
 * out = a*A + b*B + c*C with NULL fields treated as absent;
 * op selects the form:
 *   0: out = a*A + b*B + c*C      (Equal_LinComb, Equal_Mult(double,..))
 *   1: out += a*A                 (PlusEqual_Mult(double, tensor0))
 *   2: out = A*B + b*C            (Equal_LinComb(tensor0,tensor0,double,..))
 *   3: out = a / A                (Equal_Divide(double, tensor0))
 *   4: out = A / B                (Equal_Divide(tensor0, tensor0))
 *   5: out = a                    (operator=(double))
 *   6: out += a                   (operator+=(double))
 *   7: out *= a                   (operator*=(double))
 *   8: out = A                    (operator=(tensor0))
 *   9: out *= A                   (operator*=(tensor0))
 *  10: out -= A                   (operator-=(tensor0))
 *  11: out += A                   (operator+=(tensor0)) */
int fr_t0_elementwise(const fr_geom* g, int op, double* out, double a,
                      const double* A, double b, const double* B, double c,
                      const double* Cf, void* stream);
/* Reductions over the interior (tensor0.cpp:418-516):
 * op 0 sum, 1 mean, 2 max, 3 min, 4 max_abs, 5 sum_squares.
 * The result is written to *result (host memory); synchronises the stream.
 * Sums are deterministic tree sums (order differs from the reference's serial
 * i->j->k loop at the level of rounding). */
int fr_t0_reduce(const fr_geom* g, const double* f, int op, double* result,
                 void* stream);
/* fr_t0_reduce without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as fr_t0_reduce, but the partial results live in a fixed 8.3 KB
 * device array of the library (1024 partials and the result, the same for
 * every field size) instead of a workspace allocated at first use, and the
 * result is copied straight to *result.  Fields up to the limit every entry
 * point has (fewer than 2^31 ghosted cells) are accepted; other arguments
 * fail with FR_EINVAL and a message on stderr.  Calls on one GPU share that
 * array: calls that overlap on one GPU race. */
int fr_t0_reduce_noalloc(const fr_geom* g, const double* f, int op, double* result,
                         void* stream);
/* communicator::REDUCEX (communicator.cpp:635-646): out[i] = sum over
 * interior j,k of f(i,j,k) / (Ny_tot*Nz_tot), for every i; out is a device
 * array of Nx doubles. */
int fr_t0_reducex(const fr_geom* g, const double* f, double* out,
                  void* stream);

/* ------------------------------------------------------------ pressure
 * Box pressure solve: the defect-correction loop of poisson::FFT
 * (poisson.cpp:391-411) around the constant-coefficient 3-D FFT solve
 * poisson::CCP_FFT (poisson.cpp:413-462), on one device.
 * Solves div(C grad P) = RHS with P as the initial guess (poisson.h:79); C
 * is the face coefficient (1/rho_face in grid.cpp:641).  Stops after
 * `iterations` or when max|dp| <= epsilon; *num_iter receives the iteration
 * count (poisson.h:80).  Interior sizes must be powers of two in [2, 1024],
 * with Nx_tot >= 4 (the x transform is real-to-complex).
 * P's ghosts must hold its periodic images on entry; they do on exit.
 * The solver object holds its FFT workspace. */
typedef struct fr_box_poisson fr_box_poisson;
int fr_box_poisson_create(const fr_geom* g, fr_box_poisson** out);
int fr_box_poisson_destroy(fr_box_poisson* s);
int fr_box_poisson_solve(fr_box_poisson* s, const double* cx,
                         const double* cy, const double* cz, double* P,
                         const double* RHS, int iterations, double epsilon,
                         int* num_iter, void* stream);

/* Duct pressure solve: the defect-correction loop of poisson_IO::Solve
 * (poisson_IO.cpp:81-114) around poisson_IO::CCP_FFT (poisson_IO.cpp:116-206):
 * 2-D FFT over the periodic y,z directions, one Thomas solve along x per
 * (y,z) mode with Neumann ends and the (0,0) mode pinned
 * (poisson_IO.cpp:208-452), inverse FFT, P += Re/(Ny_tot*Nz_tot).  g must be
 * a duct (is_box = 0).  Same arguments and stopping rule as the box solve,
 * except that the error is max|dp| of the iteration (taken before P += dp,
 * poisson_IO.cpp:100-101).  On exit P's y,z ghosts hold periodic images and
 * its x ghosts Neumann copies (poisson_IO.cpp:102-107).  Ny_tot and Nz_tot
 * must be powers of two (Ny_tot >= 4, both <= 1024), Nx_tot >= 2. */
typedef struct fr_duct_poisson fr_duct_poisson;
int fr_duct_poisson_create(const fr_geom* g, fr_duct_poisson** out);
int fr_duct_poisson_destroy(fr_duct_poisson* s);
int fr_duct_poisson_solve(fr_duct_poisson* s, const double* cx,
                          const double* cy, const double* cz, double* P,
                          const double* RHS, int iterations, double epsilon,
                          int* num_iter, void* stream);

/* Variable-coefficient pressure solve by preconditioned conjugate gradients,
 * the counterpart of the reference's Hypre branch of poisson::Solve with
 * Solver = 2 (poisson.cpp:115-172, 219-248): the 7-point operator
 * div(C grad P) with C = 1/Rho_face, stopping when ||r||_2 <= epsilon *
 * ||RHS||_2 or after `iterations` iterations.  The preconditioner (one
 * geometric-multigrid V-cycle, red-black Gauss-Seidel, in FP32) replaces
 * BoomerAMG; the outer iteration and every residual are FP64.  Takes the face
 * densities rfx, rfy, rfz (Rho_face), works on box and duct geometries (duct:
 * closed x ends, Neumann x ghosts of P on exit), needs even interior sizes.
 * RHS must have zero mean over the interior; P is the initial guess.
 * *relres receives ||r||/||RHS|| at exit. */
typedef struct fr_pcg fr_pcg;
int fr_pcg_create(const fr_geom* g, fr_pcg** out);
int fr_pcg_destroy(fr_pcg* s);
int fr_pcg_solve(fr_pcg* s, const double* rfx, const double* rfy,
                 const double* rfz, double* P, const double* RHS,
                 int iterations, double epsilon, int* num_iter, double* relres,
                 void* stream);

/* ------------------------------------------------------------ learned preconditioner
 * No counterpart in the reference, which uses Hypre only for this solve
 * (poisson.cpp:115-388): a 3-level 3-D U-Net (16/32/64 channels, 3x3x3
 * convolutions, 2x2x2 average pooling, 2x2x2 stride-2 up-convolutions, a 1x1
 * output) with BF16 weights and activations and FP32 accumulation on the
 * matrix cores.  Input channels: the correction equation's right-hand side
 * (RHS - S P at entry), 1/rho, the previous P; output: the pressure
 * increment.  Interior sizes must be multiples of 4.
 *
 * fr_unet_forward/backward take device arrays over an nx x ny x nz interior
 * grid, x-fastest: x is [n][3] FP32 inputs, y and gy are [n] FP32.
 * fr_unet_backward follows fr_unet_forward at the same size and adds the
 * parameter gradients to the ones held (fr_unet_zero_grads clears them).
 * Parameters are one FP32 vector (fr_unet_num_params values): per 3x3x3
 * convolution (c1a 16->16 [input channels 3..15 unused], c1b 16->16,
 * c2a 16->32, c2b 32->32, c3a 32->64, c3b 64->64) W[27][ci][co] then b[co],
 * tap = (dz+1)*9 + (dy+1)*3 + (dx+1); then up2 W[8][64][32] b[32]; d2a
 * 64->32; d2b 32->32; up1 W[8][32][16] b[16]; d1a 32->16; d1b 16->16; the
 * output W[16] b[1].  (Up-convolution offset o = 4 oz + 2 oy + ox.) */
typedef struct fr_unet fr_unet;
typedef struct fr_unet_data fr_unet_data;
int fr_unet_create(int seed, fr_unet** out);
int fr_unet_destroy(fr_unet* u);
int fr_unet_num_params(fr_unet* u, long long* n);
int fr_unet_get_params(fr_unet* u, float* host);
int fr_unet_set_params(fr_unet* u, const float* host);
int fr_unet_get_grads(fr_unet* u, float* host);
int fr_unet_zero_grads(fr_unet* u, void* stream);
int fr_unet_forward(fr_unet* u, int nx, int ny, int nz, const float* x, float* y,
                    void* stream);
int fr_unet_backward(fr_unet* u, int nx, int ny, int nz, const float* gy, void* stream);
int fr_unet_adam_step(fr_unet* u, float lr, void* stream);
/* Training pairs produced by the solver (fr_pcg_set_collect / fr_sim_set_collect):
 * up to `capacity` pairs of an nx x ny x nz interior, stored in BF16. */
int fr_unet_data_create(int nx, int ny, int nz, int capacity, fr_unet_data** out);
int fr_unet_data_destroy(fr_unet_data* d);
int fr_unet_data_count(fr_unet_data* d, int* n);
/* `iters` Adam steps (mean-squared error, batches of random cx x cy x cz
 * crops of stored pairs, the previous-P channel dropped from half of them,
 * cosine-decayed learning rate); the mean batch losses of the first and the
 * last 20 steps are returned. */
int fr_unet_train(fr_unet* u, fr_unet_data* d, int iters, int batch, int cx, int cy, int cz,
                  float lr, unsigned seed, double* loss_first, double* loss_last,
                  void* stream);
/* PCG with the network: mode 0 plain; 1 learned initial iterate; 2 also the
 * flexible preconditioner z = N(r) + V(r - S N(r)), falling back to the plain
 * V-cycle when the network stops paying.  Same stopping test. */
int fr_pcg_set_learned(fr_pcg* s, fr_unet* net, int mode);
int fr_pcg_set_collect(fr_pcg* s, fr_unet_data* d);
/* Counters since creation: learned initial iterates applied, fallbacks to
 * the plain V-cycle, and the summed log10 of the squared-residual reduction
 * the applied initial iterates achieved. */
int fr_pcg_learned_stats(fr_pcg* s, int* guesses_kept, int* fallbacks, double* log10_gain);

/* ------------------------------------------------------------ whole solver
 * The box solver (grid.h:16-115 with its particle set, box poisson solver
 * and the simulation.cpp:34-71 step loop) held on one device.  Created from a
 * params.in file (params.cpp:6-64); Initial=0 reads U.bin/V.bin/W.bin from
 * ic_dir (grid.cpp:184-196).  Host buffers passed to the get/set calls use
 * the ghosted layout above; particle arrays are Np doubles (int32 for flags). */
typedef struct fr_sim fr_sim;
int fr_sim_create(const char* params_path, const char* ic_dir, fr_sim** out);
/* fr_sim_create with options.  FR_SIM_FP32 selects the precision form: every
 * field but P held and advanced in FP32 (stencils, Rho_face, the Poisson
 * right-hand side, the particle deposits), while P, the Poisson residual and
 * every sum stay FP64; particles stay FP64.  The pressure solve is FP64
 * defect correction around an FP32 inner FFT solve (Solver = 0) or FP64 PCG
 * around FP32 multigrid V-cycles (Solver = 2).  Box geometry only, else
 * FR_EINVAL.  No counterpart in the reference, which runs FP64 throughout
 * (grid.h:16-115). */
#define FR_SIM_FP32 1
/* FR_SIM_OUTPUT: the reference's run files, relative to the working
 * directory as grid.cpp writes them: touch.check ("0") and, when params.in
 * sets Statistics, the 15 stat .dat files (opened for append) at creation
 * (grid.cpp:25-49); info.txt at fr_sim_initialize (grid.cpp:229, 869-973). */
#define FR_SIM_OUTPUT 2
int fr_sim_create_ex(const char* params_path, const char* ic_dir, int flags, fr_sim** out);
int fr_sim_destroy(fr_sim* s);
/* grid::Initialize (grid.cpp:110-231) */
int fr_sim_initialize(fr_sim* s);
/* One full time step: 4 RK4 substeps, Correct_mean_RU, TimeAdvance
 * (simulation.cpp:36-70, box calls). */
int fr_sim_step(fr_sim* s);
/* One named grid method (the names of grid.h:93-114, plus "update" for
 * params::update and "Send_Recv" for particle::Send_Recv) at RK4 stage k. */
int fr_sim_call(fr_sim* s, const char* method, int rk4_count);
/* Field transfer: comp -1 for a tensor0, 0..2 for a tensor1 component.
 * Names: RU RU_int RU_new RU_np1 RHS_RU U Rho_face (tensor1); P RHS_Pois C
 * Rho Rho_int Rho_new Rho_np1 RHS_Part_Temp divergence (tensor0). */
int fr_sim_get_field(fr_sim* s, const char* name, int comp, double* host);
int fr_sim_set_field(fr_sim* s, const char* name, int comp,
                     const double* host);
/* Scalars: P0 P0_int P0_new P0_np1 T_cur C_BC U_avg_in U_avg_out, and
 * relres: the relative residual ||r||/||b|| the last Solver = 2 pressure
 * solve stopped at (setting it has no effect). */
int fr_sim_get_scalar(fr_sim* s, const char* name, double* value);
int fr_sim_set_scalar(fr_sim* s, const char* name, double value);
int fr_sim_geom(fr_sim* s, fr_geom* out);
int fr_sim_np(fr_sim* s, int* np);
/* Particle stage q (0..6 = x,y,z,u,v,w,T), s (0..3 = base,_int,_new,_np1);
 * q == 7 is RHS_T (s ignored). */
int fr_sim_get_particles(fr_sim* s, int q, int st, double* host);
int fr_sim_set_particles(fr_sim* s, int np, int q, int st,
                         const double* host);
int fr_sim_get_flags(fr_sim* s, int* host);
int fr_sim_set_flags(fr_sim* s, int np, const int* host);
/* Duct (is_box = 0) inflow, the box grid's part in grid::Update_Rho(grid&),
 * Update_RU_WOP(grid&) and Update_Particle(grid&) (grid.cpp:351-364,
 * 482-504; particle.cpp:814-851): `planes` holds 8 ghosted planes of
 * Ny*Nz doubles (j fastest) -- g.Rho_new(1,j,k), g.Rho_np1(1,j,k),
 * g.RU_new.x(2,j,k), g.RU_new.y(1,j,k), g.RU_new.z(1,j,k), then the same
 * three RU_np1 planes -- read by the following substeps (NULL keeps the
 * current planes); `inlet` holds n_inlet particles entering at the inlet
 * (the box's Rbuf1), 29 doubles each in particle::ToBuf order
 * (particle.cpp:252-262), appended at the next duct migration. */
int fr_sim_set_inflow(fr_sim* s, const double* planes, int n_inlet,
                      const double* inlet);
/* Particle capacity of the solver (default: the params.in particle count);
 * call before fr_sim_initialize.  Exceeding it fails with FR_EOVERFLOW. */
int fr_sim_reserve_particles(fr_sim* s, int capacity);
/* Solver = 2 only: the learned forms and pair collection of fr_pcg_* on the
 * grid's own pressure solve (no counterpart in the reference). */
int fr_sim_set_learned(fr_sim* s, fr_unet* net, int mode);
int fr_sim_set_collect(fr_sim* s, fr_unet_data* d);
int fr_sim_learned_stats(fr_sim* s, int* guesses_kept, int* fallbacks, double* log10_gain);
/* grid::Statistics (grid.cpp:714-866): when params.in sets Statistics, the 32
 * x-profile files stat_dir/<name>.bin (Nx_tot doubles appended per call),
 * the .dat lines "T_cur value" in the stream's default format and
 * TIME.bin, computed on the device.  Paths relative to the working
 * directory. */
int fr_sim_statistics(fr_sim* s);
/* grid::Store (grid.cpp:233-313) with particle::Store_All
 * (particle.cpp:867-917): XYZ.bin at step 1; RU_<n>.bin (+ RU_S_<n>.bin in
 * the duct) and numbers_<n>.dat every Data_freq_fast steps; Part_*_<n>.bin
 * (base stage, reference order, created exclusively: FR_EINVAL if one
 * exists), C_<n>.bin, Rho_<n>.bin (+ Rho_S_<n>.bin) every Data_freq_slow
 * steps; both kinds after a touch stop.  Restart files with the prefix
 * "Restart_" are read by fr_sim_initialize when params.in has Initial = 1
 * (grid.cpp:124-182, particle.cpp:961-1025). */
int fr_sim_store(fr_sim* s);
/* grid::Touch (grid.cpp:77-98): reads touch.check; *stop = 1 (after a store)
 * when it holds 1. */
int fr_sim_touch(fr_sim* s, int* stop);
/* grid::Write_info (grid.cpp:869-973): info.txt for one rank. */
int fr_sim_write_info(fr_sim* s);
/* Steps taken since the initial condition (grid::num_timestep). */
int fr_sim_num_timestep(fr_sim* s, int* n);
/* Steps between re-sorts of the particle storage into cell order (default 4,
 * or FR_RESORT_EVERY at creation).  A tuning knob with no counterpart in the
 * reference: storage order never changes a result, which stays in the
 * reference's particle order. */
int fr_sim_set_resort_interval(fr_sim* s, int steps);
/* Rebuild the particle cell lists after particle positions were set. */
int fr_sim_particles_changed(fr_sim* s);
/* Step diagnostics (grid.cpp:714-811 subset): out[0..9] = TKE, Rho_Mean,
 * Rho_Max, Rho_Min, C_Max, C_Min, C_Mean, mean dissipation, mean particle
 * temperature, Poisson iterations of the last solve. */
int fr_sim_diagnostics(fr_sim* s, double* out);
/* Timing of the last fr_sim_step: out[0] whole step (ms, host clock around
 * the synchronised step), [1] Poisson solves (ms, HIP events), [2] particle
 * phase (ms, HIP events), [3] Poisson iterations in the step, [4] summed
 * duration of the fused momentum kernel (ms, HIP events on its stream),
 * [5] its launch count. */
int fr_sim_timing(fr_sim* s, double* out);
/* fr_sim_timing with out[6] the summed duration (ms, HIP events on the
 * solver's stream) of the box FFT solve's forward x pass (residual + real
 * FFT, k_pass_x_fwd) in the last step and out[7] its launch count, out[8]
 * and out[9] the same for the momentum deposit kernel, out[10] and out[11]
 * for the trilinear (concentration and heat) deposit kernels; writes the
 * first min(n, 12) values. */
int fr_sim_timing_ex(fr_sim* s, double* out, int n);

/* ---- the coupled run (simulation.cpp:19-73): a periodic box (IsBox = 1)
 * feeding a duct (IsBox = 0) of the same Ny, Nz, Ly, Lz on one device, both
 * FP64.  Each substep the duct's inlet takes the box's first interior slice
 * (grid::Update_Rho(grid&), Update_RU_WOP(grid&), grid.cpp:335-537) and the
 * box's '>hi' particle list of its x pass (particle::Send_Recv(particle&),
 * particle.cpp:814-851, 860-865), all in device memory.  While a pair
 * exists its grids step only through fr_pair_step (fr_sim_step and
 * fr_sim_call return FR_EINVAL); destroy the pair before either grid. */
typedef struct fr_pair fr_pair;
int fr_pair_create(fr_sim* box, fr_sim* duct, fr_pair** out);
/* GRID.Initialize(GRID); GRID_IO.Initialize(GRID) (simulation.cpp:29-30):
 * the duct's restart (Initial = 1) takes its inlet ghosts from the box's RU
 * (grid.cpp:155-164).  Reserve the duct's particle capacity first: the box
 * particles entering per substep must fit in it. */
int fr_pair_initialize(fr_pair* p);
/* One time step of both grids in the reference's call order
 * (simulation.cpp:34-62). */
int fr_pair_step(fr_pair* p);
/* grid::Touch(grid& g_IO) (grid.cpp:77-98): one read of touch.check; on a
 * stop (*stop = 1) both grids store. */
int fr_pair_touch(fr_pair* p, int* stop);
int fr_pair_destroy(fr_pair* p);

#ifdef __cplusplus
}
#endif
#endif /* FR_AMD_H */
