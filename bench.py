"""Throughput of the full RK4 time step with particles on one MI355X.

Workload (the plain FP64 GPU configuration): triply periodic box 256^3,
4,000,000 point particles, two-way momentum and heat coupling, radiative
heating, FFT pressure solve to max|dp| <= 1e-10 (synthetic seeded solenoidal
initial field; see first_repo_amd/configs.py).  A "step" is one full RK4 step
(4 substeps: density, particles, momentum, divergence, pressure solve,
projection) on device-resident state; statistics and I/O are excluded.

Prints one JSON line (metric Mcell*steps/s, higher is better), with the
roofline of the forward x pass of the pressure solve (the residual fused with
the real FFT along x: the largest kernel of the step that runs with nothing
beside it) measured with HIP events inside the library, and the CPU oracle
timed on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_pass_x_fwd launch from rocprofv3 FETCH_SIZE and WRITE_SIZE
# (each scaled by a plain copy of known size: factors 2.0 and 1.0), measured on
# the named workloads (profiles/hbm_traffic_*.md)
X_PASS_TRAFFIC = {(256, "fp64"): 1.051e9, (512, "fp32"): 6.437e9}
METRIC = ("Mcell\u00b7steps/s full RK4 step w/ particles; "
          "Poisson iters+ms to 1e-10 vs plain")


def cpu_baseline(params, d, threads):
    """Oracle (serial restatement, OpenMP over field loops) on the same
    workload: one RK4 step after initialisation."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    og = OL.OracleGrid(params, d)
    og.initialize()
    t = time.perf_counter()
    og.step()
    wall = time.perf_counter() - t
    return wall


def stream_copy_gbs(nbytes=1 << 30, reps=5):
    """Achievable HBM bandwidth on this device: a plain device-to-device copy
    of 1 GiB, (read + write bytes) / time, best of `reps` (after the timed
    region; the library's own HIP runtime, no other work queued)."""
    import ctypes as C
    import first_repo_amd as fr
    fr.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(a), nbytes) or h.hipMalloc(C.byref(b), nbytes):
        return None
    best = None
    h.hipMemcpy(b, a, nbytes, 3)
    for _ in range(reps):
        assert h.hipDeviceSynchronize() == 0
        t = time.perf_counter()
        assert h.hipMemcpy(b, a, nbytes, 3) == 0
        assert h.hipDeviceSynchronize() == 0
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    h.hipFree(a)
    h.hipFree(b)
    return 2 * nbytes / best / 1e9


def calibrate_copy():
    """A plain device-to-device copy of 256 MiB, three times, for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes (run it under rocprofv3 --pmc)."""
    import ctypes as C
    import first_repo_amd as fr
    fr.lib()  # loads the HIP runtime the library links
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    n = 1 << 28
    a, b = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(a), n) == 0 and h.hipMalloc(C.byref(b), n) == 0
    assert h.hipMemset(a, 1, n) == 0
    for _ in range(3):
        assert h.hipMemcpy(b, a, n, 3) == 0  # device to device
    assert h.hipDeviceSynchronize() == 0
    print(f"bench: copied {n} bytes 3 times", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--particles", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--solver", default="fft", choices=["fft", "pcg"],
                    help="fft: the plain FP64 defect-correction solve (Solver=0); pcg: "
                         "PCG with the FP32 multigrid preconditioner to ||r||/||b|| <= 1e-10 "
                         "(Solver=2)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="fp32: the precision form (FP32 fields and stencils; FP64 pressure, "
                         "Poisson residual, sums and particles); its error against the FP64 "
                         "path over the same steps is measured after the timed region")
    ap.add_argument("--no-error", action="store_true",
                    help="with --precision fp32: skip the FP64 run that measures the error "
                         "(it runs after the timed region)")
    ap.add_argument("--workload", default="box256", choices=["box256", "box512", "duct512", "unet512"],
                    help="box256: the headline 256^3 box; duct512: the 512x128x128 "
                         "inflow/outflow duct (density ratio 3, 2e4 inlet particles/step)")
    ap.add_argument("--learned", type=int, default=0, choices=[0, 1, 2],
                    help="duct512 with --solver pcg: 1 = learned initial iterate, 2 = also the "
                         "learned flexible preconditioner; the network is trained first on "
                         "pairs from the solver's own run (--pairs, --train-iters)")
    ap.add_argument("--duct-fill", type=int, default=7_500_000,
                    help="duct512: particles loaded with the first step's inlet (the duct at its "
                         "working load below the 8M cap; 0 = start empty)")
    ap.add_argument("--pairs", type=int, default=400)
    ap.add_argument("--train-iters", type=int, default=300)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import first_repo_amd as fr
    from first_repo_amd.configs import BoxWorkload

    if args.workload == "duct512":
        return bench_duct(args)
    if args.workload == "unet512":
        return bench_unet(args)
    if args.workload == "box512":
        args.n, args.particles = 512, 16_000_000
    fp32 = args.precision == "fp32"

    tmp = tempfile.mkdtemp(prefix="frbench_")
    w = BoxWorkload(n=args.n, n_particles=args.particles, steps=200, mass_loading=0.5,
                    extra={"Solver": 2} if args.solver == "pcg" else {})
    print(f"bench: building {args.n}^3 workload with {args.particles} particles", file=sys.stderr,
          flush=True)
    w.build(tmp)
    params = os.path.join(tmp, "params.in")
    sim = fr.Sim(params, tmp, fp32=fp32)
    sim.initialize()
    print("bench: initialised", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        sim.step()
    t_poisson = t_part = iters = 0.0
    kern_ms = kern_n = xp_ms = xp_n = 0.0
    t0 = time.perf_counter()
    for s in range(args.steps):
        sim.step()
        tm = sim.timing()
        t_poisson += tm[1]
        t_part += tm[2]
        iters += tm[3]
        kern_ms += tm[4]
        kern_n += tm[5]
        xp_ms += tm[6]
        xp_n += tm[7]
    wall = time.perf_counter() - t0
    ncell = args.n ** 3
    ms_per_step = wall / args.steps * 1e3
    value = ncell * args.steps / wall / 1e6
    diag = sim.diagnostics()
    n_it = iters / args.steps / 4
    step_bytes = ncell * (4 * 416 + 4 * 96 * n_it) + args.particles * 4 * 288
    if fp32:  # FP32 fields: the stencil streams halve, P and the particles do not
        step_bytes = ncell * (4 * 208 + 4 * 80 * n_it) + args.particles * 4 * 264
    # dominant stencil kernel: the fused momentum update (grid::Update_RU_WOP);
    # algorithmic bytes per launch = 21 values per ghosted cell (reads U, RU_int,
    # S, RU, RU_np1: 15; writes RU_np1, RU_new: 6)
    g = sim.g
    ghosted = g.Nx * g.Ny * g.Nz
    alg_bytes = 21 * (4 if fp32 else 8) * ghosted
    per_launch_ms = kern_ms / max(kern_n, 1)
    achieved = alg_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    # roofline kernel: the forward x pass of every pressure iteration
    # (k_pass_x_fwd: residual of the defect correction + real FFT along x).
    # Algorithmic bytes per launch: reads P, RHS (FP64) and the three
    # Rho_face arrays (FP64, FP32 in the precision form) once per interior
    # cell; writes the half spectrum, (n/2+1)*n*n complex (16 B, 8 B FP32).
    nspec = (args.n // 2 + 1) * args.n * args.n
    x_bytes = ncell * (28 if fp32 else 40) + nspec * (8 if fp32 else 16)
    x_ms = xp_ms / max(xp_n, 1)
    x_gbs = x_bytes / (x_ms * 1e-3) / 1e9 if x_ms > 0 else 0.0
    if xp_n == 0:  # PCG solve: no FFT pass; the fused momentum kernel instead
        x_bytes, x_ms, x_gbs = alg_bytes, per_launch_ms, achieved
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell\u00b7steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"box {args.n}^3, {args.particles} particles, " +
                               ("precision form (FP32 fields/stencils, FP64 pressure and "
                                "particles), " if fp32 else "plain FP64, ") +
                               ("FFT pressure solve to max|dp|<=1e-10" if args.solver == "fft"
                                else "PCG + FP32 multigrid pressure solve to ||r||/||b||<=1e-10"),
                   "cells": ncell, "particles": args.particles,
                   "particle_steps_per_s": round(args.particles * args.steps / wall, 1),
                   "poisson_iters_per_substep": round(iters / args.steps / 4, 3),
                   "poisson_ms_per_step": round(t_poisson / args.steps, 3),
                   "particle_ms_per_step": round(t_part / args.steps, 3),
                   # whole-step algorithmic bytes: 416 B/cell/substep of stencils,
                   # 96 B/cell per Poisson iteration, 288 B/particle/substep
                   "step_alg_GB": round(step_bytes / 1e9, 2),
                   "step_hbm_frac": round(step_bytes / (wall / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                   "TKE": float(diag[0]), "Rho_Mean": float(diag[1]),
                   "mean_particle_T": float(diag[8])},
        "roofline": {"kernel": "k_pass_x_fwd (pressure residual + real FFT along x)" if xp_n
                     else "k_update_ru_wop_z (+ periodic ghost shells of its 6 outputs; runs "
                          "beside the particle heat branch)",
                     "bound": "hbm",
                     "achieved": round(x_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(x_gbs / HBM_PEAK_GBS, 4),
                     "traffic": X_PASS_TRAFFIC.get((args.n, args.precision))
                     if xp_n and args.particles == {256: 4_000_000, 512: 16_000_000}.get(args.n)
                     else None,
                     "alg_bytes_per_launch": x_bytes,
                     "avg_launch_ms": round(x_ms, 4), "launches": int(xp_n)},
        # the fused momentum kernel runs beside the particle heat branch on the
        # second stream, so its launch time includes that sharing
        "momentum_kernel": {"kernel": "k_update_ru_wop_z (+ periodic ghost shells of its 6 "
                                      "outputs)",
                            "alg_bytes_per_launch": alg_bytes,
                            "avg_launch_ms": round(per_launch_ms, 4),
                            "achieved_GBs": round(achieved, 1)},
    }
    sim.close()
    sgbs = stream_copy_gbs()
    if sgbs:
        # the same kernel against the bandwidth a plain copy reaches here
        out["roofline"]["stream_copy_GBs"] = round(sgbs, 1)
        out["roofline"]["frac_of_stream"] = round(x_gbs / sgbs, 4)
    if fp32 and not args.no_error:
        # error of the precision form against the FP64 path after the same steps
        print("bench: FP64 reference run for the error of the precision form", file=sys.stderr,
              flush=True)
        ref = fr.Sim(params, tmp)
        ref.initialize()
        for _ in range(args.warmup + args.steps):
            ref.step()
        dref = ref.diagnostics()
        ref.close()
        out["config"]["rel_error_vs_fp64"] = {
            name: float(abs(diag[k] - dref[k]) / abs(dref[k]))
            for k, name in ((0, "TKE"), (7, "dissipation"), (1, "Rho_Mean"),
                            (8, "mean_particle_T"))}
        out["config"]["error_steps"] = args.warmup + args.steps
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        print(f"bench: CPU oracle sample, 1 step, {threads} threads", file=sys.stderr, flush=True)
        cw = cpu_baseline(params, tmp, threads)
        out["cpu_baseline"] = {"value": round(ncell / cw / 1e6, 4), "unit": "Mcell\u00b7steps/s",
                               "cores": threads, "kind": "port",
                               "sample": f"oracle, 1 RK4 step of the same {args.n}^3 / "
                                         f"{args.particles}-particle workload after "
                                         f"initialisation ({cw:.1f} s), OpenMP field loops"}
    print(json.dumps(out), flush=True)


UNET_CONV = [(16, 16, 1), (16, 16, 1), (16, 32, 2), (32, 32, 2), (32, 64, 4), (64, 64, 4),
             (64, 32, 2), (32, 32, 2), (32, 16, 1), (16, 16, 1)]  # (ci, co, level)


def unet_flops(nx, ny, nz):
    """Multiply-adds x 2 of one forward pass of the U-Net (3x3x3 convolutions,
    up-convolutions, 1x1 output; the first convolution's padded input
    channels counted as the 3 real ones)."""
    n1 = nx * ny * nz
    f = 0
    for l, (ci, co, lv) in enumerate(UNET_CONV):
        n = n1 // (lv ** 3)
        f += 2 * 27 * (3 if l == 0 else ci) * co * n
    f += 2 * 64 * 32 * (n1 // 8) + 2 * 32 * 16 * n1  # up-convolutions (fine pixels)
    f += 2 * 16 * n1
    return f


def bench_unet(args):
    """Throughput of the learned preconditioner's network at the duct's full
    512x128x128 grid: forward passes and training steps (forward + backward)
    on the BF16 matrix cores; roofline = achieved dense BF16 TFLOP/s against
    the 2.5 PFLOP/s peak."""
    import ctypes as C
    import numpy as np
    import first_repo_amd as fr
    L = fr.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    nx, ny, nz = 512, 128, 128
    n = nx * ny * nz
    net = C.c_void_p()
    fr.check(L.fr_unet_create(1, C.byref(net)))
    x = np.random.default_rng(0).standard_normal((n, 3)).astype(np.float32)
    gy = (np.random.default_rng(1).standard_normal(n) / n).astype(np.float32)
    dx, dy, dg = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p, nb in ((dx, x.nbytes), (dy, 4 * n), (dg, gy.nbytes)):
        assert h.hipMalloc(C.byref(p), nb) == 0
    assert h.hipMemcpy(dx, x.ctypes.data, x.nbytes, 1) == 0
    assert h.hipMemcpy(dg, gy.ctypes.data, gy.nbytes, 1) == 0
    flops = unet_flops(nx, ny, nz)

    def fwd():
        fr.check(L.fr_unet_forward(net, nx, ny, nz, dx, dy, None))

    def train():
        fwd()
        fr.check(L.fr_unet_zero_grads(net, None))
        fr.check(L.fr_unet_backward(net, nx, ny, nz, dg, None))

    out = {}
    for name, fn, mult in (("forward", fwd, 1), ("train_step", train, 3)):
        for _ in range(args.warmup):
            fn()
        assert h.hipDeviceSynchronize() == 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        assert h.hipDeviceSynchronize() == 0
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        out[name] = (ms, mult * flops / (ms * 1e-3) / 1e12)
    L.fr_unet_destroy(net)
    fms, ftf = out["forward"]
    line = {
        "metric": "U-Net forward TFLOP/s at 512x128x128 (learned preconditioner)",
        "value": round(ftf, 2), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(fms, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "3-level U-Net 16/32/64 channels, 3x3x3 convolutions, "
                               "512x128x128 grid, random weights and inputs",
                   "forward_flop": flops,
                   "train_step_ms": round(out["train_step"][0], 3),
                   "train_step_TFLOPs": round(out["train_step"][1], 2)},
        "roofline": {"kernel": "U-Net forward (k_conv3 on v_mfma_f32_16x16x32_bf16)",
                     "bound": "mfma", "achieved": round(ftf, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(ftf / 2500.0, 4), "traffic": None},
    }
    print(json.dumps(line), flush=True)


def bench_duct(args):
    """The inflow/outflow duct, 512x128x128, synthetic inflow (uniform stream
    with 5% seeded perturbations), 2e4 inlet particles per step (cap 8M),
    initial density ratio 3 along the duct; plain FP64 with the 2-D FFT +
    Thomas defect-correction pressure solve.  The inflow planes and inlet
    particles of each step are prepared before the step's timer starts."""
    import numpy as np
    import first_repo_amd as fr
    from first_repo_amd.configs import DuctWorkload

    import ctypes as C
    if args.duct_fill > 0 and args.warmup < 1:
        print("bench.py: --duct-fill loads the duct in a warm-up step; use --warmup >= 1",
              file=sys.stderr)
        sys.exit(2)
    if args.learned and args.solver != "pcg":
        print("bench.py: --learned needs --solver pcg", file=sys.stderr)
        sys.exit(2)
    w = DuctWorkload(extra={"Solver": 2} if args.solver == "pcg" else {})
    tmp = tempfile.mkdtemp(prefix="frduct_")
    print("bench: building duct 512x128x128", file=sys.stderr, flush=True)
    w.build(tmp)
    params = os.path.join(tmp, "params.in")
    L = fr.lib()

    def new_sim():
        sm = fr.Sim(params, tmp)
        sm.reserve_particles(w.capacity)
        sm.initialize()
        sm.set_field("Rho", w.density())
        sm.call("Finish_Initialize")
        return sm

    def run(sm, nsteps, timed_from):
        # the load stays within the capacity over the run's own inlet particles
        nfill = max(0, min(args.duct_fill, w.capacity - nsteps * w.inject_per_step))
        flag0 = 0
        acc = [0.0, 0.0, 0.0, 0.0]  # wall, iters, poisson ms, particle ms
        for s in range(nsteps):
            inlet = w.inlet_particles(s, flag0)
            if s == 0 and nfill > 0:  # load the duct (untimed: a warm-up step)
                inlet = np.concatenate([w.fill_particles(nfill), inlet])
                inlet[:, 28] = np.arange(len(inlet))
            flag0 += len(inlet)
            sm.set_inflow(w.inflow(s), inlet)
            t0 = time.perf_counter()
            sm.step()
            dt = time.perf_counter() - t0
            if s >= timed_from:
                tm = sm.timing()
                acc[0] += dt
                acc[1] += tm[3]
                acc[2] += tm[1]
                acc[3] += tm[2]
        return acc

    learned = None
    if args.learned:
        # pairs from the solver's own run, then GPU training, then the timed
        # steps with the network; the plain run of the same steps is timed too
        nx, ny, nz = w.n
        data, net = C.c_void_p(), C.c_void_p()
        fr.check(L.fr_unet_data_create(nx, ny, nz, args.pairs, C.byref(data)))
        fr.check(L.fr_unet_create(1, C.byref(net)))
        print(f"bench: collecting {args.pairs} pairs", file=sys.stderr, flush=True)
        sc = new_sim()
        fr.check(L.fr_sim_set_collect(sc.h, data))
        run(sc, (args.pairs + 3) // 4, 10 ** 9)
        sc.close()
        cnt = C.c_int()
        L.fr_unet_data_count(data, C.byref(cnt))
        print(f"bench: training on {cnt.value} pairs", file=sys.stderr, flush=True)
        l0, l1 = C.c_double(), C.c_double()
        t0 = time.perf_counter()
        fr.check(L.fr_unet_train(net, data, args.train_iters, 1, nx // 2, ny, nz,
                                 C.c_float(3e-4), 7, C.byref(l0), C.byref(l1), None))
        t_train = time.perf_counter() - t0
        L.fr_unet_data_destroy(data)
        print("bench: plain PCG run", file=sys.stderr, flush=True)
        sp = new_sim()
        plain = run(sp, args.warmup + args.steps, args.warmup)
        sp.close()
        learned = {"mode": args.learned, "pairs": cnt.value, "train_iters": args.train_iters,
                   "train_crop": [nx // 2, ny, nz], "train_s": round(t_train, 2),
                   "loss_first": l0.value, "loss_last": l1.value,
                   "plain_poisson_iters_per_substep": round(plain[1] / args.steps / 4, 3),
                   "plain_poisson_ms_per_step": round(plain[2] / args.steps, 3),
                   "plain_ms_per_step": round(plain[0] / args.steps * 1e3, 3)}
    print("bench: timed run", file=sys.stderr, flush=True)
    sim = new_sim()
    if args.learned:
        fr.check(L.fr_sim_set_learned(sim.h, net, args.learned))
    acc = run(sim, args.warmup + args.steps, args.warmup)
    wall, iters, t_poisson, t_part = acc
    ncell = w.n[0] * w.n[1] * w.n[2]
    nfill_timed = max(0, min(args.duct_fill,
                             w.capacity - (args.warmup + args.steps) * w.inject_per_step))
    diag = sim.diagnostics()
    out = {
        "metric": METRIC,
        "value": round(ncell * args.steps / wall / 1e6, 3),
        "unit": "Mcell\u00b7steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "duct 512x128x128, plain FP64, " +
                               ("2-D FFT + Thomas pressure solve to max|dp|<=1e-10"
                                if args.solver == "fft" else
                                "PCG + FP32 multigrid pressure solve to ||r||/||b||<=1e-10") +
                               ", 2e4 inlet particles/step" +
                               (f", loaded with {nfill_timed} particles at the first step"
                                if nfill_timed > 0 else ""),
                   "cells": ncell, "particles_end": sim.np_,
                   "particle_steps_per_s": round(sim.np_ * args.steps / wall, 1),
                   "poisson_iters_per_substep": round(iters / args.steps / 4, 3),
                   "poisson_ms_per_step": round(t_poisson / args.steps, 3),
                   "particle_ms_per_step": round(t_part / args.steps, 3),
                   "TKE": float(diag[0]), "Rho_Mean": float(diag[1])},
    }
    if learned:
        out["config"]["workload"] += (", learned initial iterate" if args.learned == 1 else
                                      ", learned initial iterate + flexible preconditioner")
        out["config"]["learned"] = learned
        out["dtype"] = "f64 (network bf16, fp32 accumulate)"
        L.fr_unet_destroy(net)
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
