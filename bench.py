"""Throughput of the full RK4 time step with particles on one MI355X.

Default workload: the triply periodic box at 512^3 cells with 16,000,000
point particles (Stokes numbers uniform on [0.5, 4], mass loading 0.5),
two-way momentum and heat coupling, radiative heating, plain FP64 with the
FFT pressure solve to max|dp| <= 1e-10 (synthetic seeded solenoidal initial
field; first_repo_amd/configs.py).  A "step" is one full RK4 step (4
substeps: density, particles, momentum, divergence, pressure solve,
projection) on device-resident state; statistics and I/O are excluded.

One JSON line (metric Mcell*steps/s, higher is better) carries, beside the
timed plain FP64 run:
  * roofline: the kernel with the most time per step among those timed by
    HIP events bound to their launches (the forward x pass of the pressure
    solve, the momentum and trilinear deposits) or recorded around them (the
    fused momentum update), its algorithmic bytes per launch over its mean
    launch time; every timed kernel is listed under "kernels";
  * precision_form: the same workload in the precision form (FP32 fields and
    stencils, FP64 pressure, residual, sums and particles), its step time and
    its relative error against the FP64 run after the same steps in TKE,
    dissipation, Rho_Mean and mean particle temperature;
  * duct_poisson: the inflow/outflow duct 512x128x128 (density ratio 3, 2e4
    inlet particles per step, loaded to 7.5M particles): Poisson iterations
    and ms per step to the 1e-10 stopping test for the plain FP64 2-D FFT +
    Thomas solve, PCG with the FP32 multigrid preconditioner, and PCG with the
    learned initial iterate and flexible preconditioner (network trained on
    the GPU from the solver's own pairs first);
  * cpu_baseline: the CPU oracle on a bounded sample of the same physics.
--no-fp32, --no-duct and --no-cpu leave those legs out (profiling runs).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("Mcell·steps/s full RK4 step w/ particles; "
          "Poisson iters+ms to 1e-10 vs plain")
# HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE (each scaled
# by what the counters read for a plain copy of known size: factors 2.0 and
# 1.0), measured on the named workloads (profiles/hbm_traffic_*.md); the
# momentum update's is the mean over a step's four stage forms
TRAFFIC = {("k_pass_x_fwd", 512, 16_000_000, "fp32"): 3.775e9,
           ("k_update_ru_wop_z", 512, 16_000_000, "fp32"): 10.991e9,
           ("k_pass_x_fwd", 512, 16_000_000, "fp64"): 5.488e9,
           ("k_update_ru_wop_z", 512, 16_000_000, "fp64"): 22.150e9,
           ("k_dep_vel", 512, 16_000_000, "fp64"): 5.248e9,
           ("k_dep_cic", 512, 16_000_000, "fp64"): 2.682e9}


def box_workload(n, particles, solver="fft"):
    from first_repo_amd.configs import BoxWorkload
    return BoxWorkload(n=n, n_particles=particles, steps=200, mass_loading=0.5,
                       stokes_range=(0.5, 4.0),
                       extra={"Solver": 2} if solver == "pcg" else {})


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def run_box(params, tmp, fp32, steps, warmup, resort=None):
    """Timed steps of the box solver; returns wall time, per-step timings
    summed over the timed steps, and the diagnostics after the last step."""
    import first_repo_amd as fr
    sim = fr.Sim(params, tmp, fp32=fp32)
    if resort:
        sim.set_resort_interval(resort)
    sim.initialize()
    log(("precision form" if fp32 else "FP64") + " initialised")
    for _ in range(warmup):
        sim.step()
    acc = [0.0] * 12
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        tm = sim.timing()
        for i in range(12):
            acc[i] += tm[i]
    wall = time.perf_counter() - t0
    out = {"wall": wall, "timing": acc, "diag": sim.diagnostics(),
           "ghosted": sim.g.Nx * sim.g.Ny * sim.g.Nz}
    sim.close()
    return out


def kernel_table(n, particles, fp32, r, steps, precision="fp64"):
    """Algorithmic bytes per launch and achieved bandwidth of the timed
    kernels.  Bytes count each distinct array element read or written once
    per launch (es = field element size, ghosted = (n+2)^3 cells):
      k_pass_x_fwd: reads P, RHS (FP64) and Rho_new (the face densities of
        the residual are interpolated from it) once per interior cell, writes
        the half spectrum (n/2+1) n^2 complex;
      momentum deposit: writes 3 components and reads a bucket start per
        ghosted cell; per particle its reference index and 6 operands;
      trilinear deposits (concentration and heat alternate): 1 output and a
        bucket start per ghosted cell; per particle its reference index and 3
        weights, plus RHS_T for the heat deposit (mean of the two per launch);
      k_update_ru_wop_z: 21 values per ghosted cell (reads U, RU_int, S, RU,
        RU_np1; writes RU_np1, RU_new) in the middle stages of a step, 18 in
        the first (RU stands in for RU_np1, which is not read), 15 in the last
        (RU_np1 is not written, RU is not read): 18.75 per launch over a
        step's four."""
    es = 4 if fp32 else 8
    g = r["ghosted"]
    t = r["timing"]
    ncell = n ** 3
    nspec = (n // 2 + 1) * n * n
    rows = [
        ("k_pass_x_fwd", "pressure residual + real FFT along x (events bound to the launch)",
         ncell * (20 if fp32 else 24) + nspec * (8 if fp32 else 16), t[6], t[7]),
        ("k_dep_vel", "momentum deposit, particle.cpp:498-524 (events bound to the launch)",
         g * (3 * es + 4) + particles * 52, t[8], t[9]),
        ("k_dep_cic", "concentration and heat deposits, particle.cpp:526-658 (events bound "
         "to the launch)", g * (es + 4) + particles * 32, t[10], t[11]),
        ("k_update_ru_wop_z", "fused momentum update, grid.cpp:387-425 (events around the "
         "launch; runs beside the particle heat branch)", 75 * es * g // 4, t[4], t[5]),
    ]
    out = []
    for name, what, nbytes, ms, nl in rows:
        if nl <= 0:
            continue
        per = ms / nl
        gbs = nbytes / (per * 1e-3) / 1e9
        out.append({"kernel": name, "what": what, "alg_bytes_per_launch": int(nbytes),
                    "avg_launch_ms": round(per, 4), "launches_per_step": round(nl / steps, 2),
                    "ms_per_step": round(ms / steps, 3), "achieved_GBs": round(gbs, 1),
                    "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": TRAFFIC.get((name, n, particles, precision))})
    return out


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


def cpu_threads():
    """The host cores this process may use, at most 16 (the CPU share of one
    GPU on the box), and what nproc reports."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return min(avail, 16), os.cpu_count() or 1


def cpu_baseline(n, particles, steps=3):
    """The CPU oracle (C++ restatement of the reference path, OpenMP over the
    field loops, particle loops serial as in the reference) on a bounded
    sample of the same physics: one warm-up RK4 step, then the median of
    `steps` steps."""
    threads, nproc = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    tmp = tempfile.mkdtemp(prefix="frcpu_")
    box_workload(n, particles).build(tmp)
    og = OL.OracleGrid(os.path.join(tmp, "params.in"), tmp)
    og.initialize()
    og.step()
    walls = []
    for _ in range(steps):
        t = time.perf_counter()
        og.step()
        walls.append(time.perf_counter() - t)
    med = statistics.median(walls)
    return {"value": round(n ** 3 / med / 1e6, 4), "unit": "Mcell·steps/s",
            "cores": threads, "nproc": nproc, "kind": "port",
            "particle_steps_per_s": round(particles / med, 1),
            "sample": f"oracle (C++ restatement, -O3 -ffp-contract=off, {threads} OpenMP "
                      f"threads on the field loops) on the same physics at {n}^3 with "
                      f"{particles} particles (the timed workload's particle density): 1 warm-up "
                      f"RK4 step, median of {steps} steps ({med:.2f} s per step)"}


def run_duct(solver, steps, warmup, fill=7_500_000, learned=0, pairs=400, train_iters=300):
    """The inflow/outflow duct 512x128x128 with synthetic inflow (uniform stream
    with 5% seeded perturbations), 2e4 inlet particles per step (cap 8M),
    initial density ratio 3 along the duct.  The first warm-up step loads the
    duct to its working particle count.  learned = 1: learned initial
    iterate, 2: also the learned flexible preconditioner (PCG only), after
    collecting `pairs` pairs from the solver's own run and training."""
    import ctypes as C
    import numpy as np
    import first_repo_amd as fr
    from first_repo_amd.configs import DuctWorkload
    w = DuctWorkload(extra={"Solver": 2} if solver == "pcg" else {})
    tmp = tempfile.mkdtemp(prefix="frduct_")
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
        nfill = max(0, min(fill, w.capacity - nsteps * w.inject_per_step))
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

    info = {}
    net = None
    if learned:
        nx, ny, nz = w.n
        data, net = C.c_void_p(), C.c_void_p()
        fr.check(L.fr_unet_data_create(nx, ny, nz, pairs, C.byref(data)))
        fr.check(L.fr_unet_create(1, C.byref(net)))
        log(f"duct: collecting {pairs} training pairs from the PCG run")
        sc = new_sim()
        fr.check(L.fr_sim_set_collect(sc.h, data))
        run(sc, (pairs + 3) // 4, 10 ** 9)
        sc.close()
        cnt = C.c_int()
        L.fr_unet_data_count(data, C.byref(cnt))
        log(f"duct: training on {cnt.value} pairs")
        l0, l1 = C.c_double(), C.c_double()
        t0 = time.perf_counter()
        fr.check(L.fr_unet_train(net, data, train_iters, 1, nx // 2, ny, nz,
                                 C.c_float(3e-4), 7, C.byref(l0), C.byref(l1), None))
        info.update({"pairs": cnt.value, "train_iters": train_iters,
                     "train_crop": [nx // 2, ny, nz],
                     "train_s": round(time.perf_counter() - t0, 2),
                     "loss_first": round(l0.value, 6), "loss_last": round(l1.value, 6)})
        L.fr_unet_data_destroy(data)
    sim = new_sim()
    if net is not None:
        fr.check(L.fr_sim_set_learned(sim.h, net, learned))
    wall, iters, t_p, t_part = run(sim, warmup + steps, warmup)
    ncell = w.n[0] * w.n[1] * w.n[2]
    if net is not None:
        kept, fb, gain = C.c_int(), C.c_int(), C.c_double()
        L.fr_sim_learned_stats(sim.h, C.byref(kept), C.byref(fb), C.byref(gain))
        info.update({"guesses_kept": kept.value, "fallbacks": fb.value})
    out = {"poisson_iters_per_substep": round(iters / steps / 4, 3),
           "poisson_ms_per_step": round(t_p / steps, 3),
           "poisson_ms_per_solve": round(t_p / steps / 4, 3),
           "ms_per_step": round(wall / steps * 1e3, 3),
           "Mcell_steps_per_s": round(ncell * steps / wall / 1e6, 3),
           "particle_ms_per_step": round(t_part / steps, 3),
           "particles_end": sim.np_, **info}
    sim.close()
    if net is not None:
        L.fr_unet_destroy(net)
    return out


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
    """One duct run as the main line (--workload duct512)."""
    r = run_duct(args.solver, args.steps, args.warmup, fill=args.duct_fill,
                 learned=args.learned, pairs=args.pairs, train_iters=args.train_iters)
    ncell = 512 * 128 * 128
    line = {
        "metric": METRIC, "value": r["Mcell_steps_per_s"], "unit": "Mcell·steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if not args.learned else
        "f64 (network bf16, fp32 accumulate)", "data": "synthetic",
        "config": {"workload": "duct 512x128x128, plain FP64, " +
                               ("2-D FFT + Thomas pressure solve to max|dp|<=1e-10"
                                if args.solver == "fft" else
                                "PCG + FP32 multigrid pressure solve to ||r||/||b||<=1e-10") +
                               (", learned initial iterate" if args.learned == 1 else
                                ", learned initial iterate + flexible preconditioner"
                                if args.learned == 2 else "") +
                               ", 2e4 inlet particles/step", "cells": ncell, **r},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="box512",
                    choices=["box512", "box256", "box64", "duct512", "unet512"],
                    help="box512 (default): 512^3, 16M particles; box256: 256^3, 4M; box64: "
                         "64^3, 1e5; duct512: the 512x128x128 inflow/outflow duct alone; "
                         "unet512: the learned part's network alone")
    ap.add_argument("--solver", default="fft", choices=["fft", "pcg"],
                    help="fft: the plain FP64 defect-correction solve (Solver=0); pcg: "
                         "PCG with the FP32 multigrid preconditioner to ||r||/||b|| <= 1e-10 "
                         "(Solver=2)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="form of the main line (fp32: the precision form)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the precision-form leg")
    ap.add_argument("--no-duct", action="store_true", help="skip the duct Poisson leg")
    ap.add_argument("--no-learned", action="store_true",
                    help="duct leg without the learned form (no training)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--duct-steps", type=int, default=5)
    ap.add_argument("--learned", type=int, default=0, choices=[0, 1, 2],
                    help="--workload duct512 with --solver pcg: 1 = learned initial iterate, "
                         "2 = also the learned flexible preconditioner")
    ap.add_argument("--duct-fill", type=int, default=7_500_000)
    ap.add_argument("--pairs", type=int, default=400)
    ap.add_argument("--train-iters", type=int, default=300)
    ap.add_argument("--resort-every", type=int, default=0,
                    help="steps between particle storage re-sorts (0: the library default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "duct512":
        return bench_duct(args)
    if args.workload == "unet512":
        return bench_unet(args)

    n, particles = {"box512": (512, 16_000_000), "box256": (256, 4_000_000),
                    "box64": (64, 100_000)}[args.workload]
    fp32 = args.precision == "fp32"
    tmp = tempfile.mkdtemp(prefix="frbench_")
    log(f"building {n}^3 workload with {particles} particles")
    box_workload(n, particles, args.solver).build(tmp)
    params = os.path.join(tmp, "params.in")
    r = run_box(params, tmp, fp32, args.steps, args.warmup, args.resort_every)
    wall, t = r["wall"], r["timing"]
    ncell = n ** 3
    n_it = t[3] / args.steps / 4
    if fp32:  # FP32 fields: the stencil streams halve, P and the particles do not
        step_bytes = ncell * (4 * 208 + 4 * 80 * n_it) + particles * 4 * 264
    else:
        step_bytes = ncell * (4 * 416 + 4 * 96 * n_it) + particles * 4 * 288
    kernels = kernel_table(n, particles, fp32, r, args.steps, args.precision)
    top = max(kernels, key=lambda k: k["ms_per_step"])
    diag = r["diag"]
    form = ("precision form (FP32 fields/stencils, FP64 pressure and particles)" if fp32
            else "plain FP64")
    solve = ("FFT pressure solve to max|dp|<=1e-10" if args.solver == "fft"
             else "PCG + FP32 multigrid pressure solve to ||r||/||b||<=1e-10")
    out = {
        "metric": METRIC,
        "value": round(ncell * args.steps / wall / 1e6, 3),
        "unit": "Mcell·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"box {n}^3, {particles} particles (Stokes numbers uniform on "
                               f"[0.5, 4], mass loading 0.5), {form}, {solve}",
                   "cells": ncell, "particles": particles,
                   "particle_steps_per_s": round(particles * args.steps / wall, 1),
                   "poisson_iters_per_substep": round(n_it, 3),
                   "poisson_ms_per_step": round(t[1] / args.steps, 3),
                   "particle_ms_per_step": round(t[2] / args.steps, 3),
                   # whole-step algorithmic bytes: 416 B/cell/substep of stencils,
                   # 96 B/cell per Poisson iteration, 288 B/particle/substep
                   "step_alg_GB": round(step_bytes / 1e9, 2),
                   "step_hbm_frac": round(step_bytes / (wall / args.steps) / 1e9 /
                                          HBM_PEAK_GBS, 4),
                   "TKE": float(diag[0]), "Rho_Mean": float(diag[1]),
                   "mean_particle_T": float(diag[8])},
        "roofline": {"kernel": top["kernel"], "what": top["what"], "bound": "hbm",
                     "achieved": top["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": top["frac"],
                     "traffic": TRAFFIC.get((top["kernel"], n, particles, args.precision)),
                     "alg_bytes_per_launch": top["alg_bytes_per_launch"],
                     "avg_launch_ms": top["avg_launch_ms"]},
        "kernels": kernels,
    }
    log(f"main line: {out['value']} Mcell*steps/s, {out['ms_per_step']} ms/step")
    if not args.no_fp32 and not fp32:
        f = run_box(params, tmp, True, args.steps, args.warmup, args.resort_every)
        fd = f["diag"]
        out["precision_form"] = {
            "value": round(ncell * args.steps / f["wall"] / 1e6, 3),
            "ms_per_step": round(f["wall"] / args.steps * 1e3, 3),
            "speedup_vs_fp64": round(wall / f["wall"], 4),
            "poisson_iters_per_substep": round(f["timing"][3] / args.steps / 4, 3),
            "rel_error_vs_fp64": {name: float(abs(fd[k] - diag[k]) / abs(diag[k]))
                                  for k, name in ((0, "TKE"), (7, "dissipation"),
                                                  (1, "Rho_Mean"), (8, "mean_particle_T"))},
            "error_steps": args.warmup + args.steps}
        log(f"precision form: {out['precision_form']['ms_per_step']} ms/step")
    sgbs = stream_copy_gbs()
    if sgbs:
        out["roofline"]["stream_copy_GBs"] = round(sgbs, 1)
        out["roofline"]["frac_of_stream"] = round(top["achieved_GBs"] / sgbs, 4)
    if not args.no_duct:
        dp = {}
        for key, solver, learned in (("fft_thomas", "fft", 0), ("pcg_gmg", "pcg", 0),
                                     ("learned_fpcg", "pcg", 2)):
            if learned and args.no_learned:
                continue
            log(f"duct leg: {key}")
            dp[key] = run_duct(solver, args.duct_steps, 2, learned=learned, pairs=args.pairs,
                               train_iters=args.train_iters)
        dp["workload"] = ("duct 512x128x128, density ratio 3, 2e4 inlet particles/step, loaded "
                          "to 7.5M particles; fft_thomas: plain FP64 2-D FFT + Thomas defect "
                          "correction to max|dp|<=1e-10; pcg_gmg: PCG with an FP32 multigrid "
                          "V-cycle to ||r||/||b||<=1e-10; learned_fpcg: the same PCG with the "
                          "learned initial iterate and flexible preconditioner (multigrid "
                          "fallback)")
        out["duct_poisson"] = dp
    if not args.no_cpu:
        # same particle density as the timed workload, at most 256^3 cells
        cn = min(n, 256)
        cp = particles * cn ** 3 // n ** 3
        log(f"CPU oracle sample {cn}^3 / {cp}")
        out["cpu_baseline"] = cpu_baseline(cn, cp)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
