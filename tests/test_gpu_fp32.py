"""The precision form (FR_SIM_FP32): FP32 fields and stencils, FP64 pressure,
Poisson residual, sums and particles.  Pinned by its error against the FP64
device path (itself bit-exact with the oracle to 1e-12) per reported result:
TKE, dissipation, Rho_Mean and mean particle temperature within 1e-5
relative, and the Poisson solve still converged to the same tolerance."""
import ctypes as C

import numpy as np
import pytest

import first_repo_amd as fr
from first_repo_amd.configs import DuctWorkload
from gpu_util import make_box_case

pytestmark = pytest.mark.gpu

RESULTS = {0: "TKE", 7: "dissipation", 1: "Rho_Mean", 8: "mean particle T"}


def _run(params, d, fp32, steps):
    sim = fr.Sim(params, d, fp32=fp32)
    sim.initialize()
    diags = []
    for _ in range(steps):
        sim.step()
        diags.append(sim.diagnostics())
    rho = sim.field("Rho")
    sim.close()
    return np.array(diags), rho


def test_fp32_form_error_against_fp64(tmp_path):
    params, d = make_box_case(tmp_path, n=32, n_particles=20000)
    print("fp32 form: box 32^3, 20000 particles, 8 steps", flush=True)
    d64, rho64 = _run(params, d, False, 8)
    d32, rho32 = _run(params, d, True, 8)
    for k, name in RESULTS.items():
        err = np.abs(d32[:, k] - d64[:, k]) / np.abs(d64[:, k])
        print(f"{name}: max relative error {err.max():.3e}")
        assert err.max() <= 1e-5, (name, err)
    # the pressure solve converges as the FP64 one does (iteration count of the
    # last solve of each step)
    print("iterations fp64", d64[:, 9], "fp32", d32[:, 9])
    assert (d32[:, 9] <= d64[:, 9] + 1).all()
    inner = (slice(1, -1),) * 3
    e = np.abs(rho32[inner] - rho64[inner]).max() / np.abs(rho64[inner]).max()
    assert e < 1e-5, e


def test_fp32_form_with_multigrid_pcg(tmp_path):
    """The precision form with Solver = 2: FP64 PCG around FP32 multigrid
    V-cycles (the FP64 operator reads the FP32 faces widened).  Same bars
    against the FP64 path with the same solver, and every solve stops at the
    1e-10 relative residual within the iteration limit."""
    params, d = make_box_case(tmp_path, n=32, n_particles=20000, extra={"Solver": 2})
    print("fp32 form + PCG: box 32^3, 20000 particles, 6 steps", flush=True)
    runs = {}
    for fp32 in (False, True):
        sim = fr.Sim(params, d, fp32=fp32)
        sim.initialize()
        out = []
        for _ in range(6):
            sim.step()
            out.append(sim.diagnostics())
            assert sim.scalar("relres") <= 1e-10
            assert 0 < out[-1][9] < 30
        sim.close()
        runs[fp32] = np.array(out)
    for k, name in RESULTS.items():
        err = np.abs(runs[True][:, k] - runs[False][:, k]) / np.abs(runs[False][:, k])
        print(f"{name}: max relative error {err.max():.3e}")
        assert err.max() <= 1e-5, (name, err)


def test_fp32_form_box_only(tmp_path):
    """The precision form is offered for the box (FFT or PCG solve) only."""
    params, d = make_box_case(tmp_path, n=16, n_particles=100, extra={"Solver": 2})
    h = C.c_void_p()
    assert fr.lib().fr_sim_create_ex(params.encode(), d.encode(), fr.FR_SIM_FP32,
                                     C.byref(h)) == 0
    fr.lib().fr_sim_destroy(h)
    rc = fr.lib().fr_sim_create_ex(params.encode(), d.encode(), 4, C.byref(h))
    assert rc == 1001  # FR_EINVAL: unknown flag
    dd = str(tmp_path / "duct")
    DuctWorkload(n=(16, 8, 8), inject_per_step=0, capacity=10).build(dd)
    rc = fr.lib().fr_sim_create_ex((dd + "/params.in").encode(), dd.encode(), fr.FR_SIM_FP32,
                                   C.byref(h))
    assert rc == 1001
