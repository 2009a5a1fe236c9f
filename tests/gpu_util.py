"""Helpers for the GPU parity tests: device buffers through the HIP runtime
device memory), the oracle/device state sync and comparisons."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

import first_repo_amd as fr
from first_repo_amd.configs import BoxWorkload
import oracle_lib as OL

# live state between steps (the reference's RHS_RU and Rho_new hold dead
# intermediates there, and Update_RU_WOP leaves Rho_face shifted by Rho0 until
# Compute_RHS_Pois rewrites it; the fused kernels do not store those)
FIELDS1 = ["RU", "RU_int", "RU_np1", "Rho_face"]
FIELDS0 = ["P", "RHS_Pois", "C", "Rho", "Rho_int", "Rho_np1", "RHS_Part_Temp"]
SYNC1 = FIELDS1 + ["RU_new", "U", "RHS_RU"]
SYNC0 = FIELDS0 + ["Rho_new"]


_hip = None


def hip():
    """The HIP runtime libfr_amd.so itself links (/opt/rocm).  torch is not
    imported in GPU tests: it bundles a second HIP/HSA runtime, and two in one
    process cannot share the device."""
    global _hip
    if _hip is None:
        fr.lib()  # loads libamdhip64.so.7 as a dependency
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        n = fr.lib().fr_device_count()
        assert n > 0, "GPU test needs a HIP device"
    return _hip


class DBuf:
    """A device copy of a numpy float64 array."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.shape = a.shape
        self.nbytes = a.nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data, self.nbytes, 1) == 0  # H2D

    def get(self):
        out = np.empty(self.shape)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.p, self.nbytes, 2) == 0  # D2H
        return out

    def __del__(self):
        try:
            hip().hipFree(self.p)
        except Exception:
            pass


def dbuf(a):
    return DBuf(a)


def dptr(t):
    return t.p


def host(t):
    return t.get()


def rel(a, b):
    """Normwise relative difference max|a-b| / max|b|."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return 0.0
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    return float(num / den) if den > 0 else float(num)


def make_box_case(tmp, n=16, n_particles=2000, seed=7, **kw):
    w = BoxWorkload(n=n, n_particles=n_particles, steps=5, seed=seed, **kw)
    d = os.path.join(str(tmp), f"box{n}_{n_particles}_{seed}")
    w.build(d)
    return os.path.join(d, "params.in"), d


def sync_oracle_to_device(og: OL.OracleGrid, sim: fr.Sim):
    """Copy the complete oracle grid state into the device solver."""
    for f in SYNC1:
        for c in range(3):
            sim.set_field(f, og.field(f, c), c)
    for f in SYNC0:
        sim.set_field(f, og.field(f))
    for s in ("P0", "P0_int", "P0_new", "P0_np1", "T_cur", "C_BC"):
        sim.set_scalar(s, og.scalar(s)[0])
    for q in range(7):
        for s in range(4):
            sim.set_part(q, s, og.part(q, s))
    n = og.np_
    sim.set_part(7, 0, np.ctypeslib.as_array(OL.orc().orc_grid_part(og.h, 7, 0), shape=(n,)))
    sim.set_flags(og.flag())
    sim.particles_changed()


def compare_state(og, sim, tol=1e-12, fields1=FIELDS1, fields0=FIELDS0, particles=True,
                  stages=(0, 1, 3)):
    """Largest normwise relative difference per quantity; asserts flags equal.
    The _new stage (2) is live only between particle::update_* and
    TimeAdvance_RK4 (after which the device holds the exchanged buffer)."""
    worst = {}
    for f in fields1:
        for c in range(3):
            worst[f"{f}.{c}"] = rel(sim.field(f, c), og.field(f, c))
    for f in fields0:
        worst[f] = rel(sim.field(f), og.field(f))
    if particles:
        assert sim.np_ == og.np_
        assert np.array_equal(sim.flags(), og.flag()), "particle order differs"
        for q in range(7):
            for s in stages:
                worst[f"part{q}.{s}"] = rel(sim.part(q, s), og.part(q, s))
    bad = {k: v for k, v in worst.items() if not v <= tol}
    return worst, bad


def sim_with_deposit(params, d, form=None, heat_sum=None, **kw):
    """fr.Sim whose particle deposits use `form` ("pull": one thread per
    target cell; "tile": LDS tiles; "split": direct sums for targets with at
    most two contributions, merges for the rest; None: chosen by density).  The form is
    read from FR_DEPOSIT when the particle set is allocated.  heat_sum="field":
    the heat deposit's sum always from a reduction of the folded field
    (FR_HEAT_SUM), not from the split deposit's own partials."""
    env = {"FR_DEPOSIT": form, "FR_HEAT_SUM": heat_sum}
    old = {k: os.environ.pop(k, None) for k in env}
    try:
        for k, v in env.items():
            if v:
                os.environ[k] = v
        return fr.Sim(params, d, **kw)
    finally:
        for k, v in old.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
