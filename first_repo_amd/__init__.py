"""MI355X (gfx950) implementation of the ggeraci/first_repo time-step path.

The product is the C-ABI library ``libfr_amd.so`` (include/fr_amd.h): HIP
kernels for the RK4 staggered-grid stencils, the FFT pressure solve and the
Lagrangian particle update, plus host classes mirroring the reference's
tensor0/tensor1/particle/poisson/grid.  This module only loads it with ctypes.
Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfr_amd.so")

_lib = None


class Geom(C.Structure):
    """fr_geom of include/fr_amd.h."""
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("bs", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("is_box", C.c_int)]


def make_geom(nx, ny, nz, L=(2 * np.pi, 2 * np.pi, 2 * np.pi), box=True):
    g = Geom()
    g.Nx, g.Ny, g.Nz, g.bs = nx + 2, ny + 2, nz + 2, 1
    g.Lx, g.Ly, g.Lz = L
    g.dx, g.dy, g.dz = L[0] / nx, L[1] / ny, L[2] / nz
    g.is_box = 1 if box else 0
    return g


class FrError(RuntimeError):
    pass


def lib():
    """The loaded libfr_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FrError(f"{LIB_PATH} not built: run python -m first_repo_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, G = C.c_void_p, C.POINTER(Geom)
        i, d = C.c_int, C.c_double
        sig = {
            "fr_version": ([], C.c_char_p),
            "fr_device_count": ([], i),
            "fr_t0_div_f2c": ([G, vp, vp, vp, vp, vp], i),
            "fr_t0_del2": ([G, vp, vp, vp], i),
            "fr_t1_grad_c2f": ([G, vp, vp, vp, vp, vp], i),
            "fr_t0_interp": ([G, vp, vp, i, i, vp], i),
            "fr_t0_i_f2c": ([G, vp, vp, vp, vp, vp], i),
            "fr_t0_update_ghosts": ([G, vp, i, vp], i),
            "fr_t0_elementwise": ([G, i, vp, d, vp, d, vp, d, vp, vp], i),
            "fr_t0_reduce": ([G, vp, i, C.POINTER(d), vp], i),
            "fr_t0_reduce_noalloc": ([G, vp, i, C.POINTER(d), vp], i),
            "fr_t0_reducex": ([G, vp, vp, vp], i),
            "fr_box_poisson_create": ([G, C.POINTER(vp)], i),
            "fr_box_poisson_destroy": ([vp], i),
            "fr_box_poisson_solve": ([vp, vp, vp, vp, vp, vp, i, d, C.POINTER(i), vp], i),
            "fr_duct_poisson_create": ([G, C.POINTER(vp)], i),
            "fr_duct_poisson_destroy": ([vp], i),
            "fr_duct_poisson_solve": ([vp, vp, vp, vp, vp, vp, i, d, C.POINTER(i), vp], i),
            "fr_pcg_create": ([G, C.POINTER(vp)], i),
            "fr_pcg_destroy": ([vp], i),
            "fr_pcg_solve": ([vp, vp, vp, vp, vp, vp, i, d, C.POINTER(i), C.POINTER(d), vp], i),
            "fr_unet_create": ([i, C.POINTER(vp)], i),
            "fr_unet_destroy": ([vp], i),
            "fr_unet_num_params": ([vp, C.POINTER(C.c_longlong)], i),
            "fr_unet_get_params": ([vp, vp], i),
            "fr_unet_set_params": ([vp, vp], i),
            "fr_unet_get_grads": ([vp, vp], i),
            "fr_unet_zero_grads": ([vp, vp], i),
            "fr_unet_forward": ([vp, i, i, i, vp, vp, vp], i),
            "fr_unet_backward": ([vp, i, i, i, vp, vp], i),
            "fr_unet_adam_step": ([vp, C.c_float, vp], i),
            "fr_unet_data_create": ([i, i, i, i, C.POINTER(vp)], i),
            "fr_unet_data_destroy": ([vp], i),
            "fr_unet_data_count": ([vp, C.POINTER(i)], i),
            "fr_unet_train": ([vp, vp, i, i, i, i, i, C.c_float, C.c_uint, C.POINTER(d),
                               C.POINTER(d), vp], i),
            "fr_pcg_set_learned": ([vp, vp, i], i),
            "fr_pcg_set_collect": ([vp, vp], i),
            "fr_pcg_learned_stats": ([vp, C.POINTER(i), C.POINTER(i), C.POINTER(d)], i),
            "fr_sim_learned_stats": ([vp, C.POINTER(i), C.POINTER(i), C.POINTER(d)], i),
            "fr_sim_set_learned": ([vp, vp, i], i),
            "fr_sim_set_collect": ([vp, vp], i),
            "fr_sim_create": ([C.c_char_p, C.c_char_p, C.POINTER(vp)], i),
            "fr_sim_create_ex": ([C.c_char_p, C.c_char_p, i, C.POINTER(vp)], i),
            "fr_sim_destroy": ([vp], i),
            "fr_sim_initialize": ([vp], i),
            "fr_sim_step": ([vp], i),
            "fr_sim_call": ([vp, C.c_char_p, i], i),
            "fr_sim_get_field": ([vp, C.c_char_p, i, vp], i),
            "fr_sim_set_field": ([vp, C.c_char_p, i, vp], i),
            "fr_sim_get_scalar": ([vp, C.c_char_p, C.POINTER(d)], i),
            "fr_sim_set_scalar": ([vp, C.c_char_p, d], i),
            "fr_sim_geom": ([vp, G], i),
            "fr_sim_np": ([vp, C.POINTER(i)], i),
            "fr_sim_get_particles": ([vp, i, i, vp], i),
            "fr_sim_set_particles": ([vp, i, i, i, vp], i),
            "fr_sim_get_flags": ([vp, vp], i),
            "fr_sim_set_flags": ([vp, i, vp], i),
            "fr_sim_particles_changed": ([vp], i),
            "fr_sim_set_inflow": ([vp, vp, i, vp], i),
            "fr_sim_reserve_particles": ([vp, i], i),
            "fr_sim_set_resort_interval": ([vp, i], i),
            "fr_sim_statistics": ([vp], i),
            "fr_sim_store": ([vp], i),
            "fr_sim_touch": ([vp, C.POINTER(i)], i),
            "fr_sim_write_info": ([vp], i),
            "fr_sim_num_timestep": ([vp, C.POINTER(i)], i),
            "fr_sim_diagnostics": ([vp, vp], i),
            "fr_sim_timing": ([vp, vp], i),
            "fr_sim_timing_ex": ([vp, vp, i], i),
            "fr_pair_create": ([vp, vp, C.POINTER(vp)], i),
            "fr_pair_initialize": ([vp], i),
            "fr_pair_step": ([vp], i),
            "fr_pair_touch": ([vp, C.POINTER(i)], i),
            "fr_pair_destroy": ([vp], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        raise FrError(f"{what} failed with status {rc}")


def _p(a):
    return C.c_void_p(a.ctypes.data)


FR_SIM_FP32 = 1
FR_SIM_OUTPUT = 2


class Sim:
    """The box solver held on the device (fr_sim_* of include/fr_amd.h)."""

    def __init__(self, params_path: str, ic_dir: str | None = None, fp32: bool = False,
                 outputs: bool = False):
        """fp32 selects the precision form (FR_SIM_FP32): FP32 fields and
        stencils, FP64 pressure, residual, sums and particles.  outputs
        (FR_SIM_OUTPUT) writes the reference's run files (touch.check, stat
        files, info.txt) relative to the working directory."""
        self.L = lib()
        h = C.c_void_p()
        flags = (FR_SIM_FP32 if fp32 else 0) | (FR_SIM_OUTPUT if outputs else 0)
        check(self.L.fr_sim_create_ex(params_path.encode(), (ic_dir or ".").encode(), flags,
                                      C.byref(h)),
              "fr_sim_create_ex")
        self.h = h
        self.g = Geom()
        check(self.L.fr_sim_geom(self.h, C.byref(self.g)), "fr_sim_geom")

    @property
    def shape(self):
        return (self.g.Nz, self.g.Ny, self.g.Nx)

    def initialize(self):
        check(self.L.fr_sim_initialize(self.h), "fr_sim_initialize")

    def step(self):
        check(self.L.fr_sim_step(self.h), "fr_sim_step")

    def call(self, name, k=0):
        check(self.L.fr_sim_call(self.h, name.encode(), k), name)

    def field(self, name, comp=-1):
        a = np.empty(self.shape)
        check(self.L.fr_sim_get_field(self.h, name.encode(), comp, _p(a)), "get " + name)
        return a

    def set_field(self, name, a, comp=-1):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        check(self.L.fr_sim_set_field(self.h, name.encode(), comp, _p(a)), "set " + name)

    def scalar(self, name):
        v = C.c_double()
        check(self.L.fr_sim_get_scalar(self.h, name.encode(), C.byref(v)), name)
        return v.value

    def set_scalar(self, name, v):
        check(self.L.fr_sim_set_scalar(self.h, name.encode(), float(v)), name)

    @property
    def np_(self):
        n = C.c_int()
        check(self.L.fr_sim_np(self.h, C.byref(n)))
        return n.value

    def part(self, q, s=0):
        a = np.empty(self.np_)
        check(self.L.fr_sim_get_particles(self.h, q, s, _p(a)), "get particles")
        return a

    def set_part(self, q, s, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(self.L.fr_sim_set_particles(self.h, len(a), q, s, _p(a)), "set particles")

    def flags(self):
        a = np.empty(self.np_, dtype=np.int32)
        check(self.L.fr_sim_get_flags(self.h, _p(a)))
        return a

    def set_flags(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        check(self.L.fr_sim_set_flags(self.h, len(a), _p(a)))

    def set_inflow(self, planes, inlet=None):
        """Duct inflow: planes (8, Nz, Ny); inlet particles (n, 29)."""
        pl = None if planes is None else np.ascontiguousarray(planes, dtype=np.float64)
        n = 0 if inlet is None else len(inlet)
        ib = None if inlet is None else np.ascontiguousarray(inlet, dtype=np.float64)
        check(self.L.fr_sim_set_inflow(self.h, None if pl is None else _p(pl), n,
                                       None if ib is None else _p(ib)), "set_inflow")

    def reserve_particles(self, capacity):
        check(self.L.fr_sim_reserve_particles(self.h, int(capacity)), "reserve_particles")

    def statistics(self):
        check(self.L.fr_sim_statistics(self.h), "statistics")

    def store(self):
        """0 on success, nonzero (FR_EINVAL) when a file could not be
        written, e.g. a particle snapshot that exists."""
        return self.L.fr_sim_store(self.h)

    def touch(self):
        stop = C.c_int()
        check(self.L.fr_sim_touch(self.h, C.byref(stop)), "touch")
        return stop.value

    @property
    def num_timestep(self):
        n = C.c_int()
        check(self.L.fr_sim_num_timestep(self.h, C.byref(n)))
        return n.value

    def set_resort_interval(self, steps):
        check(self.L.fr_sim_set_resort_interval(self.h, int(steps)), "set_resort_interval")

    def particles_changed(self):
        check(self.L.fr_sim_particles_changed(self.h))

    def diagnostics(self):
        out = np.zeros(10)
        check(self.L.fr_sim_diagnostics(self.h, _p(out)), "diagnostics")
        return out

    def timing(self):
        """[step ms, Poisson ms, particle ms, Poisson iterations, momentum
        kernel ms, its launches, forward x-pass ms, its launches, momentum
        deposit kernel ms, its launches, trilinear deposit kernels ms, their
        launches]."""
        out = np.zeros(12)
        check(self.L.fr_sim_timing_ex(self.h, _p(out), 12))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.fr_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pair:
    """The coupled run (fr_pair_* of include/fr_amd.h, simulation.cpp:19-73):
    a periodic box Sim feeding a duct Sim of the same Ny, Nz, Ly, Lz.  The
    pair keeps both Sims alive; close it before them."""

    def __init__(self, box: Sim, duct: Sim):
        self.L = lib()
        self.box, self.duct = box, duct
        h = C.c_void_p()
        check(self.L.fr_pair_create(box.h, duct.h, C.byref(h)), "fr_pair_create")
        self.h = h

    def initialize(self):
        check(self.L.fr_pair_initialize(self.h), "fr_pair_initialize")

    def step(self):
        check(self.L.fr_pair_step(self.h), "fr_pair_step")

    def touch(self):
        stop = C.c_int()
        check(self.L.fr_pair_touch(self.h, C.byref(stop)), "fr_pair_touch")
        return stop.value

    def close(self):
        if getattr(self, "h", None):
            self.L.fr_pair_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
