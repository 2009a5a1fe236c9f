"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the partial
reference build (oracle/_ref/libref.so).  Test infrastructure: the oracle is
the checker, never the thing measured or shipped."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")
MPI_LIB = "/opt/conda/lib"

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Geom(C.Structure):
    """Same layout as fr_geom in include/fr_amd.h."""
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("bs", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("is_box", C.c_int)]


def make_geom(nx, ny, nz, L=(2 * np.pi, 2 * np.pi, 2 * np.pi), box=True, bs=1):
    g = Geom()
    g.Nx, g.Ny, g.Nz, g.bs = nx + 2 * bs, ny + 2 * bs, nz + 2 * bs, bs
    g.Lx, g.Ly, g.Lz = L
    g.dx, g.dy, g.dz = L[0] / nx, L[1] / ny, L[2] / nz
    g.is_box = 1 if box else 0
    return g


def shape(g):
    return (g.Nz, g.Ny, g.Nx)


_orc = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = C.CDLL(ORACLE_SO)
        G = C.POINTER(Geom)
        L.orc_div_f2c.argtypes = [G, dp, dp, dp, dp]
        L.orc_del2.argtypes = [G, dp, dp]
        L.orc_grad_c2f.argtypes = [G, dp, dp, dp, dp]
        L.orc_interp.argtypes = [G, dp, dp, C.c_int, C.c_int]
        L.orc_i_f2c.argtypes = [G, dp, dp, dp, dp]
        L.orc_update_ghosts.argtypes = [G, dp, C.c_int]
        L.orc_elementwise.argtypes = [G, C.c_int, dp, C.c_double, C.c_void_p, C.c_double,
                                      C.c_void_p, C.c_double, C.c_void_p]
        L.orc_post.argtypes = [G, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
        L.orc_post_scalars.argtypes = [G, C.c_void_p, C.c_double, C.c_void_p]
        L.orc_reduce.argtypes = [G, dp, C.c_int]
        L.orc_reduce.restype = C.c_double
        L.orc_reducex.argtypes = [G, dp, dp]
        L.orc_fft3d.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_fft2d.argtypes = [C.c_int, C.c_int, dp, dp, C.c_int]
        L.orc_rfft1d.argtypes = [dp, C.c_int, dp, dp]
        L.orc_pcg.argtypes = [G, dp, dp, dp, dp, dp, C.c_int, C.c_double,
                              C.POINTER(C.c_double)]
        L.orc_pcg.restype = C.c_int
        L.orc_irfft1d.argtypes = [dp, dp, C.c_int, dp]
        L.orc_poisson.argtypes = [G, dp, dp, dp, dp, dp, C.c_int, C.c_double]
        L.orc_poisson.restype = C.c_int
        L.orc_glibc_rand.argtypes = [C.c_int, ip]
        fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.orc_unet_num_params.argtypes = []
        L.orc_unet_num_params.restype = C.c_long
        L.orc_unet_forward.argtypes = [C.c_int, C.c_int, C.c_int, fp, C.c_int, dp, dp]
        L.orc_unet_grad.argtypes = [C.c_int, C.c_int, C.c_int, fp, C.c_int, dp, dp, dp, dp]
        L.orc_conv3.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_grid_create.argtypes = [C.c_char_p]
        L.orc_grid_create.restype = C.c_void_p
        L.orc_grid_create_ex.argtypes = [C.c_char_p, C.c_int]
        L.orc_grid_create_ex.restype = C.c_void_p
        L.orc_grid_statistics.argtypes = [C.c_void_p]
        L.orc_grid_store.argtypes = [C.c_void_p]
        L.orc_grid_store.restype = C.c_int
        L.orc_grid_touch.argtypes = [C.c_void_p]
        L.orc_grid_touch.restype = C.c_int
        L.orc_grid_write_info.argtypes = [C.c_void_p]
        L.orc_grid_num_timestep.argtypes = [C.c_void_p]
        L.orc_grid_num_timestep.restype = C.c_int
        L.orc_grid_destroy.argtypes = [C.c_void_p]
        L.orc_grid_initialize.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_grid_finish_initialize.argtypes = [C.c_void_p]
        L.orc_grid_step.argtypes = [C.c_void_p]
        L.orc_mg_vcycle.argtypes = [C.c_void_p] * 6 + [C.c_int] * 4
        L.orc_grid_sbuf0.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.orc_grid_sbuf0.restype = C.POINTER(C.c_double)
        L.orc_pair_initialize.argtypes = [C.c_void_p] * 8
        L.orc_pair_step.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_pair_touch.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_pair_touch.restype = C.c_int
        L.orc_grid_set_inflow.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_grid_call.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.orc_grid_call.restype = C.c_int
        L.orc_grid_field.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.orc_grid_field.restype = C.POINTER(C.c_double)
        L.orc_grid_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_grid_scalar.restype = C.POINTER(C.c_double)
        L.orc_grid_np.argtypes = [C.c_void_p]
        L.orc_grid_np.restype = C.c_int
        L.orc_grid_set_np.argtypes = [C.c_void_p, C.c_int]
        L.orc_grid_part.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_grid_part.restype = C.POINTER(C.c_double)
        L.orc_grid_flag.argtypes = [C.c_void_p]
        L.orc_grid_flag.restype = C.POINTER(C.c_int)
        L.orc_grid_diagnostics.argtypes = [C.c_void_p, dp]
        L.orc_grid_geom.argtypes = [C.c_void_p, C.POINTER(Geom)]
        L.orc_grid_part_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_grid_part_aux.argtypes = [C.c_void_p, C.c_int]
        L.orc_grid_part_aux.restype = C.POINTER(C.c_double)
        _orc = L
    return _orc


_ref = None


def ref_available():
    return os.path.exists(REF_SO) and os.path.exists(os.path.join(MPI_LIB, "libmpi.so.12"))


def ref():
    global _ref
    if _ref is None:
        # load the oracle first: libmpi would otherwise pull an older libgomp
        # from /opt/conda/lib that the oracle's OpenMP loops cannot use
        orc()
        C.CDLL(os.path.join(MPI_LIB, "libmpi.so.12"), mode=C.RTLD_GLOBAL)
        L = C.CDLL(REF_SO)
        vp = C.c_void_p
        L.ref_create.argtypes = [C.c_char_p]
        L.ref_create.restype = vp
        L.ref_size.argtypes = [vp]
        L.ref_size.restype = C.c_int
        L.ref_op.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.ref_reduce.argtypes = [vp, dp, C.c_int]
        L.ref_reduce.restype = C.c_double
        L.ref_reducex.argtypes = [vp, dp, dp]
        L.ref_part_set.argtypes = [vp, C.c_int, dp, ip, dp]
        L.ref_part_np.argtypes = [vp]
        L.ref_part_np.restype = C.c_int
        L.ref_part_get.argtypes = [vp, dp, ip, dp, vp, vp]
        L.ref_part_op.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, vp, vp, vp]
        L.ref_part_load_random.argtypes = [vp]
        L.ref_part_send_recv_duct.argtypes = [vp, C.c_int, vp]
        L.ref_part_send_recv_pair.argtypes = [vp, vp]
        L.ref_elementwise.argtypes = [vp, C.c_int, dp, C.c_double, vp, C.c_double, vp,
                                      C.c_double, vp]
        L.ref_post.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double]
        L.ref_post_scalars.argtypes = [vp, vp, C.c_double, vp]
        _ref = L
    return _ref


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleGrid:
    """The oracle's box grid (grid.cpp restated) driven from a params file."""

    def __init__(self, params_path, ic_dir=None, outputs=False):
        """outputs: write the reference's run files (touch.check, stat files,
        info.txt) relative to the working directory."""
        L = orc()
        self.L = L
        self.h = L.orc_grid_create_ex(params_path.encode(), 1 if outputs else 0)
        if not self.h:
            raise RuntimeError("oracle: cannot parse " + params_path)
        self.g = Geom()
        L.orc_grid_geom(self.h, C.byref(self.g))
        self.ic_dir = ic_dir

    def ic_blocks(self):
        """U.bin, V.bin, W.bin of ic_dir (kept alive on self), or Nones."""
        g = self.g
        n = (g.Nz - 2) * (g.Ny - 2) * (g.Nx - 2)
        if not self.ic_dir:
            return [None, None, None]
        self._ic = [np.fromfile(os.path.join(self.ic_dir, f), dtype="<f8", count=n)
                    for f in ("U.bin", "V.bin", "W.bin")]
        return [ptr(b) for b in self._ic]

    def initialize(self):
        self.L.orc_grid_initialize(self.h, *self.ic_blocks())

    def step(self):
        self.L.orc_grid_step(self.h)

    def set_inflow(self, planes, pbuf=None):
        """Duct inflow: planes (8, Nz, Ny) float64 or None (kept); pbuf (n, 29)
        particles."""
        if planes is not None:
            planes = np.ascontiguousarray(planes, dtype=np.float64)
        n = 0 if pbuf is None else len(pbuf)
        pb = None if pbuf is None else np.ascontiguousarray(pbuf, dtype=np.float64)
        self.L.orc_grid_set_inflow(self.h, ptr(planes), n, ptr(pb))

    def call(self, name, k=0):
        r = self.L.orc_grid_call(self.h, name.encode(), k)
        assert r == 0, name

    def field(self, name, comp=-1):
        p = self.L.orc_grid_field(self.h, name.encode(), comp)
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=shape(self.g))

    def scalar(self, name):
        return self.L.orc_grid_scalar(self.h, name.encode())

    @property
    def np_(self):
        return self.L.orc_grid_np(self.h)

    def set_np(self, n):
        self.L.orc_grid_set_np(self.h, n)

    def part(self, q, s):
        n = self.np_
        if n == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(self.L.orc_grid_part(self.h, q, s), shape=(n,))

    def flag(self):
        n = self.np_
        if n == 0:
            return np.zeros(0, np.int32)
        return np.ctypeslib.as_array(self.L.orc_grid_flag(self.h), shape=(n,))

    def part_op(self, op, k=0, pre=0.0, post=0.0, f0=None, f1=None, f2=None):
        self.L.orc_grid_part_op(self.h, op, k, pre, post, ptr(f0), ptr(f1), ptr(f2))

    def part_aux(self, which):
        n = self.np_
        return np.ctypeslib.as_array(self.L.orc_grid_part_aux(self.h, which), shape=(n,))

    def sbuf0(self):
        """Copy of the last x pass's '>hi' list, shape (n, 29)."""
        n = C.c_int()
        p = self.L.orc_grid_sbuf0(self.h, C.byref(n))
        if n.value == 0:
            return np.zeros((0, 29))
        return np.ctypeslib.as_array(p, shape=(n.value, 29)).copy()

    def statistics(self):
        self.L.orc_grid_statistics(self.h)

    def store(self):
        return self.L.orc_grid_store(self.h)

    def touch(self):
        return self.L.orc_grid_touch(self.h)

    @property
    def num_timestep(self):
        return self.L.orc_grid_num_timestep(self.h)

    def diagnostics(self):
        out = np.zeros(10)
        self.L.orc_grid_diagnostics(self.h, out)
        return out

    def __del__(self):
        try:
            self.L.orc_grid_destroy(self.h)
        except Exception:
            pass


class OraclePair:
    """The oracle's coupled box and duct (simulation.cpp:19-73): `box` and
    `duct` are OracleGrids of the same Ny, Nz; the duct's inflow planes and
    inlet particles come from the box every substep."""

    def __init__(self, box: OracleGrid, duct: OracleGrid):
        self.box, self.duct = box, duct
        self.L = box.L

    def initialize(self):
        self.L.orc_pair_initialize(self.box.h, self.duct.h, *self.box.ic_blocks(),
                                   *self.duct.ic_blocks())

    def step(self):
        self.L.orc_pair_step(self.box.h, self.duct.h)

    def touch(self):
        return self.L.orc_pair_touch(self.box.h, self.duct.h)
