"""Workload writers: ``params.in`` files in the reference's fixed field order
(params.cpp:22-48) and the synthetic initial velocity files ``U.bin``,
``V.bin``, ``W.bin`` that ``Initial=0`` reads (grid.cpp:184-196).

The velocity field is generated here because the reference's ``U/V/W.bin``
come from an external generator that is not shipped.  It is a seeded random
solenoidal field with energy spectrum E(k) ~ k^4 exp(-2 (k/4)^2), built as the
discrete curl of a smooth random vector potential on the staggered grid, so
its discrete divergence (tensor0::Equal_Div_F2C) vanishes to round-off.
"""
from __future__ import annotations

import dataclasses
import math
import os
from dataclasses import dataclass, field, asdict

import numpy as np

# Appendix-B order of params.in (params.cpp:22-48).
_FIELDS = [
    ("IsBox",), ("Nx_tot", "Ny_tot", "Nz_tot"), ("Lx", "Ly", "Lz"), ("bs",),
    ("Np_track",), ("T_final", "dt", "threshold"), ("Statistics", "Stat_print"),
    ("Data_freq_fast", "Data_freq_slow"), ("Data_dir", "Stat_dir"), ("Initial",),
    ("PreCond", "Solver"), ("Cooling",), ("Iteration1", "Iteration2", "epsilon"),
    ("TWC", "ParticleGravity"), ("A1", "A2"), ("Rho0",), ("U0",), ("T0",),
    ("Mu0", "k"), ("gx1", "gy1", "gz1", "gx2", "gy2", "gz2"), ("Cp", "Cv"),
    ("np0",), ("Dp", "Rhop"), ("Cvp",), ("epsilon_p",), ("Nu",), ("I0_1", "I0_2"),
    # extension, after the reference's last field (its parser stops before
    # it): per-particle response times uniform on [Tp_lo, Tp_hi]; 0 0 keeps
    # the single Tp = Rhop Dp^2 / (18 Mu0)
    ("Tp_lo", "Tp_hi"),
]


@dataclass
class Params:
    IsBox: int = 1
    Nx_tot: int = 64
    Ny_tot: int = 64
    Nz_tot: int = 64
    Lx: float = 2 * math.pi
    Ly: float = 2 * math.pi
    Lz: float = 2 * math.pi
    bs: int = 1
    Np_track: int = 0
    T_final: float = 1e9
    dt: float = 1e-3
    threshold: float = 1e9
    Statistics: int = 0
    Stat_print: int = 0
    Data_freq_fast: int = 1000000
    Data_freq_slow: int = 1000000
    Data_dir: str = "./data/"
    Stat_dir: str = "./stat/"
    Initial: int = 0
    PreCond: int = 0
    Solver: int = 0
    Cooling: int = 1
    Iteration1: int = 30
    Iteration2: int = 30
    epsilon: float = 1e-10
    TWC: int = 1
    ParticleGravity: int = 0
    A1: float = 0.0
    A2: float = 0.0
    Rho0: float = 1.0
    U0: float = 0.0
    T0: float = 1.0
    Mu0: float = 0.01
    k: float = 0.05
    gx1: float = 0.0
    gy1: float = 0.0
    gz1: float = 0.0
    gx2: float = 0.0
    gy2: float = 0.0
    gz2: float = 0.0
    Cp: float = 3.5
    Cv: float = 2.5
    np0: float = 0.0
    Dp: float = 1e-3
    Rhop: float = 1e3
    Cvp: float = 1.5
    epsilon_p: float = 0.4
    Nu: float = 2.0
    I0_1: float = 0.0
    I0_2: float = 0.0
    Tp_lo: float = 0.0
    Tp_hi: float = 0.0

    def text(self) -> str:
        d = asdict(self)
        lines = ["# params.in written by first_repo_amd.configs (reference field order)"]
        for grp in _FIELDS:
            lines.append("# " + " ".join(grp))
            lines.append(" ".join(_fmt(d[f]) for f in grp))
            lines.append("")
        return "\n".join(lines) + "\n"

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.text())

    # derived values, params.cpp:50-56
    @property
    def R(self) -> float:
        return self.Cp - self.Cv

    @property
    def N0(self) -> int:
        return int(self.np0 * self.Lx * self.Ly * self.Lz)

    @property
    def Tp(self) -> float:
        return self.Rhop * self.Dp * self.Dp / (18 * self.Mu0)

    @property
    def mp(self) -> float:
        return self.Rhop * math.pi * self.Dp ** 3 / 6.0


def _fmt(v) -> str:
    if isinstance(v, str):
        return v
    if isinstance(v, (int, np.integer)) and not isinstance(v, bool):
        return str(int(v))
    return repr(float(v))


def solenoidal_velocity(n: tuple[int, int, int], L: tuple[float, float, float],
                        seed: int, u_rms: float = 1.0, k_peak: float = 4.0):
    """Face velocities (U, V, W), each a [Nz][Ny][Nx] float64 block.

    U(i,j,k) lives on the left x-face of cell (i,j,k), like tensor1.x.
    A = smooth random vector potential on cell edges (|A_hat| ~
    exp(-(k/k_peak)^2), so E(k) ~ k^4 exp(-2 (k/k_peak)^2)); u = discrete
    curl A.  Scaled so that sqrt(<u.u>/3) = u_rms.
    """
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    kx = np.fft.fftfreq(nx, d=1.0 / nx) * (2 * math.pi / L[0])
    ky = np.fft.fftfreq(ny, d=1.0 / ny) * (2 * math.pi / L[1])
    kz = np.fft.rfftfreq(nz, d=1.0 / nz) * (2 * math.pi / L[2])
    # arrays are indexed [i][j][k] here and transposed at the end
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    filt = np.exp(-(kk / k_peak) ** 2)
    filt[0, 0, 0] = 0.0
    A = []
    for _ in range(3):
        a = rng.standard_normal((nx, ny, nz))
        A.append(np.fft.irfftn(np.fft.rfftn(a) * filt, s=(nx, ny, nz), axes=(0, 1, 2)))
    Ax, Ay, Az = A
    dx, dy, dz = L[0] / nx, L[1] / ny, L[2] / nz

    def fwd(a, ax):  # a(index+1) - a(index), periodic
        return np.roll(a, -1, axis=ax) - a

    u = fwd(Az, 1) / dy - fwd(Ay, 2) / dz
    v = fwd(Ax, 2) / dz - fwd(Az, 0) / dx
    w = fwd(Ay, 0) / dx - fwd(Ax, 1) / dy
    s = u_rms / math.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3.0)
    return tuple(np.ascontiguousarray((c * s).transpose(2, 1, 0)) for c in (u, v, w))


@dataclass
class BoxWorkload:
    """Triply periodic box (Initial=0), two-way coupled, radiatively heated."""
    n: int = 64
    n_particles: int = 100_000
    steps: int = 50
    seed: int = 1234
    re_lambda: float = 40.0
    prandtl: float = 0.7
    stokes: float = 1.0
    # per-particle Stokes numbers uniform on this range (the params.in
    # extension Tp_lo Tp_hi); None keeps one Stokes number for all
    stokes_range: tuple | None = None
    mass_loading: float = 0.2
    cfl: float = 0.5
    heating_rate: float = 0.5     # initial particle dT/dt from radiation, units T0/time
    epsilon: float = 1e-10
    iterations: int = 30
    extra: dict = field(default_factory=dict)

    def build(self, outdir: str) -> dict:
        """Write params.in and U/V/W.bin under outdir; return derived values."""
        os.makedirs(outdir, exist_ok=True)
        L = 2 * math.pi
        n = self.n
        U, V, W = solenoidal_velocity((n, n, n), (L, L, L), self.seed)
        for name, a in (("U.bin", U), ("V.bin", V), ("W.bin", W)):
            a.astype("<f8").tofile(os.path.join(outdir, name))
        p = self.derive(U, V, W)
        p.write(os.path.join(outdir, "params.in"))
        return {"params": p, "dt": p.dt, "N0": p.N0}

    def derive(self, U, V, W) -> Params:
        n = self.n
        L = 2 * math.pi
        dx = L / n
        rho, T0 = 1.0, 1.0
        cv, cp = 2.5, 3.5  # gamma = 1.4, R = 1
        # Taylor microscale from the discrete spectrum: lambda^2 = 15 u'^2 / <|grad u|^2>
        up2 = (np.mean(U * U) + np.mean(V * V) + np.mean(W * W)) / 3.0
        g2 = 0.0
        for a, ax in ((U, 2), (V, 1), (W, 0)):
            for axis in range(3):
                d = (np.roll(a, -1, axis=axis) - a) / dx
                g2 += np.mean(d * d)
        lam = math.sqrt(15.0 * up2 / g2) if g2 > 0 else 1.0
        nu = math.sqrt(up2) * lam / self.re_lambda
        mu = nu * rho
        eps = nu * g2  # dissipation estimate 2 nu <S_ij S_ij> ~ nu <|grad u|^2>
        tau_eta = math.sqrt(nu / eps) if eps > 0 else 1.0
        Tp = self.stokes * tau_eta
        vol = L ** 3
        N = self.n_particles
        mp = self.mass_loading * rho * vol / N
        Dp = mp / (3 * math.pi * mu * Tp)
        Rhop = 18 * mu * Tp / (Dp * Dp)
        umax = float(max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max()))
        dt = self.cfl * dx / umax
        cvp = 0.6 * cv
        eps_p = 0.4
        # radiative heating: alpha/(mp Cvp) = heating_rate
        I0 = self.heating_rate * mp * cvp / (0.25 * eps_p * math.pi * Dp * Dp)
        A = eps / (3.0 * up2) if up2 > 0 else 0.0
        p = Params(
            IsBox=1, Nx_tot=n, Ny_tot=n, Nz_tot=n, Lx=L, Ly=L, Lz=L, dt=dt,
            Initial=0, Cooling=1, Iteration1=self.iterations, Iteration2=self.iterations,
            epsilon=self.epsilon, TWC=1, A1=A, A2=A, Rho0=rho, U0=0.0, T0=T0,
            Mu0=mu, k=mu * cp / self.prandtl, Cp=cp, Cv=cv,
            np0=(N + 0.5) / vol, Dp=Dp, Rhop=Rhop, Cvp=cvp, epsilon_p=eps_p,
            Nu=2.0, I0_1=I0, I0_2=I0, T_final=dt * self.steps * 1.5,
        )
        if self.stokes_range is not None:
            p.Tp_lo = self.stokes_range[0] * tau_eta
            p.Tp_hi = self.stokes_range[1] * tau_eta
        for k, v in self.extra.items():
            setattr(p, k, v)
        assert p.N0 == N, (p.N0, N)
        return p


# Named workloads (the measured configurations).
WORKLOADS = {
    "box64": BoxWorkload(n=64, n_particles=100_000, steps=50),
    "box256": BoxWorkload(n=256, n_particles=4_000_000, steps=200, mass_loading=0.5,
                          stokes_range=(0.5, 4.0)),
    "box512": BoxWorkload(n=512, n_particles=16_000_000, steps=200, mass_loading=0.5,
                          stokes_range=(0.5, 4.0)),
}


@dataclass
class DuctWorkload:
    """Inflow/outflow duct (IsBox=0, Initial=0), periodic in y and z.

    The reference feeds the duct from a box grid (simulation.cpp:52-66); here
    the inflow is synthetic: the box planes are a uniform stream U = u_in with
    `noise` relative random perturbations (a new seeded draw every step), and
    `inject_per_step` particles enter through the inlet each step.  The initial
    state is the same stream with a streamwise density profile falling from 1 to
    1/rho_ratio (gas heated along the duct), so the pressure solve has variable
    coefficients from the first step.
    """
    n: tuple = (512, 128, 128)
    L: tuple = (8 * math.pi, 2 * math.pi, 2 * math.pi)
    steps: int = 200
    seed: int = 4321
    u_in: float = 1.0
    noise: float = 0.05
    rho_ratio: float = 3.0
    inject_per_step: int = 20_000
    capacity: int = 8_000_000
    mu: float = 2e-3
    stokes_time: float = 0.5
    heating_rate: float = 0.5
    iterations: int = 30
    epsilon: float = 1e-10
    extra: dict = field(default_factory=dict)

    @property
    def dx(self):
        return self.L[0] / self.n[0]

    def params(self) -> Params:
        nx, ny, nz = self.n
        dt = 0.5 * self.dx / (self.u_in * (1 + 4 * self.noise))
        mu, cp, cv = self.mu, 3.5, 2.5
        Tp = self.stokes_time
        # particles of the box workload's kind: mass from Dp, Rhop at Stokes time Tp
        Dp = 0.02 * self.dx
        Rhop = 18 * mu * Tp / (Dp * Dp)
        mp = Rhop * math.pi * Dp ** 3 / 6.0
        cvp = 0.6 * cv
        eps_p = 0.4
        I0 = self.heating_rate * mp * cvp / (0.25 * eps_p * math.pi * Dp * Dp)
        p = Params(
            IsBox=0, Nx_tot=nx, Ny_tot=ny, Nz_tot=nz, Lx=self.L[0], Ly=self.L[1], Lz=self.L[2],
            dt=dt, Initial=0, Cooling=0, Iteration1=self.iterations,
            Iteration2=self.iterations, epsilon=self.epsilon, TWC=1, A1=0.0, A2=0.0,
            Rho0=1.0, U0=self.u_in, T0=1.0, Mu0=mu, k=mu * cp / 0.7, Cp=cp, Cv=cv, np0=0.0,
            Dp=Dp, Rhop=Rhop, Cvp=cvp, epsilon_p=eps_p, Nu=2.0, I0_1=I0, I0_2=I0,
            T_final=dt * self.steps * 1.5)
        for k, v in self.extra.items():
            setattr(p, k, v)
        return p

    def build(self, outdir: str) -> dict:
        """params.in and U/V/W.bin (the perturbation; U0 = u_in is added by
        grid::Initialize, grid.cpp:189)."""
        os.makedirs(outdir, exist_ok=True)
        nx, ny, nz = self.n
        rng = np.random.default_rng(self.seed)
        for name in ("U.bin", "V.bin", "W.bin"):
            a = self.noise * self.u_in * rng.standard_normal((nz, ny, nx))
            a.astype("<f8").tofile(os.path.join(outdir, name))
        p = self.params()
        p.write(os.path.join(outdir, "params.in"))
        return {"params": p, "dt": p.dt}

    def density(self):
        """Ghosted initial density [Nz+2][Ny+2][Nx+2]: 1 at the inlet falling
        linearly in 1/rho to 1/rho_ratio at the outlet."""
        nx, ny, nz = self.n
        x = (np.arange(nx + 2) - 0.5) / nx
        inv = 1 + (self.rho_ratio - 1) * np.clip(x, 0, 1)
        return np.broadcast_to((1.0 / inv)[None, None, :], (nz + 2, ny + 2, nx + 2)).copy()

    def inflow(self, step: int) -> np.ndarray:
        """The 8 box planes (rho_new, rho_np1, RU_new xyz, RU_np1 xyz),
        ghosted [Nz+2][Ny+2] with periodic ghosts."""
        _, ny, nz = self.n
        rng = np.random.default_rng((self.seed, step))
        pert = self.noise * self.u_in * rng.standard_normal((3, nz, ny))
        pl = np.zeros((8, nz + 2, ny + 2))
        pl[0] = pl[1] = 1.0
        for m in range(3):
            inner = pert[m] + (self.u_in if m == 0 else 0.0)
            pl[2 + m, 1:-1, 1:-1] = inner
            pl[5 + m, 1:-1, 1:-1] = inner
        for q in range(2, 8):  # periodic ghosts in y and z
            pl[q, 0, :] = pl[q, -2, :]
            pl[q, -1, :] = pl[q, 1, :]
            pl[q, :, 0] = pl[q, :, -2]
            pl[q, :, -1] = pl[q, :, 1]
        return pl

    def inlet_particles(self, step: int, first_flag: int, n: int | None = None) -> np.ndarray:
        """n particles entering this step (ToBuf layout, 29 values): x in
        [0, dx/2), y and z uniform, velocity u_in along x, T = T0, all four
        stages equal, flags first_flag, first_flag+1, ..."""
        n = self.inject_per_step if n is None else n
        rng = np.random.default_rng((self.seed, step, 7))
        b = np.zeros((n, 29))
        pos = [rng.random(n) * 0.5 * self.dx, rng.random(n) * self.L[1], rng.random(n) * self.L[2]]
        vals = pos + [np.full(n, self.u_in), np.zeros(n), np.zeros(n), np.ones(n)]
        for q, v in enumerate(vals):
            for s in range(4):
                b[:, q * 4 + s] = v
        b[:, 28] = first_flag + np.arange(n)
        return b

    def fill_particles(self, n: int, first_flag: int = 0) -> np.ndarray:
        """n particles as the inlet would have delivered them over the last
        n / inject_per_step steps, moving with the stream: x uniform over
        the distance the stream covers in that time (capped at Lx), y and z
        uniform, velocity u_in along x, T = T0, all four stages equal (ToBuf
        layout, 29 values).  Handed over with the first step's inlet
        particles, they load the duct to its working particle count."""
        dt = self.params().dt
        xf = min(self.L[0], n / self.inject_per_step * dt * self.u_in)
        rng = np.random.default_rng((self.seed, 99))
        b = np.zeros((n, 29))
        pos = [rng.random(n) * xf, rng.random(n) * self.L[1], rng.random(n) * self.L[2]]
        vals = pos + [np.full(n, self.u_in), np.zeros(n), np.zeros(n), np.ones(n)]
        for q, v in enumerate(vals):
            for s in range(4):
                b[:, q * 4 + s] = v
        b[:, 28] = first_flag + np.arange(n)
        return b


WORKLOADS["duct512"] = DuctWorkload()


@dataclass
class PairWorkload:
    """The coupled run of the reference (simulation.cpp:19-73): a periodic
    box of n^3 cells (the box workload's turbulence and particles, convected
    at U0 = u_conv) feeding a duct of ratio*n x n x n cells and length
    ratio*L (IsBox = 0, Initial = 2: uniform stream at U0, no particles)
    through its x = Lx face.  Both grids take the box's dt, gas and particle
    properties; the duct is not cooled and has no forcing, as in the
    reference's params_IO.in, which also writes its run files to ./data_IO/
    and ./stat_IO/.  Files: box/params.in with box/U,V,W.bin and
    duct/params.in."""
    box: BoxWorkload = field(default_factory=lambda: BoxWorkload(n=64, n_particles=100_000))
    ratio: int = 4
    u_conv: float = 1.0
    cfl: float = 0.5
    duct_capacity: int = 0  # 0: ratio x the box's particles x 1.5

    def build(self, outdir: str) -> dict:
        n, L = self.box.n, 2 * math.pi
        bdir, ddir = os.path.join(outdir, "box"), os.path.join(outdir, "duct")
        os.makedirs(bdir, exist_ok=True)
        os.makedirs(ddir, exist_ok=True)
        U, V, W = solenoidal_velocity((n, n, n), (L, L, L), self.box.seed)
        for name, a in (("U.bin", U), ("V.bin", V), ("W.bin", W)):
            a.astype("<f8").tofile(os.path.join(bdir, name))
        pb = self.box.derive(U, V, W)
        umax = float(max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max()))
        pb.U0 = self.u_conv
        pb.dt = self.cfl * (L / n) / (umax + self.u_conv)
        pb.T_final = pb.dt * self.box.steps * 1.5
        pb.write(os.path.join(bdir, "params.in"))
        pd = dataclasses.replace(pb)
        pd.IsBox, pd.Nx_tot, pd.Lx = 0, self.ratio * n, self.ratio * L
        pd.Initial, pd.Cooling, pd.A1, pd.A2 = 2, 0, 0.0, 0.0
        pd.Data_dir, pd.Stat_dir = "./data_IO/", "./stat_IO/"
        pd.write(os.path.join(ddir, "params.in"))
        cap = self.duct_capacity or int(1.5 * self.ratio * self.box.n_particles)
        return {"box": bdir, "duct": ddir, "dt": pb.dt, "duct_capacity": cap}


WORKLOADS["pair64"] = PairWorkload()
