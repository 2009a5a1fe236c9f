"""Parity of the gfx950 path (through the C ABI of libfr_amd.so) with the CPU
oracle on identical inputs.

Bars: stencils, halos, deposits, migration order and cell indices are
BIT-EXACT; the FFT pressure solve uses the oracle's own radix-2 algorithm and
twiddles, so it is bit-exact too when given the same inputs; whole substeps
and steps, whose zero-mean shifts use tree reductions instead of the
reference's serial sums, agree within 1e-12 normwise relative difference
(max|a-b|/max|b|) per quantity, with particle order bit-exact.
"""
import numpy as np
import pytest

import first_repo_amd as fr
import oracle_lib as OL
from gpu_util import (dbuf, dptr, host, rel, make_box_case, sync_oracle_to_device, compare_state,
                      sim_with_deposit)

pytestmark = pytest.mark.gpu

GEOMS = [((8, 8, 8), True), ((7, 9, 10), True), ((12, 8, 6), False)]


def _g(n, box):
    return fr.make_geom(*n, L=(1.3, 0.7 * n[1] / 8, 1.1), box=box)


@pytest.mark.parametrize("geo", GEOMS, ids=str)
def test_stencils_bit_exact(geo):
    L = fr.lib()
    g = _g(*geo)
    og = OL.Geom.from_buffer_copy(g)
    shp = (g.Nz, g.Ny, g.Nx)
    rng = np.random.default_rng(1)
    a, b, c = (rng.standard_normal(shp) for _ in range(3))
    da, db, dc = dbuf(a), dbuf(b), dbuf(c)
    O = OL.orc()
    for op in ["div", "del2", "grad", "ix", "iy", "iz", "fx", "fy", "fz", "if2c", "halo", "cum"]:
        prior = rng.standard_normal(shp)
        ref = prior.copy()
        out = dbuf(prior)
        if op == "div":
            fr.check(L.fr_t0_div_f2c(g, dptr(out), dptr(da), dptr(db), dptr(dc), None))
            O.orc_div_f2c(og, ref, a, b, c)
        elif op == "del2":
            fr.check(L.fr_t0_del2(g, dptr(out), dptr(da), None))
            O.orc_del2(og, ref, a)
        elif op == "grad":
            p2, p3 = rng.standard_normal(shp), rng.standard_normal(shp)
            r2, r3 = p2.copy(), p3.copy()
            o2, o3 = dbuf(p2), dbuf(p3)
            fr.check(L.fr_t1_grad_c2f(g, dptr(out), dptr(o2), dptr(o3), dptr(da), None))
            O.orc_grad_c2f(og, ref, r2, r3, a)
            assert np.array_equal(host(o2), r2) and np.array_equal(host(o3), r3)
        elif op in ("ix", "iy", "iz", "fx", "fy", "fz"):
            axis = "xyz".index(op[1])
            direction = 0 if op[0] == "i" else 1
            fr.check(L.fr_t0_interp(g, dptr(out), dptr(da), axis, direction, None))
            O.orc_interp(og, ref, a, axis, direction)
        elif op == "if2c":
            fr.check(L.fr_t0_i_f2c(g, dptr(out), dptr(da), dptr(db), dptr(dc), None))
            O.orc_i_f2c(og, ref, a, b, c)
        else:
            cum = 1 if op == "cum" else 0
            fr.check(L.fr_t0_update_ghosts(g, dptr(out), cum, None))
            O.orc_update_ghosts(og, ref, cum)
        assert np.array_equal(host(out), ref), op


@pytest.mark.parametrize("geo", GEOMS, ids=str)
def test_elementwise_bit_exact(geo):
    """fr_t0_elementwise, every form (tensor0.cpp:93-241), against the
    oracle's tensor0 operators: bit-exact over the whole ghosted array."""
    L = fr.lib()
    g = _g(*geo)
    og = OL.Geom.from_buffer_copy(g)
    shp = (g.Nz, g.Ny, g.Nx)
    rng = np.random.default_rng(12)
    cases = [(0, 1, 1, 1), (0, 1, 1, 0), (0, 1, 0, 0), (1, 1, 0, 0), (2, 1, 1, 1),
             (3, 1, 0, 0), (4, 1, 1, 0), (5, 0, 0, 0), (6, 0, 0, 0), (7, 0, 0, 0),
             (8, 1, 0, 0), (9, 1, 0, 0), (10, 1, 0, 0), (11, 1, 0, 0)]
    for op, ua, ub, uc in cases:
        A, B, Cc = (1.5 + rng.random(shp) for _ in range(3))
        prior = rng.standard_normal(shp)
        a, b, c = 0.37, -1.25, 2.5
        dA, dB, dC = dbuf(A), dbuf(B), dbuf(Cc)
        out = dbuf(prior)
        fr.check(L.fr_t0_elementwise(g, op, dptr(out), a, dptr(dA) if ua else None, b,
                                     dptr(dB) if ub else None, c, dptr(dC) if uc else None,
                                     None), f"elementwise {op}")
        ref = prior.copy()
        OL.orc().orc_elementwise(og, op, ref, a, OL.ptr(A) if ua else None, b,
                                 OL.ptr(B) if ub else None, c, OL.ptr(Cc) if uc else None)
        assert np.array_equal(host(out), ref), (op, ua, ub, uc)


@pytest.mark.parametrize("geo", GEOMS, ids=str)
def test_reductions(geo):
    import ctypes as C
    L = fr.lib()
    g = _g(*geo)
    og = OL.Geom.from_buffer_copy(g)
    a = np.random.default_rng(2).standard_normal((g.Nz, g.Ny, g.Nx))
    da = dbuf(a)
    O = OL.orc()
    for op in range(6):
        r = C.c_double()
        fr.check(L.fr_t0_reduce(g, dptr(da), op, C.byref(r), None))
        expect = O.orc_reduce(og, a, op)
        if op in (2, 3, 4):
            assert r.value == expect, op
        else:  # tree vs serial sum order
            assert abs(r.value - expect) <= 1e-12 * np.abs(a).sum(), op
    out = dbuf(np.zeros(g.Nx))
    fr.check(L.fr_t0_reducex(g, dptr(da), dptr(out), None))
    ref = np.zeros(g.Nx)
    O.orc_reducex(og, a, ref)
    assert rel(host(out), ref) < 1e-14


@pytest.mark.parametrize("n", [(8, 8, 8), (16, 8, 32), (32, 32, 32), (64, 64, 64), (16, 256, 512),
                               (8, 512, 256)], ids=str)
def test_box_poisson_matches_oracle(n):
    """fr_box_poisson_solve vs the oracle's poisson::FFT restatement on a
    variable-coefficient problem (density ratio ~2).  The 256- and 512-point
    y and z transforms run one pencil per wave (the last two shapes), the
    others in LDS workgroup tiles."""
    import ctypes as C
    L = fr.lib()
    g = fr.make_geom(*n, L=(2 * np.pi, 1.0, 3.0))
    og = OL.Geom.from_buffer_copy(g)
    shp = (g.Nz, g.Ny, g.Nx)
    rng = np.random.default_rng(3)
    O = OL.orc()
    # smooth positive density, periodic ghosts via the oracle halo
    rho = 1.0 + 0.5 * rng.random(shp)
    O.orc_update_ghosts(og, rho, 0)
    C3 = []
    for axis in range(3):
        rf = np.zeros(shp)
        O.orc_interp(og, rf, rho, axis, 0)
        C3.append(1.0 / rf)
    rhs = rng.standard_normal(shp)
    rhs -= rhs[1:-1, 1:-1, 1:-1].mean()
    P0 = np.zeros(shp)
    ref = P0.copy()
    it_ref = O.orc_poisson(og, C3[0], C3[1], C3[2], ref, rhs, 30, 1e-10)
    h = C.c_void_p()
    fr.check(L.fr_box_poisson_create(g, C.byref(h)))
    dC = [dbuf(c) for c in C3]
    dP, dR = dbuf(P0), dbuf(rhs)
    it = C.c_int()
    fr.check(L.fr_box_poisson_solve(h, dptr(dC[0]), dptr(dC[1]), dptr(dC[2]), dptr(dP),
                                    dptr(dR), 30, 1e-10, C.byref(it), None))
    L.fr_box_poisson_destroy(h)
    assert it.value == it_ref
    assert np.array_equal(host(dP), ref), rel(host(dP), ref)


@pytest.mark.parametrize("n", [(12, 8, 4), (10, 16, 8), (48, 32, 16)], ids=str)
def test_duct_poisson_matches_oracle(n):
    """fr_duct_poisson_solve vs the oracle's poisson_IO::Solve restatement
    (2-D FFT + Thomas defect correction, Neumann x ghosts) on a
    variable-coefficient problem: bit-exact P and the same iteration count."""
    import ctypes as C
    L = fr.lib()
    g = fr.make_geom(*n, L=(3.0, 1.0, 2 * np.pi), box=False)
    og = OL.Geom.from_buffer_copy(g)
    shp = (g.Nz, g.Ny, g.Nx)
    rng = np.random.default_rng(7)
    O = OL.orc()
    rho = 1.0 + 2.0 * rng.random(shp)
    O.orc_update_ghosts(og, rho, 0)
    C3 = []
    for axis in range(3):
        rf = np.ones(shp)  # the inlet x-face ghost keeps 1 (unused by the solve)
        O.orc_interp(og, rf, rho, axis, 0)
        C3.append(1.0 / rf)
    rhs = rng.standard_normal(shp)
    rhs -= rhs[1:-1, 1:-1, 1:-1].mean()  # solvable with Neumann x ends
    P0 = np.zeros(shp)
    ref = P0.copy()
    it_ref = O.orc_poisson(og, C3[0], C3[1], C3[2], ref, rhs, 30, 1e-10)
    assert it_ref < 30
    h = C.c_void_p()
    fr.check(L.fr_duct_poisson_create(g, C.byref(h)))
    dC = [dbuf(c) for c in C3]
    dP, dR = dbuf(P0), dbuf(rhs)
    it = C.c_int()
    fr.check(L.fr_duct_poisson_solve(h, dptr(dC[0]), dptr(dC[1]), dptr(dC[2]), dptr(dP),
                                     dptr(dR), 30, 1e-10, C.byref(it), None))
    L.fr_duct_poisson_destroy(h)
    assert it.value == it_ref
    assert np.array_equal(host(dP), ref), rel(host(dP), ref)


@pytest.mark.parametrize("box", [True, False], ids=["box", "duct"])
def test_poisson_batched_solves_match_oracle(box):
    """One solver object, several solves: after the first, iterations are
    queued in batches as long as the previous solve and stopped on the
    device by the convergence test.  A looser tolerance (the batch overshoots
    the count), an iteration cap below convergence and a repeat of the first
    solve must each give the oracle's loop: the same count and P bit for
    bit."""
    import ctypes as C
    L = fr.lib()
    n = (16, 8, 32) if box else (48, 32, 16)
    g = fr.make_geom(*n, L=(3.0, 1.0, 2 * np.pi), box=box)
    og = OL.Geom.from_buffer_copy(g)
    shp = (g.Nz, g.Ny, g.Nx)
    rng = np.random.default_rng(11)
    O = OL.orc()
    rho = 1.0 + 1.5 * rng.random(shp)
    O.orc_update_ghosts(og, rho, 0)
    C3 = []
    for axis in range(3):
        rf = np.ones(shp)
        O.orc_interp(og, rf, rho, axis, 0)
        C3.append(1.0 / rf)
    rhs = rng.standard_normal(shp)
    rhs -= rhs[1:-1, 1:-1, 1:-1].mean()
    create = L.fr_box_poisson_create if box else L.fr_duct_poisson_create
    solve = L.fr_box_poisson_solve if box else L.fr_duct_poisson_solve
    destroy = L.fr_box_poisson_destroy if box else L.fr_duct_poisson_destroy
    h = C.c_void_p()
    fr.check(create(g, C.byref(h)))
    dC = [dbuf(c) for c in C3]
    dR = dbuf(rhs)
    counts = []
    for maxit, eps in [(30, 1e-10), (30, 1e-5), (3, 1e-12), (30, 1e-10)]:
        print(f"solve maxit={maxit} eps={eps}", flush=True)
        ref = np.zeros(shp)
        it_ref = O.orc_poisson(og, C3[0], C3[1], C3[2], ref, rhs, maxit, eps)
        dP = dbuf(np.zeros(shp))
        it = C.c_int()
        fr.check(solve(h, dptr(dC[0]), dptr(dC[1]), dptr(dC[2]), dptr(dP), dptr(dR), maxit, eps,
                       C.byref(it), None))
        assert it.value == it_ref, (maxit, eps, it.value, it_ref)
        assert np.array_equal(host(dP), ref), (maxit, eps, rel(host(dP), ref))
        counts.append(it.value)
    destroy(h)
    assert counts[2] == 3 and counts[1] < counts[0] == counts[3]


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    return make_box_case(tmp_path_factory.mktemp("box"), n=16, n_particles=3000)


def test_initialize_matches_oracle(case):
    params, d = case
    og = OL.OracleGrid(params, d)
    og.initialize()
    sim = fr.Sim(params, d)
    sim.initialize()
    worst, bad = compare_state(og, sim, tol=0.0)
    assert not bad, bad


# Live outputs of each grid method (grid.cpp) that the next methods read.
OUTPUTS = {
    "Update_Rho": ([], ["Rho_np1", "Rho_new"], False),
    "Update_P0": ([], [], False),
    "Update_Particle": (["U", "RHS_RU"], ["C", "RHS_Part_Temp"], True),
    "Update_RU_WOP": (["RU_np1", "RU_new"], [], False),
    "Compute_Div_U_new": ([], [], False),
    "Compute_RHS_Pois": (["Rho_face"], ["RHS_Pois"], False),
    "Solve_Poisson": ([], ["P"], False),
    "Update_RU_WP": (["RU_np1", "RU_new"], [], False),
    "TimeAdvance_RK4": (["RU_int"], ["Rho_int"], True),
    "Correct_mean_RU": (["RU_np1"], [], False),
    "TimeAdvance": (["RU"], ["Rho"], True),
}


def test_substep_methods_match_oracle(case):
    """Each grid method of one step from an identical state."""
    params, d = case
    og = OL.OracleGrid(params, d)
    og.initialize()
    og.step()  # leave the trivial initial state
    sim = fr.Sim(params, d)
    sim.initialize()
    sync_oracle_to_device(og, sim)
    og.call("update")
    sim.call("update")
    seq = ["Update_Rho", "Update_P0", "Update_Particle", "Update_RU_WOP", "Compute_Div_U_new",
           "Compute_RHS_Pois", "Solve_Poisson", "Update_RU_WP", "TimeAdvance_RK4"]
    steps = [(k, m) for k in range(4) for m in seq] + [(3, "Correct_mean_RU"), (3, "TimeAdvance")]
    for k, m in steps:
        og.call(m, k)
        sim.call(m, k)
        f1, f0, parts = OUTPUTS[m]
        stages = (0, 1, 2, 3) if m == "Update_Particle" else (0, 1, 3)
        worst, bad = compare_state(og, sim, tol=1e-12, fields1=f1, fields0=f0, particles=parts,
                                   stages=stages)
        assert not bad, (k, m, bad)
        for s in ("P0_new", "P0_np1", "P0_int", "P0"):
            a, b = sim.scalar(s), og.scalar(s)[0]
            assert abs(a - b) <= 1e-12 * abs(b), (k, m, s)


def test_steps_match_oracle(case):
    """Full RK4 steps from an identical state: per step within 1e-12 relative,
    particle order and cell indices bit-exact, diagnostics within 1e-12."""
    params, d = case
    og = OL.OracleGrid(params, d)
    og.initialize()
    sim = fr.Sim(params, d)
    sim.initialize()
    for step in range(3):
        sync_oracle_to_device(og, sim)
        og.step()
        sim.step()
        worst, bad = compare_state(og, sim, tol=1e-12)
        assert not bad, (step, bad)
        do, dg = og.diagnostics(), sim.diagnostics()
        for i in range(9):
            assert abs(dg[i] - do[i]) <= 1e-12 * max(abs(do[i]), 1e-300), (i, dg[i], do[i])
        assert dg[9] == do[9]


@pytest.mark.parametrize("interval,form",
                         [(1, "pull"), (4, "pull"), (4, "tile"), (1, "split"), (4, "split")])
def test_lockstep_steps_match_oracle(case, interval, form):
    """Six RK4 steps with no state copied between them: the device's own
    buckets, storage re-sorts (every `interval` steps, so the re-sort at step
    4 runs on ranks that migration rewrote) and migrations carried from step
    to step, with each deposit form.  Per step: fields and particles within
    1e-12 relative per step taken, particle order bit-exact, diagnostics
    likewise."""
    params, d = case
    og = OL.OracleGrid(params, d)
    og.initialize()
    sim = sim_with_deposit(params, d, form)
    sim.set_resort_interval(interval)
    sim.initialize()
    for step in range(6):
        og.step()
        sim.step()
        tol = 1e-12 * (step + 1)
        worst, bad = compare_state(og, sim, tol=tol)
        assert not bad, (interval, step, bad)
        do, dg = og.diagnostics(), sim.diagnostics()
        for i in range(9):
            assert abs(dg[i] - do[i]) <= tol * max(abs(do[i]), 1e-300), (step, i, dg[i], do[i])
        assert dg[9] == do[9]


def test_deposit_forms_agree_bit_for_bit(tmp_path):
    """The split form (direct sums of targets with at most two contributions,
    merges of the rest, operand records) and the pull form (every target
    merged) give the reference's ordered sums, so three full steps agree bit
    for bit in the momentum and density fields and the particles, and the
    last momentum deposit RHS_RU agrees over the interior.  The heat sum comes
    from the folded field in both (the split deposit's own partials add the
    same values in another order)."""
    params, d = make_box_case(tmp_path, n=16, n_particles=4000, seed=13)
    a = sim_with_deposit(params, d, "split", heat_sum="field")
    a.initialize()
    b = sim_with_deposit(params, d, "pull", heat_sum="field")
    b.initialize()
    for step in range(3):
        a.step()
        b.step()
        for f in ("RU", "RU_np1", "RU_int", "Rho"):
            for c in range(3 if f != "Rho" else 1):
                fa = a.field(f, c) if f != "Rho" else a.field(f)
                fb = b.field(f, c) if f != "Rho" else b.field(f)
                assert np.array_equal(fa, fb), (step, f, c)
        assert np.array_equal(a.flags(), b.flags())
        for q in range(7):
            assert np.array_equal(a.part(q, 0), b.part(q, 0)), (step, q)
    n = 18
    for c in range(3):
        ra = a.field("RHS_RU", c).reshape(n, n, n)[1:-1, 1:-1, 1:-1]
        rb = b.field("RHS_RU", c).reshape(n, n, n)[1:-1, 1:-1, 1:-1]
        assert np.array_equal(ra, rb), c
        assert np.abs(rb).max() > 0


def test_per_particle_stokes_steps_match_oracle(tmp_path):
    """The per-particle Stokes-number extension (params Tp_lo Tp_hi, Stokes
    numbers uniform on [0.5, 4] as in the 256^3 workload): three lockstep
    steps at 16^3 with 3000 particles, same bars as the plain steps."""
    params, d = make_box_case(tmp_path, n=16, n_particles=3000, seed=5,
                              stokes_range=(0.5, 4.0))
    og = OL.OracleGrid(params, d)
    og.initialize()
    sim = fr.Sim(params, d)
    sim.initialize()
    for step in range(3):
        og.step()
        sim.step()
        tol = 1e-12 * (step + 1)
        worst, bad = compare_state(og, sim, tol=tol)
        assert not bad, (step, bad)
        do, dg = og.diagnostics(), sim.diagnostics()
        for i in range(9):
            assert abs(dg[i] - do[i]) <= tol * max(abs(do[i]), 1e-300), (step, i, dg[i], do[i])


def test_migration_order_many_leavers(case):
    """particle::Send_Recv with a large fraction of leavers in every axis."""
    params, d = case
    og = OL.OracleGrid(params, d)
    og.initialize()
    sim = fr.Sim(params, d)
    sim.initialize()
    rng = np.random.default_rng(9)
    L = og.g.Lx
    for trial in range(3):
        for q in range(3):
            new = og.part(q, 2)
            sel = rng.random(new.size) < 0.3
            new[sel] += rng.choice([-1.0, 1.0], sel.sum()) * rng.uniform(0.01, 0.4, sel.sum()) * L
        sync_oracle_to_device(og, sim)
        og.call("Send_Recv")
        sim.call("Send_Recv")
        assert np.array_equal(sim.flags(), og.flag())
        for q in range(7):
            for s in range(4):
                assert np.array_equal(sim.part(q, s), og.part(q, s)), (q, s)


@pytest.mark.parametrize("form", ["tile", "pull", "split"])
@pytest.mark.parametrize("npart", [30000, 70000])
def test_dense_deposits_match_oracle(tmp_path, npart, form):
    """particle::Update_Particle at about 7 and 17 particles per cell: the
    trilinear deposits' tiles hold more contributions than their LDS list,
    so they take the dense path, which merges per target from the tile's
    particles staged in LDS (7 per cell) or, when they do not fit, from
    global memory (17 per cell); the pull form merges every target's eight
    buckets from global memory, and the split form does so for nearly every
    target (three or more contributions), compacted per workgroup.  The
    concentration deposit is bit-exact; the rest within the substep bar."""
    params, d = make_box_case(tmp_path, n=16, n_particles=npart)
    og = OL.OracleGrid(params, d)
    og.initialize()
    og.step()
    sim = sim_with_deposit(params, d, form)
    sim.initialize()
    sync_oracle_to_device(og, sim)
    og.call("update")
    sim.call("update")
    for m in ("Update_Rho", "Update_P0", "Update_Particle"):
        print(f"npart={npart} {m}", flush=True)
        og.call(m, 0)
        sim.call(m, 0)
    f1, f0, parts = OUTPUTS["Update_Particle"]
    worst, bad = compare_state(og, sim, tol=1e-12, fields1=f1, fields0=f0, particles=parts,
                               stages=(0, 1, 2, 3))
    assert not bad, bad
    assert worst["C"] == 0.0, worst["C"]


@pytest.mark.parametrize("form", ["tile", "pull", "split"])
def test_clustered_deposits_match_oracle(tmp_path, form):
    """A sparse particle set (mean tile below half the LDS list) with all of
    its particles gathered into 3x3x3 cells, about 37 per cell: in the tiled
    form the few dense tiles merge per target inside the sparse deposit
    kernel; in the pull and split forms those targets merge long bucket
    ranges (the split form sums the lone and paired contributions at the
    cluster's edge directly)."""
    params, d = make_box_case(tmp_path, n=16, n_particles=1000)
    og = OL.OracleGrid(params, d)
    og.initialize()
    og.step()
    rng = np.random.default_rng(5)
    for q, L in enumerate((og.g.Lx, og.g.Ly, og.g.Lz)):
        x = (0.3 + rng.random(og.np_) * 3.0 / 16.0) * L
        for s in range(4):
            og.part(q, s)[:] = x
    sim = sim_with_deposit(params, d, form)
    sim.initialize()
    sync_oracle_to_device(og, sim)
    og.call("update")
    sim.call("update")
    for m in ("Update_Rho", "Update_P0", "Update_Particle"):
        print(f"clustered {m}", flush=True)
        og.call(m, 0)
        sim.call(m, 0)
    f1, f0, parts = OUTPUTS["Update_Particle"]
    worst, bad = compare_state(og, sim, tol=1e-12, fields1=f1, fields0=f0, particles=parts,
                               stages=(0, 1, 2, 3))
    assert not bad, bad
    assert worst["C"] == 0.0, worst["C"]


def test_resort_moves_distinct_stage_slots(tmp_path):
    """The storage re-sort at the top of a time step reads stage slot 0 only
    when slots 1 and 3 equal it bit for bit (every step of the solver leaves
    them so), and all three slots otherwise.  Here a synced state has slots 1
    and 3 apart from slot 0; the first stage's particle update, which
    re-sorts, still matches the oracle (particles bit-exact in order)."""
    params, d = make_box_case(tmp_path, n=16, n_particles=3000, seed=11)
    og = OL.OracleGrid(params, d)
    og.initialize()
    og.step()
    rng = np.random.default_rng(4)
    Ls = (og.g.Lx, og.g.Ly, og.g.Lz)
    for q in range(7):
        scale = 1e-3 * (Ls[q] if q < 3 else 1.0)
        for s in (1, 3):
            a = og.part(q, s)
            # positions stay inside the box (a particle outside it has no
            # defined gather cell)
            keep = (a > 2 * scale) & (a < Ls[q] - 2 * scale) if q < 3 else np.ones(a.size, bool)
            a[keep] += rng.uniform(-scale, scale, int(keep.sum()))
    sim = fr.Sim(params, d)
    sim.set_resort_interval(1)
    sim.initialize()
    sync_oracle_to_device(og, sim)
    og.call("update")
    sim.call("update")
    for m in ("Update_Rho", "Update_P0", "Update_Particle"):
        og.call(m, 0)
        sim.call(m, 0)
    f1, f0, parts = OUTPUTS["Update_Particle"]
    worst, bad = compare_state(og, sim, tol=1e-12, fields1=f1, fields0=f0, particles=parts,
                               stages=(0, 1, 2, 3))
    assert not bad, bad
    assert np.array_equal(sim.flags(), og.flag())


def test_pressure_solve_64_matches_oracle(tmp_path):
    """grid::Solve_Poisson on a 64^3 box (C = 1/Rho_face formed inside the
    residual of the x pass) after the substep's earlier methods, from an
    identical state."""
    params, d = make_box_case(tmp_path, n=64, n_particles=20000)
    og = OL.OracleGrid(params, d)
    og.initialize()
    og.step()
    sim = fr.Sim(params, d)
    sim.initialize()
    sync_oracle_to_device(og, sim)
    og.call("update")
    sim.call("update")
    for m in ("Update_Rho", "Update_P0", "Update_Particle", "Update_RU_WOP", "Compute_Div_U_new",
              "Compute_RHS_Pois", "Solve_Poisson"):
        print(f"64^3 {m}", flush=True)
        og.call(m, 0)
        sim.call(m, 0)
    worst, bad = compare_state(og, sim, tol=1e-12, fields1=[], fields0=["P", "RHS_Pois"],
                               particles=False)
    assert not bad, bad


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True], ids=["fp64", "fp32"])
def test_cell_density_residual_equals_face_form(tmp_path, fp32):
    """Inside the box's method sequence the FFT solve's residual forms the face
    densities from Rho_new (as Compute_RHS_Pois forms Rho_face); a field read
    in between sends the next solve through the stored Rho_face instead.  Both
    must give the same P bit for bit and the same iteration count."""
    params, d = make_box_case(tmp_path, n=32, n_particles=4000)
    sim = fr.Sim(params, d, fp32=fp32)
    sim.initialize()
    sim.step()
    sim.call("update")
    for m in ("Update_Rho", "Update_P0", "Update_Particle", "Update_RU_WOP", "Compute_Div_U_new"):
        sim.call(m, 0)
    p0 = sim.field("P")
    sim.call("Compute_RHS_Pois", 0)
    it0 = sim.timing()[3]  # iterations counted so far (not a field read)
    sim.call("Solve_Poisson", 0)  # cell-density form
    it1 = sim.timing()[3]
    p_cell = sim.field("P")
    sim.set_field("P", p0)
    sim.call("Solve_Poisson", 0)  # the field reads above: stored Rho_face
    it2 = sim.timing()[3]
    p_face = sim.field("P")
    print(f"32^3 fp32={fp32} max|P_cell - P_face| = {np.max(np.abs(p_cell - p_face))}, "
          f"iterations {it1 - it0} and {it2 - it1}", flush=True)
    assert np.array_equal(p_cell, p_face)
    assert not np.array_equal(p_cell, p0)
    assert it1 - it0 == it2 - it1 > 0


@pytest.mark.gpu
def test_heat_sum_from_deposit_partials(tmp_path):
    """The split heat deposit sums its own outputs before the cumulative fold
    (each deposited value lands in one interior cell after it); the steps
    agree with those that reduce the folded field to rounding."""
    params, d = make_box_case(tmp_path, n=16, n_particles=4000, seed=5)
    a = sim_with_deposit(params, d, "split")
    a.initialize()
    b = sim_with_deposit(params, d, "split", heat_sum="field")
    b.initialize()
    for _ in range(3):
        a.step()
        b.step()
    for f in ("Rho", "RHS_Part_Temp", "P"):
        e = rel(a.field(f), b.field(f))
        print(f"heat sum partials vs field: {f} {e:.3e}", flush=True)
        assert e <= 1e-12, f
    for c in range(3):
        assert rel(a.field("RU", c), b.field("RU", c)) <= 1e-12
    assert np.array_equal(a.flags(), b.flags())
