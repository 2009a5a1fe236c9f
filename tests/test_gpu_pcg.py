"""Parity of the PCG + FP32 multigrid pressure solve on the GPU with the
oracle's restatement (same V-cycle arithmetic in float, same FP64 outer
iteration and reduction order): bit-exact solution, identical iteration
count and relative residual."""
import ctypes as C

import numpy as np
import pytest

import first_repo_amd as fr
import oracle_lib as OL
from gpu_util import dbuf, dptr, host, rel

pytestmark = pytest.mark.gpu


def _problem(n, box, seed=3, ratio=2.0):
    L = (2 * np.pi,) * 3 if box else (4 * np.pi, np.pi, np.pi)
    g = fr.make_geom(*n, L=L, box=box)
    og = OL.Geom.from_buffer_copy(g)
    shp = (g.Nz, g.Ny, g.Nx)
    rng = np.random.default_rng(seed)
    O = OL.orc()
    rho = 1.0 + ratio * rng.random(shp)
    O.orc_update_ghosts(og, rho, 0)
    rf = []
    for a in range(3):
        f = np.ones(shp)
        O.orc_interp(og, f, rho, a, 0)
        rf.append(f)
    rhs = rng.standard_normal(shp)
    inner = (slice(1, -1),) * 3
    rhs[inner] -= rhs[inner].mean()
    return g, og, rf, rhs


@pytest.mark.parametrize("n,box", [((16, 16, 16), True), ((32, 16, 16), False),
                                   ((64, 32, 32), True)], ids=str)
def test_pcg_matches_oracle(n, box):
    L = fr.lib()
    g, og, rf, rhs = _problem(n, box)
    shp = (g.Nz, g.Ny, g.Nx)
    P0 = np.zeros(shp)
    ref = P0.copy()
    rr = C.c_double()
    it_ref = OL.orc().orc_pcg(og, rf[0], rf[1], rf[2], ref, rhs, 100, 1e-10, C.byref(rr))
    h = C.c_void_p()
    fr.check(L.fr_pcg_create(g, C.byref(h)))
    d = [dbuf(a) for a in rf]
    dP, dR = dbuf(P0), dbuf(rhs)
    it = C.c_int()
    rel_dev = C.c_double()
    fr.check(L.fr_pcg_solve(h, dptr(d[0]), dptr(d[1]), dptr(d[2]), dptr(dP), dptr(dR), 100,
                            1e-10, C.byref(it), C.byref(rel_dev), None))
    L.fr_pcg_destroy(h)
    assert it.value == it_ref
    assert rel_dev.value == rr.value
    assert rr.value <= 1e-10
    assert np.array_equal(host(dP), ref), rel(host(dP), ref)


def test_box_steps_with_pcg_match_oracle(tmp_path):
    """Full RK4 steps of the box with Solver = 2 (PCG + FP32 multigrid
    pressure solve) on the GPU and in the oracle."""
    from gpu_util import make_box_case, compare_state
    params, d = make_box_case(tmp_path, n=16, n_particles=2000, extra={"Solver": 2})
    og = OL.OracleGrid(params, d)
    og.initialize()
    sim = fr.Sim(params, d)
    sim.initialize()
    for _ in range(2):
        og.step()
        sim.step()
    worst, bad = compare_state(og, sim)
    assert not bad, bad
    dg = sim.diagnostics()
    assert dg[9] > 0  # iterations of the last solve
