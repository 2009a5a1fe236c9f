"""The learned preconditioner's network on the GPU (unet.hip) against the
oracle's restatement with BF16 storage emulated (oracle/src/unet.cpp, itself
pinned to torch and finite differences): forward output and every
parameter-gradient block; then the learned PCG forms on the duct with the
variable-coefficient solve (Solver = 2): pairs collected from the solver's
own runs, trained on the GPU, and the iterations to 1e-10 with and without
the network.

Tolerances: BF16 activations round at 2^-9 relative; the GPU accumulates in
FP32 (matrix cores, fixed per launch but not the oracle's order) and the
oracle in FP64, so a rounding tie can land on the other side; normwise
relative differences of 2e-2 (output) and 5e-2 (gradient blocks) bound that.
"""
import ctypes as C

import numpy as np
import pytest

import first_repo_amd as fr
import oracle_lib as OL
from gpu_util import dbuf, dptr, host
from unet_layout import offsets, make_params

pytestmark = pytest.mark.gpu


def _net(seed):
    L = fr.lib()
    h = C.c_void_p()
    fr.check(L.fr_unet_create(seed, C.byref(h)), "fr_unet_create")
    n = C.c_longlong()
    fr.check(L.fr_unet_num_params(h, C.byref(n)))
    assert n.value == offsets()["n"]
    return h


def _fbuf(a):
    """Device copy of a float32 array (via a float64 staging of the bits)."""
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    if a.size % 2:
        a = np.concatenate([a, np.zeros(1, np.float32)])
    return dbuf(a.view(np.float64).copy())


def _fhost(buf, n):
    return host(buf).view(np.float32)[:n].copy()


@pytest.mark.parametrize("dims", [(16, 16, 8), (24, 12, 8)], ids=str)
def test_unet_forward_and_gradients_match_oracle(dims):
    nx, ny, nz = dims
    n = nx * ny * nz
    L = fr.lib()
    print(f"unet parity {dims}", flush=True)
    p = make_params(7)
    h = _net(1)
    fr.check(L.fr_unet_set_params(h, p.ctypes.data_as(C.c_void_p)))
    rng = np.random.default_rng(8)
    x = rng.standard_normal((n, 3)).astype(np.float32)
    gy = rng.standard_normal(n).astype(np.float32)
    dx, dy, dg = _fbuf(x), _fbuf(np.zeros(n, np.float32)), _fbuf(gy)
    fr.check(L.fr_unet_forward(h, nx, ny, nz, dptr(dx), dptr(dy), None), "forward")
    fr.check(L.fr_unet_zero_grads(h, None))
    fr.check(L.fr_unet_backward(h, nx, ny, nz, dptr(dg), None), "backward")
    y = _fhost(dy, n)
    g = np.zeros(offsets()["n"], np.float32)
    fr.check(L.fr_unet_get_grads(h, g.ctypes.data_as(C.c_void_p)))
    L.fr_unet_destroy(h)
    yo = np.zeros(n)
    go = np.zeros(offsets()["n"])
    assert OL.orc().orc_unet_grad(nx, ny, nz, p, 1, x.astype(np.float64),
                                  gy.astype(np.float64), yo, go) == 0
    ey = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("output rel", ey)
    assert ey < 2e-2
    worst = 0.0
    for a, b in offsets()["blocks"]:
        if np.linalg.norm(go[a:b]) == 0:
            assert np.abs(g[a:b]).max() == 0
            continue
        e = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        worst = max(worst, e)
        assert e < 5e-2, (a, b, e)
    print("worst gradient block rel", worst)


def _duct(tmp, tag):
    from first_repo_amd.configs import DuctWorkload
    d = str(tmp / tag)
    w = DuctWorkload(n=(64, 32, 32), inject_per_step=200, capacity=20000, rho_ratio=3.0, seed=5,
                     extra={"Solver": 2})
    w.build(d)
    sim = fr.Sim(d + "/params.in", d)
    sim.reserve_particles(w.capacity)
    sim.initialize()
    sim.set_field("Rho", w.density())
    sim.call("Finish_Initialize")
    return w, sim


def _run(w, sim, steps):
    flag0, its = 0, []
    for s in range(steps):
        inlet = w.inlet_particles(s, flag0)
        flag0 += len(inlet)
        sim.set_inflow(w.inflow(s), inlet)
        sim.step()
        its.append(sim.timing()[3])
    return np.array(its)


def test_unet_training_fits_a_learnable_target():
    """The GPU trainer (forward, backward, Adam on the BF16 matrix-core
    network) fits a local target: y = 0.5 * (input channel 0) on smooth
    random fields, the loss falls by more than 10x in 300 steps."""
    L = fr.lib()
    nx = ny = nz = 16
    n = nx * ny * nz
    rng = np.random.default_rng(0)
    net = _net(3)

    def smooth(a):
        for ax in range(3):
            a = (np.roll(a, 1, ax) + 2 * a + np.roll(a, -1, ax)) / 4
        return a

    losses = []
    for it in range(300):
        f = smooth(smooth(rng.standard_normal((nz, ny, nx))))
        f /= f.std()
        x = np.zeros((n, 3), np.float32)
        x[:, 0] = f.reshape(-1)
        x[:, 1] = 1.0
        t = 0.5 * f.reshape(-1)
        dx, dy = _fbuf(x), _fbuf(np.zeros(n, np.float32))
        fr.check(L.fr_unet_forward(net, nx, ny, nz, dptr(dx), dptr(dy), None))
        e = _fhost(dy, n) - t
        losses.append(float((e * e).mean()))
        fr.check(L.fr_unet_zero_grads(net, None))
        fr.check(L.fr_unet_backward(net, nx, ny, nz, dptr(_fbuf((2 * e / n).astype(np.float32))),
                                    None))
        fr.check(L.fr_unet_adam_step(net, C.c_float(1e-3), None))
    L.fr_unet_destroy(net)
    first, last = np.mean(losses[:10]), np.mean(losses[-10:])
    print("loss", first, "->", last)
    assert last < 0.1 * first


def test_learned_pcg_on_the_duct(tmp_path):
    """Ablation on the duct's variable-coefficient solve (Solver = 2,
    64x32x32, density ratio 3): pairs collected from the solver's own run
    train the network on the GPU (full-domain samples); the learned initial
    iterate (mode 1) and the learned flexible preconditioner (mode 2) are
    then run on the same steps as the plain multigrid PCG.  Both must reach
    the same 1e-10 relative residual (the step test and the fallback keep
    them safe); the iteration counts of all three are printed.  On this
    duct the increment is dominated by the x-direction (Neumann, inflow /
    outflow) long-wave part that the network's receptive field does not
    span, and the measured gain is within +-1% of plain."""
    L = fr.lib()
    print("learned PCG: collect", flush=True)
    w, sim = _duct(tmp_path, "collect")
    data = C.c_void_p()
    fr.check(L.fr_unet_data_create(64, 32, 32, 64, C.byref(data)))
    fr.check(L.fr_sim_set_collect(sim.h, data))
    plain = _run(w, sim, 12)
    cnt = C.c_int()
    L.fr_unet_data_count(data, C.byref(cnt))
    assert cnt.value == 48
    sim.close()
    net = _net(3)
    l0, l1 = C.c_double(), C.c_double()
    print("learned PCG: train", flush=True)
    fr.check(L.fr_unet_train(net, data, 400, 2, 64, 32, 32, C.c_float(3e-4), 9, C.byref(l0),
                             C.byref(l1), None), "train")
    print("loss", l0.value, "->", l1.value, flush=True)
    assert np.isfinite(l1.value)
    res = {}
    for mode in (1, 2):
        print(f"learned PCG: mode {mode}", flush=True)
        w2, s2 = _duct(tmp_path, f"mode{mode}")
        fr.check(L.fr_sim_set_learned(s2.h, net, mode))
        res[mode] = _run(w2, s2, 12)
        kept, fb, gain = C.c_int(), C.c_int(), C.c_double()
        fr.check(L.fr_sim_learned_stats(s2.h, C.byref(kept), C.byref(fb), C.byref(gain)))
        print(f"mode {mode}: guesses applied {kept.value}, fallbacks {fb.value}, mean log10 "
              f"residual^2 gain {gain.value / max(kept.value, 1):.3f}", flush=True)
        s2.close()
    print("iterations per step: plain", plain.tolist(), "mode1", res[1].tolist(), "mode2",
          res[2].tolist())
    L.fr_unet_data_destroy(data)
    L.fr_unet_destroy(net)
    for mode in (1, 2):
        assert res[mode].sum() <= 1.02 * plain.sum(), (mode, res[mode].sum(), plain.sum())
