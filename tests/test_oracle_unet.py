"""Pins of the oracle's U-Net (oracle/src/unet.cpp), the checker of the GPU
network: against torch's CPU float64 conv3d / avg_pool3d /
conv_transpose3d / autograd (the committed fixture written by
tests/golden/make_unet_fixture.py), against finite differences of its own
FP64 forward pass, and the BF16 emulation against the exact network."""
import os

import numpy as np
import pytest

import oracle_lib as OL
from unet_layout import offsets, make_params

FIX = os.path.join(os.path.dirname(__file__), "golden", "unet_torch_fp64.npz")


def _grad(nx, ny, nz, p, x, gy, bf16=0):
    O = OL.orc()
    n = nx * ny * nz
    y = np.zeros(n)
    g = np.zeros(offsets()["n"])
    assert O.orc_unet_grad(nx, ny, nz, p, bf16, np.ascontiguousarray(x), gy, y, g) == 0
    return y, g


def test_layout_matches_oracle():
    assert OL.orc().orc_unet_num_params() == offsets()["n"]


def test_oracle_unet_matches_torch_fixture():
    f = np.load(FIX)
    nx, ny, nz = (int(v) for v in f["dims"])
    p = make_params(int(f["seed"]))
    assert p.astype(np.float64).sum() == pytest.approx(float(f["param_sum"]), rel=1e-12)
    y, g = _grad(nx, ny, nz, p, f["x"], f["gy"])
    assert np.abs(y - f["y"]).max() <= 1e-11 * np.abs(f["y"]).max()
    gs = g[::37]
    assert np.abs(gs - f["grad_sub"]).max() <= 1e-10 * np.abs(f["grad_sub"]).max()
    norms = np.array([np.linalg.norm(g[a:b]) for a, b in offsets()["blocks"]])
    np.testing.assert_allclose(norms, f["grad_norms"], rtol=1e-10, atol=1e-14)


def test_oracle_conv3_matches_direct_sum():
    """One convolution against a literal triple loop over taps (zero padding,
    tap = (dz+1)*9 + (dy+1)*3 + (dx+1), W[tap][ci][co])."""
    rng = np.random.default_rng(5)
    nx, ny, nz, ci, co = 5, 4, 3, 2, 3
    W = rng.standard_normal((27, ci, co))
    b = rng.standard_normal(co)
    x = rng.standard_normal((nz, ny, nx, ci))
    y = np.zeros((nz, ny, nx, co))
    assert OL.orc().orc_conv3(nx, ny, nz, ci, co, W.reshape(-1).copy(), b, x.reshape(-1).copy(),
                              y.reshape(-1)) == 0
    ref = np.zeros_like(y) + b
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                t = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                sh = xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
                ref += sh @ W[t]
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_oracle_unet_gradient_finite_differences():
    """Central differences of L = sum(gy * y) in FP64 for parameters of every
    block kind (3x3x3 weight, bias, up-convolution, output)."""
    nx, ny, nz = 8, 4, 4
    p = make_params(11)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((nx * ny * nz, 3))
    gy = rng.standard_normal(nx * ny * nz)
    _, g = _grad(nx, ny, nz, p, x, gy)
    off = offsets()
    O = OL.orc()

    def L(q):
        y = np.zeros(nx * ny * nz)
        O.orc_unet_forward(nx, ny, nz, q, 0, np.ascontiguousarray(x), y)
        return float(gy @ y)

    picks = [off["cw"][0] + 5, off["cw"][4] + 1234, off["cw"][9] + 77, off["cb"][3] + 2,
             off["uw"][0] + 999, off["uw"][1] + 321, off["ub"][1] + 3, off["ow"] + 7, off["ob"]]
    for i in picks:
        e = np.float32(2.0 ** -18)
        qp, qm = p.copy(), p.copy()
        qp[i] += e
        qm[i] -= e
        fd = (L(qp) - L(qm)) / (float(qp[i]) - float(qm[i]))
        assert fd == pytest.approx(g[i], rel=2e-3, abs=1e-6 * np.abs(g).max()), i


def test_bf16_emulation_close_to_exact():
    nx, ny, nz = 8, 8, 4
    p = make_params(3)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((nx * ny * nz, 3))
    y0 = np.zeros(nx * ny * nz)
    y1 = np.zeros(nx * ny * nz)
    O = OL.orc()
    O.orc_unet_forward(nx, ny, nz, p, 0, x, y0)
    O.orc_unet_forward(nx, ny, nz, p, 1, x, y1)
    e = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert 1e-5 < e < 3e-2, e
