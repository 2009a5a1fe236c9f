"""Writes tests/golden/unet_torch_fp64.npz: the U-Net of include/fr_amd.h
evaluated by torch (CPU, float64: conv3d, avg_pool3d, conv_transpose3d,
relu, autograd) on seeded parameters and inputs, to pin the oracle's
restatement (tests/test_oracle_unet.py).

The parameters are regenerated by the test from the stored seed (numpy
PCG64 standard normals scaled per layer); their checksum is stored too.
Stored: output y, the gradient of sum(gy * y) subsampled every 37th entry,
and per-layer gradient norms.  Run from the repository root:
    python tests/golden/make_unet_fixture.py
"""
import os
import sys

import numpy as np
import torch
import torch.nn.functional as F

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from unet_layout import CONV, UP, offsets, make_params  # noqa: E402

SEED, NX, NY, NZ = 20261203, 8, 8, 4


def torch_unet(p, x):
    """p: flat float64 tensor (requires grad); x: [n][3] -> y [n]."""
    off = offsets()
    t = torch.zeros(1, 16, NZ, NY, NX, dtype=torch.float64)
    t[0, :3] = x.reshape(NZ, NY, NX, 3).permute(3, 0, 1, 2)

    def conv(l, h):
        ci, co = CONV[l]
        W = p[off["cw"][l]:off["cw"][l] + 27 * ci * co].reshape(3, 3, 3, ci, co)
        b = p[off["cb"][l]:off["cb"][l] + co]
        return F.relu(F.conv3d(h, W.permute(4, 3, 0, 1, 2), b, padding=1))

    def up(k, h):
        ci, co = UP[k]
        W = p[off["uw"][k]:off["uw"][k] + 8 * ci * co].reshape(2, 2, 2, ci, co)
        b = p[off["ub"][k]:off["ub"][k] + co]
        return F.conv_transpose3d(h, W.permute(3, 4, 0, 1, 2), b, stride=2)

    e1 = conv(1, conv(0, t))
    e2 = conv(3, conv(2, F.avg_pool3d(e1, 2)))
    b2 = conv(5, conv(4, F.avg_pool3d(e2, 2)))
    d2 = conv(7, conv(6, torch.cat([up(0, b2), e2], 1)))
    d1 = conv(9, conv(8, torch.cat([up(1, d2), e1], 1)))
    w = p[off["ow"]:off["ow"] + 16].reshape(1, 16, 1, 1, 1)
    y = F.conv3d(d1, w, p[off["ob"]:off["ob"] + 1])
    return y.reshape(-1)


def main():
    params = make_params(SEED)
    rng = np.random.default_rng(SEED + 1)
    n = NX * NY * NZ
    x = rng.standard_normal((n, 3))
    gy = rng.standard_normal(n)
    p = torch.tensor(params.astype(np.float64), requires_grad=True)
    y = torch_unet(p, torch.tensor(x))
    (y * torch.tensor(gy)).sum().backward()
    g = p.grad.numpy()
    off = offsets()
    norms = np.array([np.linalg.norm(g[a:b]) for a, b in off["blocks"]])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "unet_torch_fp64.npz")
    np.savez_compressed(out, seed=SEED, dims=np.array([NX, NY, NZ]),
                        param_sum=np.float64(params.astype(np.float64).sum()),
                        x=x, gy=gy, y=y.detach().numpy(), grad_sub=g[::37], grad_norms=norms)
    print("wrote", out, "params", params.size)


if __name__ == "__main__":
    main()
