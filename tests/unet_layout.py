"""Parameter layout of the U-Net of include/fr_amd.h (fr_unet_*), shared by
the fixture generator and the tests."""
import numpy as np

CONV = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64),
        (64, 64), (64, 32), (32, 32), (32, 16), (16, 16)]
UP = [(64, 32), (32, 16)]


def offsets():
    o, cw, cb, uw, ub, blocks = 0, [], [], [], [], []
    for ci, co in CONV:
        cw.append(o); blocks.append((o, o + 27 * ci * co)); o += 27 * ci * co
        cb.append(o); blocks.append((o, o + co)); o += co
    for ci, co in UP:
        uw.append(o); blocks.append((o, o + 8 * ci * co)); o += 8 * ci * co
        ub.append(o); blocks.append((o, o + co)); o += co
    ow = o; blocks.append((o, o + 16)); o += 16
    ob = o; blocks.append((o, o + 1)); o += 1
    return {"cw": cw, "cb": cb, "uw": uw, "ub": ub, "ow": ow, "ob": ob, "n": o,
            "blocks": blocks}


def make_params(seed):
    """Seeded parameters (float32): He-scaled weights, small biases; the
    unused input channels 3..15 of the first convolution are zero."""
    off = offsets()
    rng = np.random.default_rng(seed)
    p = np.zeros(off["n"], np.float32)
    for l, (ci, co) in enumerate(CONV):
        real = 3 if l == 0 else ci
        W = rng.standard_normal((27, ci, co)) * np.sqrt(2.0 / (27 * real))
        if l == 0:
            W[:, 3:, :] = 0
        p[off["cw"][l]:off["cw"][l] + W.size] = W.reshape(-1)
        p[off["cb"][l]:off["cb"][l] + co] = 0.1 * rng.standard_normal(co)
    for k, (ci, co) in enumerate(UP):
        p[off["uw"][k]:off["uw"][k] + 8 * ci * co] = rng.standard_normal(8 * ci * co) / np.sqrt(ci)
        p[off["ub"][k]:off["ub"][k] + co] = 0.1 * rng.standard_normal(co)
    p[off["ow"]:off["ow"] + 16] = 0.5 * rng.standard_normal(16)
    p[off["ob"]] = 0.1
    return p
