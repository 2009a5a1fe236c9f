"""Builds first_repo_amd/libfr_amd.so (gfx950 kernels + host classes + C ABI)
with hipcc, and the C++ test programs of tests/cpp against it.  Usage:
python -m first_repo_amd.build"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FR_OFFLOAD_ARCH", "gfx950")

KERNELS = ["fields.hip", "box.hip", "duct.hip", "fft_poisson.hip", "duct_poisson.hip",
           "mg_pcg.hip", "unet.hip", "particles.hip", "diag.hip", "stats.hip"]
HOST = ["host.cpp", "host_io.cpp", "members.cpp"]

# -ffp-contract=off: the FP64 path reproduces the reference's unfused
# multiply-add sequences.  The packed-FP32 feature is disabled for all device
# code.
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value", "-I" + os.path.join(ROOT, "include")]

LIB = os.path.join(HERE, "libfr_amd.so")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out)
    if r.returncode:
        raise RuntimeError(f"build failed: {cmd[-1]}")


def build(jobs: int = 8) -> str:
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", f) for f in KERNELS] + \
           [os.path.join(HERE, "host", f) for f in HOST]
    procs, objs = [], []
    for s in srcs:
        o = os.path.join(obj_dir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", s, "-o", o]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print("\n".join(lines))
        if p.returncode:
            failed = True
            print("FAILED:", " ".join(cmd))
    if failed:
        raise RuntimeError("hipcc compile failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    build_test_programs()
    return LIB


# C++ programs that drive the library through its host classes (tests/cpp),
# linked against libfr_amd.so and found beside it at run time.
TEST_PROGRAMS = ["member_substep.cpp"]


def build_test_programs() -> None:
    src_dir = os.path.join(ROOT, "tests", "cpp")
    out_dir = os.path.join(src_dir, "_build")
    os.makedirs(out_dir, exist_ok=True)
    for f in TEST_PROGRAMS:
        exe = os.path.join(out_dir, os.path.splitext(f)[0])
        _run([HIPCC, *[x for x in FLAGS if x != "-fPIC"], "-x", "hip", os.path.join(src_dir, f),
              "-x", "none", LIB, "-Wl,-rpath,$ORIGIN/../../../first_repo_amd", "-o", exe])


if __name__ == "__main__":
    build()
    sys.exit(0)
