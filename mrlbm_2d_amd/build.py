"""Builds libmrlbm.so (HIP, gfx950) in-tree: mrlbm_2d_amd/_lib/libmrlbm.so."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "libmrlbm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["mrlbm_kernels.hip", "mrlbm_capi.hip", "mrlbm_adjoint.hip", "mrlbm_batch.hip", "mrlbm_geometry.cpp"]
# FP64 must not contract mul+add into FMA: the FP64 path is bit-compared with the
# reference, which is built without FMA.  Packed FP32 instructions are disabled.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas"]


FORTRAN_DIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(OUT_DIR, "mrlbm_gpu")
FC = os.environ.get("FC_MRLBM", "flang")


def build_driver(verbose: bool = False) -> str:
    """Fortran drop-in driver (iso_c_binding) linked against libmrlbm.so."""
    bdir = os.path.join(FORTRAN_DIR, "_build")
    os.makedirs(bdir, exist_ok=True)
    iface_o = os.path.join(bdir, "mrlbm_iface.o")
    steps = [
        [FC, "-O2", "-J", bdir, "-c", os.path.join(FORTRAN_DIR, "mrlbm_iface.f90"), "-o", iface_o],
        [FC, "-O2", "-J", bdir, os.path.join(FORTRAN_DIR, "mrlbm_driver.f90"), iface_o,
         "-L", OUT_DIR, "-lmrlbm", "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed:\n{r.stdout}{r.stderr}")
        if verbose and (r.stdout + r.stderr).strip():
            print(r.stdout + r.stderr, file=sys.stderr)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", "-I", os.path.join(REPO, "include"), "-I", CSRC, "-o", tmp,
           *[os.path.join(CSRC, s) for s in SOURCES]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{msgs}")
    if verbose and msgs.strip():
        print(msgs, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


def build_all(verbose: bool = False):
    return build(verbose), build_driver(verbose)


if __name__ == "__main__":
    print(build_all(verbose=True))
