"""Builds libmom6hip.so (every HIP source under csrc/) for gfx950 with hipcc,
and the Fortran host side (fortran/*.F90: the iso_c_binding modules and the
rk2_hip_driver program) with amdflang into bin/rk2_hip_driver.

Flags: -ffp-contract=off keeps the reference's FP64 rounding (no FMA
contraction); packed FP32 instructions are disabled; no fast-math.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmom6hip.so")
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(HERE, "bin", "rk2_hip_driver")
ARCH = os.environ.get("MOM6H_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(verbose: bool = False) -> str:
    """Compile the library in place (always recompiles)."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, *sources()]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran modules and driver against libmom6hip.so (run
    build() first).  Returns the driver path, or None without amdflang."""
    fc = os.environ.get("FC_HIP", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        return None
    mod_dir = os.path.join(FORTRAN, "_build")
    os.makedirs(mod_dir, exist_ok=True)
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    srcs = [os.path.join(FORTRAN, "mom6_hip.F90"), os.path.join(FORTRAN, "rk2_hip_driver.F90")]
    tmp = DRIVER + ".tmp"
    cmd = [fc, "-O2", "-module-dir", mod_dir, *srcs, "-L" + HERE, "-lmom6hip",
           "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=mod_dir)
    if res.returncode != 0:
        raise RuntimeError(f"amdflang failed ({res.returncode}):\n{res.stderr[-6000:]}")
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
