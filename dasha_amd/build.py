"""Build recipe for the HIP step library (gfx950) and the Fortran driver.

``python -m dasha_amd.build`` compiles ``csrc/mhd_gpu.hip`` into
``dasha_amd/libmhd_gpu.so`` in-tree with hipcc, and, when AMD flang is
present, the ISO_C_BINDING driver in ``fortran/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("MHD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-std=c++17",
    # rounding follows the written expression order (no FMA contraction)
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build_hip(verbose: bool = False) -> str:
    src = os.path.join(HERE, "csrc", "mhd_gpu.hip")
    out = os.path.join(HERE, "libmhd_gpu.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), src, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for mhd_gpu.hip")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING module and driver (skipped if flang is absent)."""
    if not (os.path.exists(FLANG) or shutil.which("flang")):
        return None
    fc = FLANG if os.path.exists(FLANG) else shutil.which("flang")
    fdir = os.path.join(HERE, "fortran")
    mod = os.path.join(fdir, "mhd_gpu_mod.f90")
    io = os.path.join(fdir, "mhd_io.f90")
    drv = os.path.join(fdir, "mhd_driver.f90")
    if not (os.path.exists(mod) and os.path.exists(io) and os.path.exists(drv)):
        return None
    bdir = os.path.join(fdir, "build")
    os.makedirs(bdir, exist_ok=True)
    exe = os.path.join(fdir, "mhd_driver")
    cmd = [fc, "-O2", "-J", bdir, mod, io, drv, "-o", exe, f"-L{HERE}", "-lmhd_gpu",
           f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("flang failed for the Fortran driver")
    return exe


def main():
    print(build_hip(verbose=True))
    exe = build_fortran()
    if exe:
        print(exe)


if __name__ == "__main__":
    main()
