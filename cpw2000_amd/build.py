"""Build recipe for the native libraries of cpw2000_amd (in-tree).

  libcpwgpu.so          HIP kernels + C ABI (include/cpwgpu.h), gfx950
  libcpwgpu_fortran.so  Fortran drop-ins hk_psi_c16, hk_psi_spin_c16 and ditsp_c16, and the all-k
                        hk_psi_allk_c16, over the C ABI (flang)
  libcpwprobe.so        VALU issue-rate probe (measurement only: ``bench.py --issue-rates``)

Run ``python -m cpw2000_amd.build``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libcpwgpu.so")
FLIB = os.path.join(PKG, "libcpwgpu_fortran.so")
PLIB = os.path.join(PKG, "libcpwprobe.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang")
ARCH = os.environ.get("CPWGPU_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    srcs = [os.path.join(CSRC, "cpwgpu.hip")]
    tmp = LIB + ".tmp"
    # rocSOLVER (zheevd, zpotrf) and rocBLAS (ztrsm) for the small dense
    # eigenproblems of the iterative diagonalisation
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs,
          "-lrocsolver", "-lrocblas"])
    os.replace(tmp, LIB)


def build_probe():
    tmp = PLIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, os.path.join(CSRC, "issue_rates.hip")])
    os.replace(tmp, PLIB)


def build_fortran():
    srcs = [os.path.join(PKG, "fortran", f)
            for f in ("hk_psi_c16.f90", "hk_psi_spin_c16.f90", "ditsp_c16.f90", "hk_psi_allk_c16.f90",
                      "proj_nl_kb_c16.f90")]
    tmp = FLIB + ".tmp"
    objdir = os.path.join(PKG, "fortran")
    _run([FLANG, "-O2", "-fPIC", "-shared", "-module-dir", objdir, "-o", tmp, *srcs,
          "-L", PKG, "-lcpwgpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FLIB)


def build(fortran: bool = True):
    build_hip()
    build_probe()
    if fortran:
        build_fortran()


if __name__ == "__main__":
    build(fortran="--no-fortran" not in sys.argv)
