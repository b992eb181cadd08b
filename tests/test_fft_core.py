"""The register DFTs and two-pass line transforms of the specialised HIP
kernels (cpw2000_amd/csrc/fft_core.hpp), compiled for the host with g++ and
checked against a long-double DFT (tests/fft_core_check.cpp)."""

import os
import subprocess

from tests.conftest import ROOT


def test_fft_core_matches_dft(tmp_path):
    exe = os.path.join(tmp_path, "fft_core_check")
    src = os.path.join(ROOT, "tests", "fft_core_check.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-o", exe, src], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    err = float(out.stdout.split("max_rel_err")[1])
    assert err < 1e-14
