"""Two-component H|psi> on the GPU (cpwgpu_plan_apply_spin and the drop-in
cpwgpu_hk_psi_spin_c16) against the spinor oracle (oracle/spin.py, pinned
to the scalar oracle, a dense G-space spinor H and the Pauli form).

FP64: relative 2-norm error <= 1e-12; FP32-mixed: <= 2e-6.  nsp = 1, 2, 4
on the 2-atom cell (24^3) through the device API, nsp = 4 through the
host drop-in and on the 64-atom cell (72^3); nsp = 4 with both coupled plane
kernels (LDS two-pass and register lines) in both precisions."""

import numpy as np
import pytest

from cpw2000_amd import synth
from oracle import spin

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 2e-6


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def hk():
    from cpw2000_amd import hk_psi

    return hk_psi


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _plan(hk, d, s, nsp):
    plan = hk.Plan.from_inputs(d, projectors=False)
    plan.set_potential_spin(s["vscr_sp"], nsp)
    plan.set_projectors_spin(s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    return plan


def _spin_kernels(hk, plan, nsp):
    """Coupled plane kernels to run: both for nsp = 4 (the register one is
    built on these meshes), the default otherwise."""
    if nsp != 4:
        return [-1]
    for prec in (hk.PREC_FP64, hk.PREC_FP32):
        plan.set_variant(prec, hk.TUNE_SPIN, 1)
        assert plan.get_variant(prec, hk.TUNE_SPIN) == 1
    return [0, 1]


@pytest.mark.parametrize("nsp", [1, 2, 4])
def test_apply_spin_matches_oracle(torch_dev, hk, nsp):
    torch = torch_dev
    d = synth.make_config(0, nbands=2)
    s = synth.spin_inputs(d, 6)
    mtxd = d["mtxd"]
    ref = spin.hk_psi_spin(d, s["psi_sp"], s["vscr_sp"], nsp, s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    plan = _plan(hk, d, s, nsp)
    pt = torch.from_numpy(np.ascontiguousarray(s["psi_sp"][: 2 * mtxd].T)).to("cuda")
    for kern in _spin_kernels(hk, plan, nsp):
        for prec, tol in ((hk.PREC_FP64, TOL64), (hk.PREC_FP32, TOL32)):
            plan.set_variant(prec, hk.TUNE_SPIN, kern)
            ht = torch.zeros_like(pt)
            plan.apply_spin(pt, ht, prec)
            torch.cuda.synchronize()
            assert _rel(ht.cpu().numpy().T, ref) < tol, (nsp, kern, prec)
    # without projectors: local + kinetic only
    plan.set_projectors_spin(s["anlsp"], s["xnlkbsp"], 0)
    ref0 = spin.hk_psi_spin(d, s["psi_sp"], s["vscr_sp"], nsp)
    ht = torch.zeros_like(pt)
    plan.apply_spin(pt, ht, hk.PREC_FP64)
    torch.cuda.synchronize()
    assert _rel(ht.cpu().numpy().T, ref0) < TOL64
    plan.close()


def test_tune_spin_keeps_results(torch_dev, hk):
    """The two-component tuner picks valid kernels and leaves H psi of the
    chosen configuration in hpsi; later applies agree with the oracle."""
    torch = torch_dev
    d = synth.make_config(0, nbands=2)
    s = synth.spin_inputs(d, 6)
    mtxd = d["mtxd"]
    ref = spin.hk_psi_spin(d, s["psi_sp"], s["vscr_sp"], 4, s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    plan = _plan(hk, d, s, 4)
    pt = torch.from_numpy(np.ascontiguousarray(s["psi_sp"][: 2 * mtxd].T)).to("cuda")
    for prec, tol in ((hk.PREC_FP64, TOL64), (hk.PREC_FP32, TOL32)):
        ht = torch.zeros_like(pt)
        choice = plan.tune_spin(pt, ht, prec)
        torch.cuda.synchronize()
        assert choice["spin"] in (0, 1) and 0 <= choice["proj"] <= 5 and 0 <= choice["back"] <= 4
        assert choice["splitk"] in (128, 256, 512, 1024)
        assert _rel(ht.cpu().numpy().T, ref) < tol, (prec, choice)
        h2 = torch.zeros_like(pt)
        plan.apply_spin(pt, h2, prec)
        torch.cuda.synchronize()
        assert _rel(h2.cpu().numpy().T, ref) < tol, (prec, choice)
    # the potential must be set first
    plain = hk.Plan.from_inputs(d)
    with pytest.raises(hk.CpwGpuError):
        plain.tune_spin(pt, torch.zeros_like(pt), hk.PREC_FP32)
    plain.close()
    plan.close()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32mixed"])
def test_hk_psi_spin_c16_drop_in(hk, prec):
    """The reference's argument list, host arrays, mxddim > mtxd rows."""
    d = synth.make_config(0, nbands=2, mxddim_pad=5)
    s = synth.spin_inputs(d, 4)
    ref = spin.hk_psi_spin(d, s["psi_sp"], s["vscr_sp"], 4, s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    hpsi = np.full_like(s["psi_sp"], 3.0 + 0j, order="F")
    hk.set_precision(prec)
    try:
        hk.hk_psi_spin_c16(d["mtxd"], 4, s["psi_sp"], hpsi, True, d["ng"],
                           np.asfortranarray(d["kgv"], np.int32), d["ekpg"],
                           np.asfortranarray(d["isort"], np.int32), s["vscr_sp"],
                           np.asfortranarray(d["kmscr"], np.int32), 4, s["anlsp"], s["xnlkbsp"],
                           s["nanlsp"], d["mxddim"], 4, s["mxdasp"], d["mxdgve"], d["mxdscr"],
                           s["mxdnsp"])
    finally:
        hk.set_precision(hk.PREC_FP64)
    mtxd = d["mtxd"]
    assert _rel(hpsi[: 2 * mtxd], ref) < (TOL64 if prec == 0 else TOL32)
    assert np.all(hpsi[2 * mtxd:] == 3.0)  # rows beyond 2*mtxd untouched


def test_noncollinear_si64(torch_dev, hk):
    torch = torch_dev
    d = synth.make_config(1, nbands=2)
    s = synth.spin_inputs(d, 4)
    mtxd = d["mtxd"]
    ref = spin.hk_psi_spin(d, s["psi_sp"], s["vscr_sp"], 4, s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    plan = _plan(hk, d, s, 4)
    pt = torch.from_numpy(np.ascontiguousarray(s["psi_sp"][: 2 * mtxd].T)).to("cuda")
    for kern in _spin_kernels(hk, plan, 4):
        for prec, tol in ((hk.PREC_FP64, TOL64), (hk.PREC_FP32, TOL32)):
            plan.set_variant(prec, hk.TUNE_SPIN, kern)
            ht = torch.zeros_like(pt)
            plan.apply_spin(pt, ht, prec)
            torch.cuda.synchronize()
            assert _rel(ht.cpu().numpy().T, ref) < tol, (kern, prec)
    plan.close()


def test_fortran_shim_hk_psi_spin_c16(hk):
    """The Fortran drop-in hk_psi_spin_c16_ (libcpwgpu_fortran.so), by reference."""
    import ctypes as C

    lib = C.CDLL(hk.FLIB_PATH)
    d = synth.make_config(0, nbands=2, mxddim_pad=2)
    s = synth.spin_inputs(d, 3)
    ref = spin.hk_psi_spin(d, s["psi_sp"], s["vscr_sp"], 4, s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    hpsi = np.zeros_like(s["psi_sp"], order="F")
    i = lambda v: C.byref(C.c_int(int(v)))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    kgv = np.asfortranarray(d["kgv"], np.int32)
    lnew = C.c_int(1)
    lib.hk_psi_spin_c16_(i(d["mtxd"]), i(3), p(s["psi_sp"]), p(hpsi), C.byref(lnew), i(d["ng"]), p(kgv),
                         p(d["ekpg"]), p(d["isort"]), p(s["vscr_sp"]), p(d["kmscr"]), i(4), p(s["anlsp"]),
                         p(s["xnlkbsp"]), i(s["nanlsp"]), i(d["mxddim"]), i(3), i(s["mxdasp"]), i(d["mxdgve"]),
                         i(d["mxdscr"]), i(s["mxdnsp"]))
    assert lnew.value == 1
    assert _rel(hpsi[: 2 * d["mtxd"]], ref) < TOL64


def _spin_golden():
    """Inputs and the reference's own hk_psi_spin_c16 outputs (flang build of
    Src/hk_psi/reference/hk_psi_spin_c16.f90, nsp = 1, 2, 4; golden vectors
    tests/golden/hkpsi_spin_cfg0.npz)."""
    import json
    import os

    from tests.conftest import GOLDEN
    from tests.golden.make_golden import digest

    z = np.load(os.path.join(GOLDEN, "hkpsi_spin_cfg0.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    d = synth.make_config(0, nbands=2)
    dg = digest(d)
    for k, v in meta["digest"].items():
        np.testing.assert_allclose(dg[k], v, rtol=1e-12, atol=1e-12, err_msg=f"generator drifted ({k})")
    s = synth.spin_inputs(d, meta["nbands"], seed=meta["seed"])
    return d, s, meta, z


@pytest.mark.parametrize("nsp", [1, 2, 4])
def test_spin_matches_reference_golden(torch_dev, hk, nsp):
    """Device plan (every coupled kernel for nsp = 4) and host drop-in against
    the reference's output, FP64 to 1e-12 and FP32-mixed to 2e-6."""
    torch = torch_dev
    d, s, meta, z = _spin_golden()
    ref = z[f"hpsi_nsp{nsp}"]
    mtxd, nb = d["mtxd"], meta["nbands"]
    plan = _plan(hk, d, s, nsp)
    pt = torch.from_numpy(np.ascontiguousarray(s["psi_sp"][: 2 * mtxd].T)).to("cuda")
    for kern in _spin_kernels(hk, plan, nsp):
        for prec, tol in ((hk.PREC_FP64, TOL64), (hk.PREC_FP32, TOL32)):
            plan.set_variant(prec, hk.TUNE_SPIN, kern)
            ht = torch.zeros_like(pt)
            plan.apply_spin(pt, ht, prec)
            torch.cuda.synchronize()
            assert _rel(ht.cpu().numpy().T, ref) < tol, (nsp, kern, prec)
    plan.close()
    hpsi = np.zeros_like(s["psi_sp"], order="F")
    hk.hk_psi_spin_c16(mtxd, nb, s["psi_sp"], hpsi, True, d["ng"], np.asfortranarray(d["kgv"], np.int32),
                       d["ekpg"], np.asfortranarray(d["isort"], np.int32), s["vscr_sp"],
                       np.asfortranarray(d["kmscr"], np.int32), nsp, s["anlsp"], s["xnlkbsp"], s["nanlsp"],
                       d["mxddim"], nb, s["mxdasp"], d["mxdgve"], d["mxdscr"], s["mxdnsp"])
    err = _rel(hpsi[: 2 * mtxd], ref)
    print(f"hk_psi_spin_c16 nsp {nsp}: drop-in FP64 vs reference {err:.2e}")
    assert err < TOL64
