"""Atomic forces on the GPU (adjoint of the local and non-local terms) against
the force oracle (restatement of for_str_nl_kb / for_str_local_force),
against central finite differences of the GPU FP64 energy (h = 1e-4 bohr,
tolerance 1e-5 Ha/bohr), and through torch.autograd.gradcheck."""

import numpy as np
import pytest

from cpw2000_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


def _psi_t(torch, d):
    return torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")


def _gpu_forces(torch, d):
    from cpw2000_amd import hk_psi

    plan = hk_psi.Plan.from_inputs(d)
    f, enl = plan.forces(_psi_t(torch, d), d["rat"], d["avec"], d["proj_atom"], d["floc_q"],
                         d["floc_f"], d["species"])
    plan.close()
    return f, enl


# (workload, kmscr(1:3) override): 24^3 and 80^3 run the register-line density
# pass with per-plane partial transforms, 96^3 (the 2-atom box enlarged) the LDS
# density pass with per-group partial transforms
@pytest.mark.parametrize("cfg,kmax", [("si2f", None), (4, None), ("si2f", 47)])
def test_forces_match_oracle(torch_dev, cfg, kmax):
    from oracle import forces as Fo

    d = synth.make_config(cfg, kmax_override=None if kmax is None else [kmax] * 3)
    f, enl = _gpu_forces(torch_dev, d)
    fo, enlo = Fo.forces(d, d["psi"])
    assert np.abs(fo).max() > 1e-4
    assert np.abs(f - fo).max() < 1e-9
    assert abs(enl - enlo) < 1e-10 * max(1.0, abs(enlo))


def test_forces_match_finite_differences_sige64(torch_dev):
    from cpw2000_amd.autograd import _energy_and_plan

    torch = torch_dev
    d = synth.make_config(4)
    f, _ = _gpu_forces(torch, d)
    psi_t = _psi_t(torch, d)
    h = 1e-4
    spec = synth.config_spec(4)
    pos0 = [t[1] for t in spec.types]
    for atom, sp, loc in ((5, 0, 5), (40, 1, 8)):     # one Si, one Ge
        for j in range(3):
            dfrac = np.linalg.solve(d["avec"], h * np.eye(3)[j])
            es = []
            for sgn in (1, -1):
                pos = [p.copy() for p in pos0]
                pos[sp][loc] += sgn * dfrac
                dd = synth.make_config(4, positions=pos, with_psi=False)
                e, plan = _energy_and_plan(dd, psi_t)
                plan.close()
                es.append(e)
            assert abs(f[atom, j] + (es[0] - es[1]) / (2 * h)) < 1e-5


def test_autograd_gradcheck_two_atoms(torch_dev):
    torch = torch_dev
    from cpw2000_amd.autograd import PotentialEnergy

    d0 = synth.make_config("si2f")
    psi_t = _psi_t(torch, d0)
    avec = d0["avec"]

    def builder(tau_cart):
        frac = np.linalg.solve(avec, np.asarray(tau_cart).T).T
        return synth.make_config("si2f", positions=[frac])

    tau = torch.from_numpy(d0["rat"] @ avec.T).requires_grad_(True)
    fn = lambda t: PotentialEnergy.apply(t, builder, psi_t)
    assert torch.autograd.gradcheck(fn, (tau,), eps=1e-4, atol=1e-6, rtol=1e-4)


def test_forces_match_reference_routines(torch_dev):
    """Against cpw2000's own force routines (flang build of the reference,
    golden vectors tests/golden/forces_sige64.npz): for_str_nl_kb for the
    non-local term with the projectors built on the device from the same KB
    tables, for_str_local_force for the local term (one star per q of the
    form factors, the band density as den), 32 bands with occupations 2 ..
    0.5.  The reference's covariant components (per unit of the fractional
    coordinate) become Cartesian through inv(avec)^T, the step of
    Src/cpw_lib/force_stress_kb.f90:163-167 (adotm1 = bdot / 4 pi^2) followed
    by avec.  Tolerance 1e-9 Ha/bohr."""
    import json
    import os

    from cpw2000_amd import hk_psi
    from tests.conftest import GOLDEN
    from tests.golden.make_golden import digest

    torch = torch_dev
    z = np.load(os.path.join(GOLDEN, "forces_sige64.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    d = synth.make_config(4)
    dg = digest(d)
    for k, v in meta["digest"].items():
        np.testing.assert_allclose(dg[k], v, rtol=1e-12, atol=1e-12, err_msg=f"generator drifted ({k})")
    nb, occ = meta["nbands"], z["occ"]
    psi = _psi_t(torch, d)[:nb]
    to_cart = lambda f: f @ np.linalg.inv(d["avec"])
    # non-local term, projectors from the KB tables on the device
    pp = synth.kb_pseudo(synth.config_spec(4))
    plan = hk_psi.Plan.from_inputs(d, projectors=False)
    plan.build_projectors_kb(np.asarray(d["kpt"], float), pp["nanl"], pp["nqnl"], pp["delqnl"], pp["vkb"],
                             pp["nkb"], pp["ntype"], pp["natom"], pp["rat"], pp["adot"])
    fnl, _ = plan.forces(psi, d["rat"], d["avec"], d["proj_atom"], occ=occ)
    plan.close()
    ref = to_cart(z["fnl_cov"])
    err = np.abs(fnl - ref).max()
    print(f"non-local forces vs for_str_nl_kb: max |dF| {err:.2e} of max |F| {np.abs(ref).max():.2e} Ha/bohr")
    assert np.abs(ref).max() > 1e-5
    assert err < 1e-9
    # local term alone (no projectors on this plan)
    plan = hk_psi.Plan.from_inputs(d, projectors=False)
    floc, _ = plan.forces(psi, d["rat"], d["avec"], np.zeros(0, np.int32), d["floc_q"], d["floc_f"],
                          d["species"], occ=occ)
    plan.close()
    ref = to_cart(z["floc_cov"])
    err = np.abs(floc - ref).max()
    print(f"local forces vs for_str_local_force: max |dF| {err:.2e} of max |F| {np.abs(ref).max():.2e} Ha/bohr")
    assert np.abs(ref).max() > 1e-3
    assert err < 1e-9
