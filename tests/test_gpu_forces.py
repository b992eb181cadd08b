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
