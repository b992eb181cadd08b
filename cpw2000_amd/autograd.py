"""Differentiable potential energy of the bands with respect to atomic positions.

``PotentialEnergy`` is a ``torch.autograd.Function``:

    forward:  E(tau) = sum_n occ_n <psi_n| V_loc(tau) + V_NL(tau) |psi_n>   (psi fixed)
              from H|psi> on the GPU in FP64 (E = <psi|H psi> - kinetic)
    backward: dE/dtau = -F(tau), F from the adjoint kernels behind
              cpwgpu_plan_forces (structure-factor derivative of V_loc, d beta/d tau
              projections of V_NL)

``builder(tau_cart)`` returns the hk_psi inputs (a dict with the keys of
cpw2000_amd.synth.make_config) for Cartesian positions ``tau_cart`` (natom, 3).
This is the derivative that Src/ALFA/for_str_nl_kb.f90 and
Src/ALFA/for_str_local_force.f90 compute for the same energy.
"""

from __future__ import annotations

import numpy as np
import torch

from . import hk_psi


def _energy_and_plan(d, psi_t):
    plan = hk_psi.Plan.from_inputs(d)
    h = torch.zeros_like(psi_t)
    plan.apply(psi_t, h, hk_psi.PREC_FP64)
    torch.cuda.synchronize()
    mtxd = d["mtxd"]
    ek = torch.from_numpy(np.ascontiguousarray(d["ekpg"][:mtxd])).to(psi_t.device)
    e = torch.sum((psi_t.conj() * h).real) - torch.sum(ek[None, :] * psi_t.abs() ** 2)
    return float(e), plan


class PotentialEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tau_cart, builder, psi_t):
        d = builder(tau_cart.detach().cpu().numpy())
        e, plan = _energy_and_plan(d, psi_t)
        ctx.save_for_backward(tau_cart)
        ctx.builder = builder
        ctx.psi_t = psi_t
        ctx.d = d
        ctx.plan = plan
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_e):
        d = ctx.d
        f, _ = ctx.plan.forces(ctx.psi_t, d["rat"], d["avec"], d["proj_atom"], d.get("floc_q"),
                               d.get("floc_f"), d.get("species"))
        # the plan stays alive with ctx: backward may run more than once
        g = -torch.from_numpy(f) * grad_e
        return g, None, None
