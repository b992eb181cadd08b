"""cpwgpu_plan_apply_noalloc against cpwgpu_plan_apply, through the C ABI.

Two calls in a row on one plan, the second on fewer bands.  Where the bands
fit in one chunk of the fixed workspace the kernels and launch shapes are
those of cpwgpu_plan_apply, so the results are equal bit for bit; on the
64-atom cell (128 bands in several chunks) the split-K shape follows the
chunk, so the comparison is to the FP64 / FP32-mixed tolerances of the other
parity tests.  Device free memory is unchanged across the calls."""

import numpy as np
import pytest

from cpw2000_amd import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


def _psi(torch, d, nb=None):
    m = d["mtxd"]
    p = d["psi"][:m] if nb is None else d["psi"][:m, :nb]
    return torch.from_numpy(np.ascontiguousarray(p.T)).to("cuda")


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_noalloc_equals_apply_bitwise_small(torch_dev):
    torch = torch_dev
    from cpw2000_amd import hk_psi

    d = synth.make_config(0)
    plan = hk_psi.Plan.from_inputs(d)
    psi = _psi(torch, d)
    for prec in (hk_psi.PREC_FP64, hk_psi.PREC_FP32):
        for nb in (16, 7):  # second call: fewer bands
            x = psi[:nb].contiguous()
            ref = torch.zeros_like(x)
            plan.apply(x, ref, prec)  # also makes the FP32 projector copy
            got = torch.full_like(x, float("nan"))
            plan.apply_noalloc(x, got, prec)
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (prec, nb)
    plan.close()


def test_noalloc_chunked_64_atoms_no_device_allocation(torch_dev):
    torch = torch_dev
    from cpw2000_amd import hk_psi

    d = synth.make_config(1)
    plan = hk_psi.Plan.from_inputs(d)
    psi = _psi(torch, d)
    ref64 = torch.zeros_like(psi)
    plan.apply(psi, ref64, hk_psi.PREC_FP64)
    ref32 = torch.zeros_like(psi[:40])
    plan.apply(psi[:40].contiguous(), ref32, hk_psi.PREC_FP32)
    x40 = psi[:40].contiguous()
    got64 = torch.full_like(psi, float("nan"))
    got32 = torch.full_like(x40, float("nan"))
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    plan.apply_noalloc(psi, got64, hk_psi.PREC_FP64)
    plan.apply_noalloc(x40, got32, hk_psi.PREC_FP32)  # second call: fewer bands
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 >= free0, f"device free memory fell by {free0 - free1} bytes"
    assert _rel(got64, ref64) < TOL64
    assert _rel(got32, ref32) < TOL32
    # and both against FP64 on the same bands
    assert _rel(got32, ref64[:40]) < TOL32
    plan.close()
