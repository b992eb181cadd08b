"""The per-kernel timing hook (cpwgpu_profile / cpwgpu_profile_read): with it on,
every kernel launch carries start and stop events in its dispatch.  The
results stay bit-identical to applies with the hook off, every kernel of the
call is counted once per launch, and no kernel's summed duration exceeds the
wall time of the applies it was timed in."""

import numpy as np
import pytest

from cpw2000_amd import synth

pytestmark = pytest.mark.gpu

STEPS = 4


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("sched", [0, 1, 2])
def test_profiled_applies_equal_plain_ones(torch_dev, prec, sched):
    torch = torch_dev
    from cpw2000_amd import hk_psi as hk

    d = synth.make_config(1)
    plan = hk.Plan.from_inputs(d)
    try:
        plan.set_variant(prec, hk.TUNE_SCHED, sched)
        x = torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")
        plain = torch.zeros_like(x)
        plan.apply(x, plain, prec)
        torch.cuda.synchronize()
        hk.profile_read()
        hk.profile(True)
        try:
            timed = torch.full_like(x, float("nan"))
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(STEPS):
                plan.apply(x, timed, prec)
            e1.record()
            torch.cuda.synchronize()
        finally:
            hk.profile(False)
        wall = e0.elapsed_time(e1)
        rec = hk.profile_read()
        assert torch.equal(timed, plain)
        for k in ("k_col_bwd", "k_plane", "k_col_fwd", "k_proj", "k_proj_reduce", "k_back"):
            ms, n = rec[k]
            assert n >= STEPS and n % STEPS == 0, (k, n)
            assert 0.0 < ms <= wall, (k, ms, wall)
        assert rec["k_proj_kb"][1] == 0
        # the hook off: nothing more is recorded
        plan.apply(x, timed, prec)
        torch.cuda.synchronize()
        assert all(n == 0 for _, n in hk.profile_read().values())
    finally:
        plan.close()


def test_profiled_projector_build(torch_dev):
    """The projector construction kernel, timed by its dispatch events, builds
    the same projectors as without the hook (same H|psi>)."""
    torch = torch_dev
    from cpw2000_amd import hk_psi as hk

    d = synth.make_config(0)
    pp = synth.kb_pseudo(synth.config_spec(0))
    plan = hk.Plan.from_inputs(d, projectors=False)
    rk = np.asarray(d["kpt"], float)
    x = torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")

    def build_and_apply():
        plan.build_projectors_kb(rk, pp["nanl"], pp["nqnl"], pp["delqnl"], pp["vkb"], pp["nkb"], pp["ntype"],
                                 pp["natom"], pp["rat"], pp["adot"])
        out = torch.zeros_like(x)
        plan.apply(x, out, 0)
        torch.cuda.synchronize()
        return out

    try:
        plain = build_and_apply()
        hk.profile_read()
        hk.profile(True, ["k_proj_kb"])
        try:
            timed = build_and_apply()
        finally:
            hk.profile(False)
        rec = hk.profile_read()
        assert torch.equal(timed, plain)
        ms, n = rec["k_proj_kb"]
        assert n >= 1 and ms > 0.0
        assert all(rec[k][1] == 0 for k in rec if k != "k_proj_kb")
    finally:
        plan.close()
