"""Several k-points (uneven plane-wave counts) recorded into one HIP graph:
every replay equals the eager result (bit for bit: same kernels, same
order) with eager work and synchronisations between the replays, and the
FP64 products equal the oracle."""

import numpy as np
import pytest

from cpw2000_amd import synth

pytestmark = pytest.mark.gpu

KPTS = [(0.0, 0.0, 0.0), (0.1, 0.2, -0.15), (0.25, -0.25, 0.125)]


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("sched", [0, 1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replays_equal_eager(torch_dev, orc, prec, sched):
    """Under each stream schedule (the fork and join of the two streams are
    captured into the graph as well)."""
    torch = torch_dev
    from cpw2000_amd import hk_psi

    ds = [synth.make_config(0, kpt=k) for k in KPTS]
    assert len({d["mtxd"] for d in ds}) > 1
    plans = [hk_psi.Plan.from_inputs(d) for d in ds]
    for p in plans:
        p.set_variant(prec, hk_psi.TUNE_SCHED, sched)
    psis = [torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda") for d in ds]
    eager = []
    for p, x in zip(plans, psis):
        h = torch.zeros_like(x)
        p.apply(x, h, prec)
        eager.append(h)
    torch.cuda.synchronize()
    outs = [torch.zeros_like(x) for x in psis]
    g = hk_psi.Graph(plans, psis, outs, prec)
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for e, o in zip(eager, outs):
            assert torch.equal(e, o), f"replay {rep}"
        # eager work between replays
        h = torch.zeros_like(psis[1])
        plans[1].apply(psis[1], h, prec)
        torch.cuda.synchronize()
    if prec == 0:
        for d, o in zip(ds, outs):
            st, ho = orc.orc_hk_psi(d)
            ref = ho[: d["mtxd"]]
            assert np.linalg.norm(o.cpu().numpy().T - ref) < 1e-12 * np.linalg.norm(ref)
    g.close()
    for p in plans:
        p.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_follows_projector_rebuild(torch_dev, prec):
    """Projectors rebuilt on the device at moved atoms (same count) after the
    capture: the replay equals an eager apply with the new projectors in both
    precisions (the FP32 copy is converted again in place).  A rebuild with
    another projector count makes the launch fail instead of replaying."""
    torch = torch_dev
    from cpw2000_amd import hk_psi

    d = synth.make_config(0)
    pp = synth.kb_pseudo(synth.config_spec(0))
    plan = hk_psi.Plan.from_inputs(d, projectors=False)
    rk = np.asarray(d["kpt"], float)

    def build(rat, nkb=pp["nkb"], nanl=pp["nanl"]):
        plan.build_projectors_kb(rk, nanl, pp["nqnl"], pp["delqnl"], pp["vkb"], nkb, pp["ntype"], pp["natom"],
                                 rat, pp["adot"])

    build(pp["rat"])
    x = torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")
    out = torch.zeros_like(x)
    g = hk_psi.Graph([plan], [x], [out], prec)
    try:
        g.launch()
        torch.cuda.synchronize()
        before = out.clone()
        moved = pp["rat"].copy(order="F")
        moved[:, 0, 0] += (0.01, -0.02, 0.015)
        build(moved)
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        eager = torch.zeros_like(x)
        plan.apply(x, eager, prec)
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
        assert not torch.equal(before, eager)   # the projectors did change
        # s projectors only: another count, the launch is refused
        nkb_s = pp["nkb"].copy(order="F")
        nkb_s[1] = 0
        build(moved, nkb=nkb_s, nanl=pp["nanl"] // 4)
        with pytest.raises(hk_psi.CpwGpuError):
            g.launch()
    finally:
        g.close()
        plan.close()
