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


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replays_equal_eager(torch_dev, orc, prec):
    torch = torch_dev
    from cpw2000_amd import hk_psi

    ds = [synth.make_config(0, kpt=k) for k in KPTS]
    assert len({d["mtxd"] for d in ds}) > 1
    plans = [hk_psi.Plan.from_inputs(d) for d in ds]
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
