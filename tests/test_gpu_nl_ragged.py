"""Every projection and back-projection kernel on ragged shapes: projector
and band counts that are multiples of no kernel's chunk or tile, so that
full chunks are followed by a partial one and the edge tiles are partial,
checked against the oracle's H|psi> with the same random projectors.

FP64: relative 2-norm error <= 1e-12; FP32-mixed: <= 2e-6 (the tolerances
of the golden cases)."""

import numpy as np
import pytest

from tests.conftest import rel

from cpw2000_amd import synth

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6


@pytest.mark.parametrize("nanl,neig", [(37, 23), (61, 70), (100, 5)])
def test_nl_kernels_on_ragged_shapes(nanl, neig):
    import torch

    from cpw2000_amd import hk_psi as hk
    from oracle import oracle as O

    assert torch.cuda.is_available()
    d = synth.make_config(0, nbands=neig)
    mtxd, mxddim = d["mtxd"], d["mxddim"]
    rng = np.random.default_rng(1000 * nanl + neig)
    an = np.zeros((mxddim, nanl), np.complex128, order="F")
    an[:mtxd] = (rng.standard_normal((mtxd, nanl)) + 1j * rng.standard_normal((mtxd, nanl))) / np.sqrt(mtxd)
    d = dict(d, anlga=an, xnlkb=rng.uniform(-2.0, 2.0, nanl), nanl=nanl, mxdanl=nanl)
    st, want = O.orc_hk_psi(d)
    assert st == 0
    want = want[:mtxd]
    plan = hk.Plan.from_inputs(d)
    pt = torch.from_numpy(np.ascontiguousarray(d["psi"][:mtxd, :neig].T)).to("cuda")
    print(f"ragged: nanl={nanl} neig={neig} mtxd={mtxd}", flush=True)
    for prec, tol in ((hk.PREC_FP64, TOL64), (hk.PREC_FP32, TOL32)):
        for pv in range(hk.NPROJ_VAR):
            plan.set_variant(prec, hk.TUNE_PROJ, pv)
            for bv in range(hk.NBACK_VAR):
                plan.set_variant(prec, hk.TUNE_BACK, bv)
                ht = torch.zeros_like(pt)
                plan.apply(pt, ht, prec)
                torch.cuda.synchronize()
                err = rel(ht.cpu().numpy().T, want)
                assert err < tol, (prec, pv, bv, err)
        plan.set_variant(prec, hk.TUNE_PROJ, -1)
        plan.set_variant(prec, hk.TUNE_BACK, -1)
    plan.close()
