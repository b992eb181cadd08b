"""H|psi> on the specialised mesh lengths 48, 64, 80, 96, 108 and 128 (the
column and plane kernels compiled for each length; 24 and 72 are the golden
cases) against the oracle on the same inputs, which is pinned there by the
G-space convolution form and the reference build (test_oracle_meshes.py).

Every plane-stage variant built for the length (the LDS two-pass kernel and
the register shapes whose line split divides it), with the LDS-staged and a
direct-load projection GEMM: FP64 relative 2-norm error <= 1e-12,
FP32-mixed <= 2e-6."""

import numpy as np
import pytest

from tests.conftest import ENLARGED_MESHES, mesh_config, rel

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 2e-6
NB = 4


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def hk():
    from cpw2000_amd import hk_psi

    return hk_psi


def _built(hk, plan, prec):
    out = []
    for v in range(hk.NPLANE_VAR):
        try:
            plan.set_variant(prec, hk.TUNE_PLANE, v)
        except hk.CpwGpuError as e:
            assert e.code == hk.ERR_MODE
            continue
        out.append(v)
    return out


@pytest.mark.parametrize("cfg,kmax,n", ENLARGED_MESHES, ids=[f"n{m[2]}" for m in ENLARGED_MESHES])
def test_specialised_mesh_matches_oracle(torch_dev, hk, orc, cfg, kmax, n):
    torch = torch_dev
    d = mesh_config(cfg, kmax, NB)
    assert [int(x) for x in d["kmscr"][3:6]] == [n, n, n]
    st, href = orc.orc_hk_psi(d)
    assert st == 0
    mtxd = d["mtxd"]
    href = href[:mtxd].T
    plan = hk.Plan.from_inputs(d)
    info = plan.info()
    assert (info["n1"], info["n2"], info["n3"]) == (n, n, n)
    pt = torch.from_numpy(np.ascontiguousarray(d["psi"][:mtxd].T)).to("cuda")
    for prec, tol in ((hk.PREC_FP64, TOL64), (hk.PREC_FP32, TOL32)):
        vs = _built(hk, plan, prec)
        assert 0 in vs, vs
        for v in vs:
            plan.set_variant(prec, hk.TUNE_PLANE, v)
            for jv in (0, 2):
                plan.set_variant(prec, hk.TUNE_PROJ, jv)
                ht = torch.zeros_like(pt)
                plan.apply(pt, ht, prec)
                torch.cuda.synchronize()
                err = rel(ht.cpu().numpy(), href)
                assert err < tol, (n, prec, v, jv, err)
        plan.set_variant(prec, hk.TUNE_PLANE, -1)
        plan.set_variant(prec, hk.TUNE_PROJ, -1)
    plan.close()
