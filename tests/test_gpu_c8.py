"""Single-precision drop-in cpwgpu_hk_psi_c8 (complex*8 interface) against the
FP64 C oracle evaluated on the same FP32-rounded inputs: relative 2-norm
error <= 2e-6 (the FP32-mixed tolerance), rows beyond mtxd untouched."""

import numpy as np
import pytest

from cpw2000_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hk():
    import torch

    assert torch.cuda.is_available()
    from cpw2000_amd import hk_psi

    return hk_psi


@pytest.mark.parametrize("cfg,nb,pad", [(0, 8, 0), (0, 5, 3), (1, 4, 0)])
def test_hk_psi_c8_matches_oracle_on_rounded_inputs(hk, cfg, nb, pad):
    d = synth.make_config(cfg, nbands=nb, mxddim_pad=pad)
    f = lambda a, t: np.asfortranarray(np.asarray(a).astype(t))
    psi, anl = f(d["psi"], np.complex64), f(d["anlga"], np.complex64)
    ek, vs, xn = f(d["ekpg"], np.float32), f(d["vscr"], np.float32), f(d["xnlkb"], np.float32)
    hpsi = np.full_like(psi, 5.0 + 0j, order="F")
    hk.hk_psi_c8(d["mtxd"], nb, psi, hpsi, True, d["ng"], np.asfortranarray(d["kgv"], np.int32), ek,
                 np.asfortranarray(d["isort"], np.int32), vs, np.asfortranarray(d["kmscr"], np.int32), anl,
                 xn, d["nanl"], d["mxddim"], nb, d["mxdanl"], d["mxdgve"], d["mxdscr"])
    # the exact H of the rounded inputs
    dr = dict(d, ekpg=ek.astype(np.float64), vscr=vs.astype(np.float64),
              anlga=np.asfortranarray(anl.astype(np.complex128)), xnlkb=xn.astype(np.float64))
    st, ref = orc.orc_hk_psi(dr, psi=np.asfortranarray(psi.astype(np.complex128)))
    assert st == 0
    m = d["mtxd"]
    err = np.linalg.norm(hpsi[:m] - ref[:m]) / np.linalg.norm(ref[:m])
    assert err < 2e-6, err
    assert np.all(hpsi[m:] == 5.0)
