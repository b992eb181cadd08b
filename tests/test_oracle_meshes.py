"""The oracle on the other specialised mesh lengths of the device kernels
(48, 64, 80, 96, 108, 128; 24 and 72 are the golden cases): the 2-atom and
64-atom cells with the FFT box enlarged (kmscr(1:3) raised until size_fft
gives the length).  The potential is band-limited, so H|psi> does not depend
on the mesh: the C restatement must equal the exact G-space convolution form
(no mesh, no FFT) and, where it is built, the reference's own hk_psi_c16
(oracle/_ref), to 1e-13 relative 2-norm."""

import pytest

from tests.conftest import ENLARGED_MESHES, mesh_config, rel


@pytest.mark.parametrize("cfg,kmax,n", ENLARGED_MESHES, ids=[f"n{m[2]}" for m in ENLARGED_MESHES])
def test_restatement_on_enlarged_meshes(orc, cfg, kmax, n):
    d = mesh_config(cfg, kmax, 2)
    assert [int(x) for x in d["kmscr"][3:6]] == [n, n, n]
    st, h = orc.orc_hk_psi(d)
    assert st == 0
    mtxd = d["mtxd"]
    assert rel(h[:mtxd], orc.conv_apply(d, d["psi"])) < 1e-13
    if orc.have_ref():
        hr = orc.ref_hk_psi(d)
        assert rel(h[:mtxd], hr[:mtxd]) < 1e-13
