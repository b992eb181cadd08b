"""H|psi> for several k-points in one call with the all-k arrays of the
self-consistency loop (cpwgpu_hk_psi_allk_c16, and the Fortran
hk_psi_allk_c16), against the oracle applied to each k-point alone.

Three k-points of the 2-atom cell (ragged plane-wave counts 411 / 401 / 410,
different band counts), arrays padded beyond every dimension: FP64 relative
2-norm error <= 1e-12 per k-point, FP32-mixed <= 2e-6; entries beyond
mtxd(k), nband(k) and nrk keep what they held; a second call (cached plans)
gives the same result."""

import numpy as np
import pytest

from cpw2000_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

KPTS = ((0.0, 0.0, 0.0), (0.25, 0.0, 0.0), (0.125, 0.25, 0.375))
NBANDS = (8, 5, 7)
TOL64, TOL32 = 1e-12, 2e-6


@pytest.fixture(scope="module")
def hk():
    from cpw2000_amd import hk_psi

    return hk_psi


def _allk():
    ds = [synth.make_config(0, nbands=nb, kpt=k) for k, nb in zip(KPTS, NBANDS)]
    nrk = len(ds)
    mxddim = max(d["mtxd"] for d in ds) + 3
    mxdbnd, mxdnrk = max(NBANDS) + 1, nrk + 1
    mxdanl = max(d["nanl"] for d in ds) + 2
    a = dict(nrk=nrk, mxddim=mxddim, mxdbnd=mxdbnd, mxdnrk=mxdnrk, mxdanl=mxdanl, ng=ds[0]["ng"],
             mxdgve=ds[0]["mxdgve"], mxdscr=ds[0]["mxdscr"],
             kgv=np.asfortranarray(ds[0]["kgv"], np.int32), vscr=np.asfortranarray(ds[0]["vscr"]),
             kmscr=np.asfortranarray(ds[0]["kmscr"], np.int32))
    f = lambda shape, dt: np.zeros(shape, dt, order="F")
    a.update(mtxd_allk=f(mxdnrk, np.int32), nband=f(mxdnrk, np.int32), nanl_allk=f(mxdnrk, np.int32),
             psi_allk=f((mxddim, mxdbnd, mxdnrk), np.complex128),
             hpsi_allk=np.full((mxddim, mxdbnd, mxdnrk), 9.0 + 0j, order="F"),
             ekpg_allk=f((mxddim, mxdnrk), np.float64), isort_allk=f((mxddim, mxdnrk), np.int32),
             anlga_allk=f((mxddim, mxdanl, mxdnrk), np.complex128), xnlkb_allk=f((mxdanl, mxdnrk), np.float64))
    refs = []
    for k, (d, nb) in enumerate(zip(ds, NBANDS)):
        m, na = d["mtxd"], d["nanl"]
        assert np.array_equal(d["kmscr"], ds[0]["kmscr"]) and np.array_equal(d["vscr"], ds[0]["vscr"])
        a["mtxd_allk"][k], a["nband"][k], a["nanl_allk"][k] = m, nb, na
        a["psi_allk"][:m, :nb, k] = d["psi"][:m, :nb]
        a["ekpg_allk"][:m, k] = d["ekpg"][:m]
        a["isort_allk"][:m, k] = d["isort"][:m]
        a["anlga_allk"][:m, :na, k] = d["anlga"][:m, :na]
        a["xnlkb_allk"][:na, k] = d["xnlkb"][:na]
        st, h = orc.orc_hk_psi(d, psi=d["psi"][:, :nb])
        assert st == 0
        refs.append(h[:m, :nb])
    return a, refs


def _check(a, refs, tol):
    h = a["hpsi_allk"]
    for k, ref in enumerate(refs):
        m, nb = ref.shape
        err = np.linalg.norm(h[:m, :nb, k] - ref) / np.linalg.norm(ref)
        assert err < tol, (k, err)
        assert np.all(h[m:, :, k] == 9.0) and np.all(h[:, nb:, k] == 9.0)
    assert np.all(h[:, :, len(refs):] == 9.0)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "fp32mixed"])
def test_allk_matches_oracle_per_kpoint(hk, prec):
    a, refs = _allk()
    hk.set_precision(prec)
    try:
        hk.hk_psi_allk_c16(**a)
        _check(a, refs, TOL64 if prec == 0 else TOL32)
        first = a["hpsi_allk"].copy()
        hk.hk_psi_allk_c16(**a)  # plans cached from the first call
        assert np.array_equal(a["hpsi_allk"], first)
    finally:
        hk.set_precision(hk.PREC_FP64)


def test_fortran_shim_hk_psi_allk_c16(hk):
    """The Fortran hk_psi_allk_c16_ (libcpwgpu_fortran.so), by reference."""
    import ctypes as C

    a, refs = _allk()
    lib = C.CDLL(hk.FLIB_PATH)
    i = lambda v: C.byref(C.c_int(int(v)))
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    lib.hk_psi_allk_c16_(i(a["nrk"]), p(a["mtxd_allk"]), p(a["nband"]), p(a["psi_allk"]), p(a["hpsi_allk"]),
                         i(a["ng"]), p(a["kgv"]), p(a["ekpg_allk"]), p(a["isort_allk"]), p(a["vscr"]),
                         p(a["kmscr"]), p(a["anlga_allk"]), p(a["xnlkb_allk"]), p(a["nanl_allk"]),
                         i(a["mxddim"]), i(a["mxdbnd"]), i(a["mxdanl"]), i(a["mxdgve"]), i(a["mxdscr"]),
                         i(a["mxdnrk"]))
    _check(a, refs, TOL64)
