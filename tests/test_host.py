"""Host-side checks that need no GPU: the synthetic workloads have the sizes of
the benchmark configurations, the C-ABI library exports every declared entry
point, and argument errors are reported with the reference's conditions."""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

from cpw2000_amd import synth


@pytest.mark.parametrize("cfg,mtxd,mesh,ncol,nanl", [
    (0, 411, 24, 73, 8),
    (1, 23847, 72, 997, 256),
    (4, 25677, 80, 1049, 256),
])
def test_workload_sizes(cfg, mtxd, mesh, ncol, nanl):
    d = synth.make_config(cfg, nbands=2)
    assert d["mtxd"] == mtxd
    assert tuple(d["kmscr"][3:6]) == (mesh,) * 3
    g = d["kgv"][:, d["isort"][:mtxd] - 1]
    n1, n2 = d["kmscr"][3:5]
    assert len(set(zip(g[0] % n1, g[1] % n2))) == ncol
    assert d["nanl"] == nanl


def test_kpoint_workload_is_ragged():
    spec = synth.config_spec(3)
    sizes = {synth.build_kspace(spec, k)["mtxd"] for k in spec.kpts}
    assert min(sizes) >= 23960 and max(sizes) <= 23990 and len(sizes) > 1


def test_synthetic_inputs_are_physical():
    d = synth.make_config(0)
    p = d["psi"][: d["mtxd"]]
    np.testing.assert_allclose(p.conj().T @ p, np.eye(p.shape[1]), atol=1e-13)
    n1, n2, n3 = d["kmscr"][3:6]
    v = d["vscr"].reshape(n3, n2, n1)
    vg = np.fft.fftn(v) / v.size
    assert abs(vg[0, 0, 0]) < 1e-14                      # V(0) = 0
    assert np.abs(v).max() > 0.1


def test_size_fft_rule():
    # Src/ALFA/size_fft.f90: 2*k+1 -> smallest listed size
    assert synth.size_fft([10, 17, 35])[0] == [24, 36, 72]
    assert synth.size_fft([36, 53, 5])[0] == [80, 108, 12]


def _declared():
    src = open(os.path.join(ROOT, "include", "cpwgpu.h")).read()
    return sorted(set(re.findall(r"\b(cpwgpu_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from cpw2000_amd import hk_psi

    L = hk_psi.lib()
    names = _declared()
    assert set(names) == set(hk_psi.EXPORTS)
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", hk_psi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n


def test_fortran_shim_exports_hk_psi_c16_and_ditsp_c16():
    from cpw2000_amd import hk_psi

    assert os.path.exists(hk_psi.FLIB_PATH)
    out = subprocess.run(["nm", "-D", "--defined-only", hk_psi.FLIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    assert re.search(r"\bT hk_psi_c16_\b", out)
    assert re.search(r"\bT ditsp_c16_\b", out)
    assert re.search(r"\bT hk_psi_spin_c16_\b", out)
    assert re.search(r"\bT hk_psi_allk_c16_\b", out)
    assert re.search(r"\bT proj_nl_kb_c16_\b", out)


def test_proj_nl_kb_checks_before_any_device_work():
    """cpwgpu_proj_nl_kb_c16 and cpwgpu_plan_build_projectors_kb stop where
    the reference stops (Src/omp/proj_nl_kb_c16.f90:117, :143-157; the cell of
    Src/crystal/adot_to_avec.f90), and reject tables that a read at
    nint(|q|/delqnl) + 1 would overrun, before they touch the device."""
    from cpw2000_amd import hk_psi
    from tests.kb_case import mixed_case

    _, _, d = mixed_case()
    m, nanl = d["mtxd"], d["nanl"]

    def call(**over):
        a = dict(d, anlga=np.zeros((m, nanl), np.complex128, order="F"), xnlkb=np.zeros(nanl),
                 mxddim=m, mxdanl=nanl, mxdgve=d["kgv"].shape[1])
        a.update(over)
        hk_psi.proj_nl_kb_c16(**a)

    bad_adot = np.asfortranarray(np.diag([1.0, 1.0, -1.0]))
    short = np.asfortranarray(d["nqnl"] + np.array([2, 0], np.int32))   # 701 + 2 > mxdlqp + 1
    zero_dq = np.array([0.01, 0.0])
    for over, code in ((dict(nanl=nanl - 1, mxdanl=nanl), hk_psi.ERR_ARG),
                       (dict(mxdanl=nanl - 1), hk_psi.ERR_ARG),
                       (dict(ng=d["kgv"].shape[1] + 1), hk_psi.ERR_NG),
                       (dict(mxddim=m - 1, anlga=np.zeros((m - 1, nanl), np.complex128, order="F")),
                        hk_psi.ERR_MTXD),
                       (dict(adot=bad_adot), hk_psi.ERR_ARG),
                       (dict(nqnl=short), hk_psi.ERR_ARG),
                       (dict(delqnl=zero_dq), hk_psi.ERR_ARG),
                       (dict(ntype=3), hk_psi.ERR_ARG)):
        with pytest.raises(hk_psi.CpwGpuError) as e:
            call(**over)
        assert e.value.code == code, over
    with pytest.raises(hk_psi.CpwGpuError, match="inconsistent value of nanl"):
        call(nanl=nanl + 1, anlga=np.zeros((m, nanl + 1), np.complex128, order="F"), xnlkb=np.zeros(nanl + 1),
             mxdanl=nanl + 1)


def test_allk_argument_checks_before_any_device_work():
    """cpwgpu_hk_psi_allk_c16 checks every k-point's sizes (and the shared mesh)
    before it touches the device; nrk = 0 is a no-op."""
    from cpw2000_amd import hk_psi

    d = synth.make_config(0, nbands=2)
    m, nanl = d["mtxd"], d["nanl"]
    f = lambda shape, dt: np.zeros(shape, dt, order="F")

    def args(**over):
        a = dict(nrk=1, mtxd_allk=np.array([m, 0], np.int32), nband=np.array([2, 0], np.int32),
                 psi_allk=f((m, 2, 2), np.complex128), hpsi_allk=f((m, 2, 2), np.complex128), ng=d["ng"],
                 kgv=np.asfortranarray(d["kgv"], np.int32), ekpg_allk=f((m, 2), np.float64),
                 isort_allk=f((m, 2), np.int32), vscr=np.asfortranarray(d["vscr"]),
                 kmscr=np.asfortranarray(d["kmscr"], np.int32), anlga_allk=f((m, nanl, 2), np.complex128),
                 xnlkb_allk=f((nanl, 2), np.float64), nanl_allk=np.array([nanl, 0], np.int32), mxddim=m,
                 mxdbnd=2, mxdanl=nanl, mxdgve=d["mxdgve"], mxdscr=d["mxdscr"], mxdnrk=2)
        a.update(over)
        return a

    hk_psi.hk_psi_allk_c16(**args(nrk=0))
    for over, code in ((dict(nrk=3), hk_psi.ERR_ARG),
                       (dict(mtxd_allk=np.array([m + 1, 0], np.int32)), hk_psi.ERR_MTXD),
                       (dict(nband=np.array([3, 0], np.int32)), hk_psi.ERR_ARG),
                       (dict(nanl_allk=np.array([nanl + 1, 0], np.int32)), hk_psi.ERR_ARG),
                       (dict(mxdgve=d["ng"] - 1), hk_psi.ERR_NG),
                       (dict(mxdscr=100), hk_psi.ERR_MXDSCR)):
        with pytest.raises(hk_psi.CpwGpuError) as e:
            hk_psi.hk_psi_allk_c16(**args(**over))
        assert e.value.code == code, over


def _ditsp_args(d, neig, **over):
    """Arguments of ditsp_c16 for a config dict (host arrays, Fortran layout)."""
    mtxd, mxddim = d["mtxd"], d["mxddim"]
    a = dict(ipr=0, icmax=50, iguess=0, epspsi=1e-6, lnewanl=True, neig=neig, mtxd=mtxd,
             mtxds=3 * neig + 20, mxdsml=3 * neig + 20, mxdbnd=neig)
    a.update(over)
    psi = np.zeros((mxddim, a["mxdbnd"]), np.complex128, order="F")
    out = dict(
        a, psi=psi, hpsi=np.zeros_like(psi, order="F"), ei=np.zeros(a["mxdbnd"]),
        ekpg=np.asfortranarray(d["ekpg"]), isort=np.asfortranarray(d["isort"], np.int32),
        vscr=np.asfortranarray(d["vscr"]), kmscr=np.asfortranarray(d["kmscr"], np.int32),
        ng=d["ng"], kgv=np.asfortranarray(d["kgv"], np.int32),
        anlga=np.asfortranarray(d["anlga"], np.complex128), xnlkb=np.asfortranarray(d["xnlkb"]),
        nanl=d["nanl"], hamsm=np.zeros((a["mxdsml"], a["mxdsml"]), np.complex128, order="F"),
        hdiag=np.asfortranarray(d["ekpg"].copy()), mxddim=mxddim, mxdgve=d["mxdgve"],
        mxdscr=d["mxdscr"], mxdanl=max(d["mxdanl"], 1))
    out.update(over)
    return out


def test_ditsp_argument_checks_of_the_reference():
    """cpwgpu_ditsp_c16 reports the stop conditions of Src/ditsp/ditsp_c16.f90:151-212
    before any device work; neig < 1 returns with ifail = 0."""
    from cpw2000_amd import hk_psi

    d = synth.make_config(0, nbands=2)
    assert hk_psi.ditsp_c16(**_ditsp_args(d, 0, mtxds=8, mxdsml=8, mxdbnd=1)) == 0
    for over, code in ((dict(mtxds=4), hk_psi.ERR_ARG),                 # neig > mtxds
                       (dict(mxdbnd=4), hk_psi.ERR_ARG),                # neig > mxdbnd
                       (dict(mxdsml=20, mtxds=20), hk_psi.ERR_ARG),     # 3 neig > mxdsml
                       (dict(mtxds=60), hk_psi.ERR_ARG),                # mtxds > mxdsml
                       (dict(mtxd=d["mxddim"] + 1), hk_psi.ERR_MTXD),   # mtxd > mxddim
                       (dict(mxdgve=d["ng"] - 1), hk_psi.ERR_NG)):      # ng > mxdgve
        args = _ditsp_args(d, 8, **over)
        with pytest.raises(hk_psi.CpwGpuError) as e:
            hk_psi.ditsp_c16(**args)
        assert e.value.code == code, over


def test_argument_errors_before_any_device_work():
    """The drop-in reports the reference's stop conditions as status codes."""
    from cpw2000_amd import hk_psi

    d = synth.make_config(0, nbands=2)
    with pytest.raises(hk_psi.CpwGpuError) as e:
        hk_psi.hk_psi_c16_dict(dict(d, mxdgve=d["ng"] - 1))
    assert e.value.code == hk_psi.ERR_NG
    with pytest.raises(hk_psi.CpwGpuError) as e:
        hk_psi.hk_psi_c16_dict(dict(d, mxdscr=100))
    assert e.value.code == hk_psi.ERR_MXDSCR
    km = d["kmscr"].copy()
    km[3:7] = (10, 10, 10, 10)
    with pytest.raises(hk_psi.CpwGpuError) as e:
        hk_psi.hk_psi_c16_dict(dict(d, kmscr=km))
    assert e.value.code == hk_psi.ERR_WRAP
    km = d["kmscr"].copy()
    km[3:7] = (28, 24, 24, 28)                           # 28 = 4*7: not 2^a 3^b 5^c
    with pytest.raises(hk_psi.CpwGpuError) as e:
        hk_psi.hk_psi_c16_dict(dict(d, kmscr=km))
    assert e.value.code == hk_psi.ERR_SIZE
    with pytest.raises(hk_psi.CpwGpuError) as e:
        hk_psi.set_precision(7)
    assert e.value.code == hk_psi.ERR_MODE


def test_hk_psi_spin_argument_checks_of_the_reference():
    """cpwgpu_hk_psi_spin_c16 reports the stops of
    Src/hk_psi/reference/hk_psi_spin_c16.f90:106-141 before any device work."""
    from cpw2000_amd import hk_psi

    d = synth.make_config(0, nbands=2)
    s = synth.spin_inputs(d, 2)

    def call(nsp=4, mxdnsp=4, mxdgve=None, mxdscr=None):
        hpsi = np.zeros_like(s["psi_sp"], order="F")
        hk_psi.hk_psi_spin_c16(d["mtxd"], 2, s["psi_sp"], hpsi, True, d["ng"],
                               np.asfortranarray(d["kgv"], np.int32), d["ekpg"],
                               np.asfortranarray(d["isort"], np.int32), s["vscr_sp"],
                               np.asfortranarray(d["kmscr"], np.int32), nsp, s["anlsp"],
                               s["xnlkbsp"], s["nanlsp"], d["mxddim"], 2, s["mxdasp"],
                               d["mxdgve"] if mxdgve is None else mxdgve,
                               d["mxdscr"] if mxdscr is None else mxdscr, mxdnsp)

    for kw, code in ((dict(mxdgve=d["ng"] - 1), hk_psi.ERR_NG),
                     (dict(nsp=4, mxdnsp=2), hk_psi.ERR_ARG),
                     (dict(nsp=3), hk_psi.ERR_MODE),
                     (dict(mxdscr=100), hk_psi.ERR_MXDSCR)):
        with pytest.raises(hk_psi.CpwGpuError) as e:
            call(**kw)
        assert e.value.code == code, kw
