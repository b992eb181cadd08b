"""Inputs of proj_nl_kb_c16 (Src/omp/proj_nl_kb_c16.f90:23-73) for the
projector tests: the plane waves of one k-point (synth.build_kspace) and the
pseudopotential arrays of synth.kb_pseudo in the reference's layouts."""

from __future__ import annotations

import numpy as np

from cpw2000_amd import synth


def _case(spec, ks, pp):
    return dict(pp, rkpt=np.asarray(ks["kpt"], float), mtxd=ks["mtxd"], isort=ks["isort"], ng=ks["ng"],
                kgv=ks["kgv"])


def oracle_call(orc_kb, d, mxddim=None, mxdanl=None):
    mxddim = d["mtxd"] if mxddim is None else mxddim
    mxdanl = d["nanl"] if mxdanl is None else mxdanl
    return orc_kb.proj_nl_kb_c16(d["rkpt"], d["mtxd"], d["isort"], d["nanl"], d["ng"], d["kgv"], d["nqnl"],
                                 d["delqnl"], d["vkb"], d["nkb"], d["ntype"], d["natom"], d["rat"], d["adot"],
                                 mxddim, mxdanl, d["kgv"].shape[1])


def gauss_case(cfg, kpt=(0.0, 0.0, 0.0), delq=0.002, ecut=None):
    """s and p projectors with the Gaussian radial part of the synthetic
    generator, tabulated finely enough that the interpolation error stays
    below 1e-8."""
    spec = synth.config_spec(cfg)
    if ecut is not None:
        spec.ecut = ecut
    ks = synth.build_kspace(spec, kpt)
    return spec, ks, _case(spec, ks, synth.kb_pseudo(spec, delq))


def mixed_case(kpt=(0.13, -0.21, 0.37), seed=5):
    """The 2-atom fcc cell (a non-orthogonal metric) split into two species:
    species 1 with l = 0, 1, 3 (signs +1, -1, +1), species 2 with l = 0, 2, 3;
    random tables, species 2's cut short so that part of the sphere lies
    beyond its last interpolation point; mxdatm padded."""
    spec = synth.config_spec(0)
    frac = spec.types[0][1]
    spec.types = [("A", frac[:1], 1.8), ("B", frac[1:], 2.0)]
    ks = synth.build_kspace(spec, kpt)
    rng = np.random.default_rng(seed)
    mxdlqp = 700
    tab = rng.standard_normal((2, 4, mxdlqp + 3))
    pp = synth.kb_pseudo(spec, 0.01, lsets=[(0, 1, 3), (0, 2, 3)],
                         tables=lambda t, l, q: tab[t, l, :q.shape[0]], nqnl=[mxdlqp + 1, 300],
                         nkb_sign=[{0: 1, 1: -1, 3: 1}, {0: 1, 2: 1, 3: -1}], mxdatm_pad=2)
    assert pp["vkb"].shape[0] == mxdlqp + 3
    return spec, ks, _case(spec, ks, pp)


def golden_cases(path):
    """(inputs, reference anlga, reference xnlkb) of each case stored by
    tests/golden/make_golden.py (make_projkb_fixture), after checking that
    the generator still produces the stored inputs."""
    import json

    g = np.load(path)
    out = []
    for mt in json.loads(str(g["meta"])):
        _, _, d = mixed_case(**mt["kwargs"])
        vkb = np.asfortranarray(d["vkb"])
        w = np.cos(0.37 * np.arange(vkb.size))
        assert abs(float(np.dot(w, vkb.reshape(-1, order="F"))) - mt["digest_vkb"]) < 1e-9
        assert int(np.dot(np.arange(d["mtxd"]) % 97, d["isort"].astype(np.int64))) == mt["digest_isort"]
        assert (d["mtxd"], d["nanl"]) == (mt["mtxd"], mt["nanl"])
        out.append((d, g["anlga_" + mt["tag"]], g["xnlkb_" + mt["tag"]]))
    return out
