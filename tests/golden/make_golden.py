"""Writes the golden vectors of the H|psi> path from the reference itself.

The reference hk_psi_c16 (Src/hk_psi/reference/hk_psi_c16.f90 with the NCAR
FFT backend and netlib zgemm) is compiled from the cpw2000 sources by
``make -C oracle ref`` into oracle/_ref/libcpwref.so and called here on the
seeded synthetic inputs of cpw2000_amd.synth.  Each fixture holds the
generator arguments, digests of the inputs (to detect generator drift) and
the reference output.  Also stored: the reference's pruned 3-D FFT
(cfft_wf_c16) on a random sphere, both directions, and the reference's
projector construction (proj_nl_kb_c16, with ylm_rc, adot_to_avec and
adot_to_bdot from the same build) on the two-species fcc case of
tests/kb_case.py.

Run in the build container only (it needs the reference sources):
    python tests/golden/make_golden.py
"""

from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cpw2000_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

# name -> (make_config kwargs, number of bands applied, projectors kept)
CASES = {
    "cfg0": (dict(cfg=0), 16, True),
    "cfg0_k": (dict(cfg=0, kpt=(0.1, 0.2, -0.15)), 16, True),
    "cfg0_fullmesh": (dict(cfg=0, kmax_override=(5, 5, 5)), 16, True),
    "cfg0_pad": (dict(cfg=0, idshift=3, mxddim_pad=7), 16, True),
    "cfg0_nonl": (dict(cfg=0), 16, False),
    "ortho": (dict(cfg="ortho"), 8, True),
    "cfg1_b2": (dict(cfg=1), 2, True),
}
# large cases: only every ROW_STRIDE-th plane wave of the output is stored
LARGE_CASES = {
    "cfg2_b2": (dict(cfg=2), 2, True),
}
ROW_STRIDE = 4


def digest(d):
    """Order-sensitive digests of the inputs."""
    w = lambda n: np.cos(0.37 * np.arange(n))
    out = {}
    for k in ("psi", "vscr", "anlga", "ekpg"):
        a = np.asarray(d[k]).reshape(-1, order="F")
        out[k] = [float(np.real(np.dot(w(a.size), a))), float(np.imag(np.dot(w(a.size), a)))]
    out["isort"] = int(np.dot(np.arange(d["isort"].size) % 97, d["isort"].astype(np.int64)))
    return out


def inputs(name):
    kw, nb, proj = {**CASES, **LARGE_CASES}[name]
    d = synth.make_config(**kw)
    if not proj:
        d = dict(d)
        d["nanl"] = 0
    psi = np.asfortranarray(d["psi"][:, :nb])
    return d, psi


def make_fft_fixture():
    """Reference cfft_wf_c16 on a random sphere (12 x 16 x 20 mesh, kd = 3,4,5)."""
    rng = np.random.Generator(np.random.PCG64(20261234))
    n1, n2, n3 = 16, 20, 24
    kd = (3, 4, 5)
    chd = np.zeros((n3, n2, n1), complex)
    for k3 in range(-kd[2], kd[2] + 1):
        for k2 in range(-kd[1], kd[1] + 1):
            for k1 in range(-kd[0], kd[0] + 1):
                chd[k3 % n3, k2 % n2, k1 % n1] = rng.normal() + 1j * rng.normal()
    full = rng.normal(size=(n3, n2, n1)) + 1j * rng.normal(size=(n3, n2, n1))
    lib = O.ref_lib()
    i = lambda v: C.byref(C.c_int(v))
    out = {}
    for mode, src in ((-1, chd), (1, full)):
        buf = np.ascontiguousarray(src).copy()
        wrk = np.zeros(4 * max(n1 * n2, n1 * n3, n2 * n3))
        lib.cfft_wf_c16_(buf.ctypes.data_as(C.c_void_p), i(n1), i(n1), i(n2), i(n3), i(kd[0]),
                         i(kd[1]), i(kd[2]), i(mode), wrk.ctypes.data_as(C.c_void_p), i(wrk.size))
        out[mode] = buf
    np.savez_compressed(os.path.join(HERE, "fft_wf_16x20x24.npz"), ns=np.array([n1, n2, n3]),
                        kd=np.array(kd), chd_in_bwd=chd, chd_out_bwd=out[-1], chd_in_fwd=full,
                        chd_out_fwd=out[1])


def make_projkb_fixture():
    """Reference proj_nl_kb_c16 on tests/kb_case.mixed_case (fcc metric, two
    species, l = 0..3, signs of nkb, random tables, one ending inside the
    sphere, k off the origin) with padded mxddim / mxdanl / mxdatm."""
    from tests.kb_case import mixed_case

    lib = O.ref_lib()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    i = lambda v: C.byref(C.c_int(int(v)))
    meta = []
    out = {}
    for tag, kw in (("a", dict()), ("b", dict(kpt=(0.0, 0.5, 0.0), seed=9))):
        _, _, d = mixed_case(**kw)
        m, nanl = d["mtxd"], d["nanl"]
        mxddim, mxdanl = m + 3, nanl + 2
        anl = np.zeros((mxddim, mxdanl), complex, order="F")
        xnl = np.zeros(mxdanl)
        vkb = np.asfortranarray(d["vkb"])
        lib.proj_nl_kb_c16_(p(np.ascontiguousarray(d["rkpt"])), i(m), p(d["isort"]), i(nanl), i(d["ng"]),
                            p(np.asfortranarray(d["kgv"], np.int32)), p(d["nqnl"]), p(d["delqnl"]), p(vkb),
                            p(np.asfortranarray(d["nkb"], np.int32)), i(d["ntype"]), p(d["natom"]),
                            p(np.asfortranarray(d["rat"])), p(np.asfortranarray(d["adot"])), p(anl), p(xnl),
                            i(vkb.shape[3]), i(d["rat"].shape[1]), i(vkb.shape[0] - 3), i(mxddim), i(mxdanl),
                            i(d["kgv"].shape[1]))
        w = np.cos(0.37 * np.arange(vkb.size))
        meta.append(dict(tag=tag, kwargs=kw, mtxd=int(m), nanl=int(nanl),
                         digest_vkb=float(np.dot(w, vkb.reshape(-1, order="F"))),
                         digest_isort=int(np.dot(np.arange(m) % 97, d["isort"].astype(np.int64)))))
        out[f"anlga_{tag}"] = anl[:m, :nanl]
        out[f"xnlkb_{tag}"] = xnl[:nanl]
        print("proj_nl_kb_c16", tag, m, nanl, "|anlga| =", np.linalg.norm(anl))
    np.savez_compressed(os.path.join(HERE, "projkb_fcc2.npz"), meta=json.dumps(meta), **out)


# two-component case: hk_psi_spin_c16 on the 2-atom cell (24^3) with the
# inputs of synth.spin_inputs (random sigma_z / sigma_x / sigma_y potentials,
# spin-mixing projectors), 6 two-component bands
SPIN_BANDS, SPIN_SEED = 6, 21


def make_spin_fixture():
    """Reference hk_psi_spin_c16 for nsp = 1, 2 and 4."""
    lib = O.ref_lib()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    i = lambda v: C.byref(C.c_int(int(v)))
    d = synth.make_config(0, nbands=2)
    sp = synth.spin_inputs(d, SPIN_BANDS, seed=SPIN_SEED)
    kgv = np.asfortranarray(d["kgv"], np.int32)
    out = {}
    for nsp in (1, 2, 4):
        psi = np.asfortranarray(sp["psi_sp"])
        h = np.zeros_like(psi, order="F")
        lnew = C.c_int(1)
        lib.hk_psi_spin_c16_(i(d["mtxd"]), i(SPIN_BANDS), p(psi), p(h), C.byref(lnew), i(d["ng"]), p(kgv),
                             p(np.ascontiguousarray(d["ekpg"])), p(np.ascontiguousarray(d["isort"], np.int32)),
                             p(np.asfortranarray(sp["vscr_sp"])), p(np.ascontiguousarray(d["kmscr"], np.int32)),
                             i(nsp), p(np.asfortranarray(sp["anlsp"])), p(np.ascontiguousarray(sp["xnlkbsp"])),
                             i(sp["nanlsp"]), i(d["mxddim"]), i(SPIN_BANDS), i(sp["mxdasp"]), i(d["mxdgve"]),
                             i(d["mxdscr"]), i(sp["mxdnsp"]))
        out[f"hpsi_nsp{nsp}"] = h[: 2 * d["mtxd"]]
        print("hk_psi_spin_c16 nsp", nsp, "|hpsi| =", np.linalg.norm(h))
    w = np.cos(0.37 * np.arange(sp["vscr_sp"].size))
    meta = dict(nbands=SPIN_BANDS, seed=SPIN_SEED, digest=digest(d),
                digest_vsp=float(np.dot(w, sp["vscr_sp"].reshape(-1, order="F"))),
                digest_psi=[float(np.real(np.sum(sp["psi_sp"] * np.cos(np.arange(sp["psi_sp"].shape[0]))[:, None]))),
                            float(np.imag(np.sum(sp["psi_sp"] * np.cos(np.arange(sp["psi_sp"].shape[0]))[:, None])))])
    np.savez_compressed(os.path.join(HERE, "hkpsi_spin_cfg0.npz"), meta=json.dumps(meta), **out)


# forces: the Si/Ge 64-atom cell, FORCE_BANDS bands with occupations
# FORCE_OCC, the s+p KB tables of synth.kb_pseudo
FORCE_BANDS = 32


def force_occ(nb=FORCE_BANDS):
    return np.linspace(2.0, 0.5, nb)


def make_forces_fixture():
    """The reference's force routines on the Si/Ge cell.

    * for_str_nl_kb (non-local, covariant, per unit of the fractional
      coordinate) with the KB tables of synth.kb_pseudo, the first
      FORCE_BANDS bands and occupations force_occ();
    * for_str_local_force with one star per q vector of the local form
      factors (mstar = 1, phase = 1, conj = 1; star 1 is q = 0), den(q) the
      band density's Fourier coefficients (numpy FFT of |psi(r)|^2, the
      convention rho(q) = (1/N) sum_r rho(r) exp(-i q.r)), vql = the form
      factors, no core charge."""
    from oracle import forces as Fo

    lib = O.ref_lib()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    i = lambda v: C.byref(C.c_int(int(v)))
    d = synth.make_config(4)
    spec = synth.config_spec(4)
    pp = synth.kb_pseudo(spec)
    nb = FORCE_BANDS
    psi = np.asfortranarray(d["psi"][:, :nb])
    occ = np.ascontiguousarray(force_occ(nb))
    ntype, mxdatm = int(pp["ntype"]), pp["rat"].shape[1]
    vkb = np.asfortranarray(pp["vkb"])
    kgv = np.asfortranarray(d["kgv"], np.int32)
    fnl = np.zeros((3, mxdatm, ntype), order="F")
    snl = np.zeros((3, 3), order="F")
    lib.for_str_nl_kb_(p(fnl), p(snl), i(d["mtxd"]), p(np.ascontiguousarray(d["kpt"], float)), i(nb), p(occ),
                       p(np.ascontiguousarray(d["isort"], np.int32)), p(psi), p(kgv), p(pp["nqnl"]), p(pp["delqnl"]),
                       p(vkb), p(np.asfortranarray(pp["nkb"], np.int32)), i(ntype), p(pp["natom"]),
                       p(np.asfortranarray(pp["rat"])), p(np.asfortranarray(pp["adot"])), i(ntype), i(mxdatm),
                       i(kgv.shape[1]), i(vkb.shape[0] - 3), i(d["mxddim"]), i(nb))
    # local term: one star per q (star 1 = q 0)
    q = np.asarray(d["floc_q"], np.int64)
    keep = np.any(q != 0, axis=1)
    q, fq = q[keep], np.asarray(d["floc_f"])[keep]
    n1, n2, n3 = [int(x) for x in d["kmscr"][3:6]]
    rq = Fo.band_density_q(d, psi, occ)
    den = np.concatenate([[0.0], rq[q[:, 2] % n3, q[:, 1] % n2, q[:, 0] % n1]]).astype(complex)
    ns = q.shape[0] + 1
    kloc = np.asfortranarray(np.concatenate([np.zeros((1, 3), np.int64), q]).T.astype(np.int32))
    vql = np.asfortranarray(np.concatenate([np.zeros((1, ntype)), fq]).T)
    phase = np.ones(ns, complex)
    conj = np.ones(ns)
    mstar = np.ones(ns, np.int32)
    vxc = np.zeros(ns, complex)
    dnc = np.zeros((ntype, ns), order="F")
    floc = np.zeros((3, mxdatm, ntype), order="F")
    lib.for_str_local_force_(p(floc), i(ns), p(kloc), p(phase), p(conj), i(ns), p(mstar), p(vxc), p(den),
                             i(ntype), p(pp["natom"]), p(np.asfortranarray(pp["rat"])), p(vql), p(dnc), i(ns),
                             i(ns), i(ntype), i(mxdatm))
    natom = pp["natom"]
    glob = lambda f: np.concatenate([f[:, :natom[t], t].T for t in range(ntype)])
    meta = dict(nbands=nb, digest=digest(d), nq=int(q.shape[0]))
    np.savez_compressed(os.path.join(HERE, "forces_sige64.npz"), meta=json.dumps(meta), fnl_cov=glob(fnl),
                        floc_cov=glob(floc), occ=occ, den=den[1:])
    print("for_str_nl_kb |f| =", np.abs(fnl).max(), " for_str_local_force |f| =", np.abs(floc).max())


def main(only=None):
    """Writes every fixture, or only the hk_psi cases named in ``only``."""
    if not O.build_ref():
        raise SystemExit("reference sources not found: cannot write golden vectors")
    for name in CASES:
        if only and name not in only:
            continue
        d, psi = inputs(name)
        hpsi = O.ref_hk_psi(d, psi=psi)
        kw, nb, proj = CASES[name]
        np.savez_compressed(os.path.join(HERE, f"hkpsi_{name}.npz"),
                            meta=json.dumps(dict(kwargs=kw, nbands=nb, projectors=proj,
                                                 digest=digest(d))),
                            hpsi=hpsi[: d["mtxd"]])
        print(name, d["mtxd"], nb, "|hpsi| =", np.linalg.norm(hpsi))
    for name, (kw, nb, proj) in LARGE_CASES.items():
        if only and name not in only:
            continue
        d, psi = inputs(name)
        hpsi = O.ref_hk_psi(d, psi=psi)
        np.savez_compressed(os.path.join(HERE, f"hkpsi_{name}.npz"),
                            meta=json.dumps(dict(kwargs=kw, nbands=nb, projectors=proj, row_stride=ROW_STRIDE,
                                                 digest=digest(d))),
                            hpsi=hpsi[: d["mtxd"]:ROW_STRIDE])
        print(name, d["mtxd"], nb, "|hpsi| =", np.linalg.norm(hpsi))
    if only:
        if "spin" in only:
            make_spin_fixture()
        if "forces" in only:
            make_forces_fixture()
        return
    make_fft_fixture()
    make_projkb_fixture()
    make_spin_fixture()
    make_forces_fixture()


if __name__ == "__main__":
    main(sys.argv[1:] or None)
