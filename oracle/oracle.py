"""ORACLE -- test infrastructure only.

Python access to the three CPU checkers of the H|psi> path:

* ``orc_*``    the C restatement in ``hk_psi_oracle.c`` (built as ``liborc.so``),
               following Src/hk_psi/reference/hk_psi_c16.f90 step by step.
* ``ref_*``    the reference Fortran ``hk_psi_c16`` itself, compiled from the
               cpw2000 sources by ``make ref`` into ``_ref/libcpwref.so``
               (present only where it was built).
* ``dense_*`` / ``conv_*``  textbook forms of the same operator that use no FFT:
               H_ij = ekpg_i d_ij + V(G_i - G_j) + sum_a anlga_ia D_a conj(anlga_ja)
               (the reference's own dense representation,
               Src/ALFA/hamilt_local.f90:119-120, Src/ALFA/hamilt_kb_non_local.f90:109-114),
               applied either as a dense matrix or as a sparse convolution over
               the nonzero V(q) of the synthetic potential.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(HERE, "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libcpwref.so")
# the projector construction alone, built with OpenMP (its all-core baseline)
REF_OMP_SO = os.path.join(HERE, "_ref", "libcpwref_omp.so")

_orc = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build_orc():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def build_ref(refsrc="/root/reference/Src"):
    """Build the reference library (only where the cpw2000 sources exist)."""
    if not os.path.isdir(refsrc):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref", f"REFSRC={refsrc}"], check=True)
    return os.path.exists(REF_SO)


def orc_lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_SO):
            build_orc()
        _orc = C.CDLL(ORC_SO)
        _orc.orc_hk_psi_c16.restype = C.c_int
        _orc.orc_cfft_wf_c16.restype = C.c_int
        _orc.orc_num_threads.restype = C.c_int
        _orc.orc_set_num_threads.restype = None
        _orc.orc_set_num_threads.argtypes = [C.c_int]
    return _orc


def have_ref():
    return os.path.exists(REF_SO)


_ref_omp = None


def ref_omp_lib():
    """The OpenMP build of the reference proj_nl_kb_c16, or None where not built."""
    global _ref_omp
    if _ref_omp is None and os.path.exists(REF_OMP_SO):
        _ref_omp = C.CDLL(REF_OMP_SO)
    return _ref_omp


def ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
    return _ref


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _args(d, psi, hpsi, neig):
    return dict(
        mtxd=int(d["mtxd"]), neig=int(neig), psi=psi, hpsi=hpsi, ng=int(d["ng"]),
        kgv=np.asfortranarray(d["kgv"], np.int32), ekpg=np.ascontiguousarray(d["ekpg"], np.float64),
        isort=np.ascontiguousarray(d["isort"], np.int32),
        vscr=np.ascontiguousarray(d["vscr"], np.float64),
        kmscr=np.ascontiguousarray(d["kmscr"], np.int32),
        anlga=np.asfortranarray(d["anlga"], np.complex128),
        xnlkb=np.ascontiguousarray(d["xnlkb"], np.float64), nanl=int(d["nanl"]),
        mxddim=int(d["mxddim"]), mxdbnd=max(int(neig), 1), mxdanl=max(int(d["mxdanl"]), 1),
        mxdgve=int(d["mxdgve"]), mxdscr=int(d["mxdscr"]))


def orc_hk_psi(d, psi=None, neig=None, hpsi=None):
    """H|psi> by the C restatement.  Returns (status, hpsi)."""
    psi = np.asfortranarray(d["psi"] if psi is None else psi, np.complex128)
    neig = psi.shape[1] if neig is None else neig
    if hpsi is None:
        hpsi = np.zeros_like(psi, order="F")
    a = _args(d, psi, hpsi, neig)
    st = orc_lib().orc_hk_psi_c16(
        a["mtxd"], a["neig"], _ptr(psi), _ptr(hpsi), a["ng"], _ptr(a["kgv"]), _ptr(a["ekpg"]),
        _ptr(a["isort"]), _ptr(a["vscr"]), _ptr(a["kmscr"]), _ptr(a["anlga"]), _ptr(a["xnlkb"]),
        a["nanl"], a["mxddim"], a["mxdbnd"], a["mxdanl"], a["mxdgve"], a["mxdscr"])
    return st, hpsi


def ref_hk_psi(d, psi=None, neig=None, hpsi=None):
    """H|psi> by the reference Fortran (by-reference argument passing)."""
    psi = np.asfortranarray(d["psi"] if psi is None else psi, np.complex128)
    neig = psi.shape[1] if neig is None else neig
    if hpsi is None:
        hpsi = np.zeros_like(psi, order="F")
    a = _args(d, psi, hpsi, neig)
    i = lambda v: C.byref(C.c_int(v))
    lnew = C.c_int(1)
    ref_lib().hk_psi_c16_(
        i(a["mtxd"]), i(a["neig"]), _ptr(psi), _ptr(hpsi), C.byref(lnew), i(a["ng"]),
        _ptr(a["kgv"]), _ptr(a["ekpg"]), _ptr(a["isort"]), _ptr(a["vscr"]), _ptr(a["kmscr"]),
        _ptr(a["anlga"]), _ptr(a["xnlkb"]), i(a["nanl"]), i(a["mxddim"]), i(a["mxdbnd"]),
        i(a["mxdanl"]), i(a["mxdgve"]), i(a["mxdscr"]))
    return hpsi


def orc_fft1d(x, isign):
    """Unnormalised 1-D DFT of the restatement (isign>0: exp(-i), isign<0: exp(+i))."""
    buf = np.ascontiguousarray(np.asarray(x, np.complex128)).copy()
    orc_lib().orc_fft1d(C.c_int(buf.shape[0]), _ptr(buf), C.c_int(isign))
    return buf


def orc_cfft_wf(chd, ns, kd, mode, idd=None):
    """3-D transform of the restatement on chd laid out as (k3*n2+k2)*id+k1."""
    n1, n2, n3 = ns
    idd = n1 if idd is None else idd
    buf = np.ascontiguousarray(chd, np.complex128).copy()
    st = orc_lib().orc_cfft_wf_c16(_ptr(buf), idd, n1, n2, n3, kd[0], kd[1], kd[2], mode)
    return st, buf


# ---------------------------------------------------------------------------
# FFT-free forms of the operator
# ---------------------------------------------------------------------------

def pw_vectors(d):
    """Integer G of each plane wave (mtxd, 3)."""
    return d["kgv"].T[d["isort"][: d["mtxd"]] - 1].astype(np.int64)


def vq_from_mesh(d):
    """V(q) = (1/N) sum_r V(r) exp(-i q.r) from vscr (the transform the path applies)."""
    n1, n2, n3, idd = [int(x) for x in d["kmscr"][3:7]]
    v = d["vscr"][: idd * n2 * n3].reshape(n3, n2, idd)[:, :, :n1]
    return np.fft.fftn(v) / (n1 * n2 * n3)          # indexed [q3, q2, q1] mod n


def dense_h(d):
    """Dense FP64 Hamiltonian (mtxd x mtxd); V(G_i - G_j) read from the mesh transform."""
    g = pw_vectors(d)
    n1, n2, n3 = [int(x) for x in d["kmscr"][3:6]]
    vg = vq_from_mesh(d)
    dg = g[:, None, :] - g[None, :, :]
    h = vg[dg[..., 2] % n3, dg[..., 1] % n2, dg[..., 0] % n1].astype(np.complex128)
    h[np.diag_indices_from(h)] += d["ekpg"][: d["mtxd"]]
    nanl = d["nanl"]
    a = d["anlga"][: d["mtxd"], :nanl]
    h += (a * d["xnlkb"][None, :nanl]) @ a.conj().T
    return h


def conv_apply(d, psi, vq_g=None, vq=None):
    """H psi with V applied as the exact convolution sum_q V(q) psi(G - q) over the
    synthetic V(q) list (no mesh, no FFT)."""
    vq_g = d["vq_g"] if vq_g is None else vq_g
    vq = d["vq"] if vq is None else vq
    mtxd = d["mtxd"]
    g = pw_vectors(d)
    lo = g.min(axis=0)
    span = g.max(axis=0) - lo + 1
    lut = -np.ones(span, np.int64)
    lut[tuple((g - lo).T)] = np.arange(mtxd)
    psi = np.asarray(psi)[:mtxd]
    out = d["ekpg"][:mtxd, None] * psi
    for q, v in zip(np.asarray(vq_g, np.int64), vq):
        src = g - q[None, :] - lo
        ok = np.all((src >= 0) & (src < span), axis=1)
        idx = np.full(mtxd, -1)
        idx[ok] = lut[tuple(src[ok].T)]
        sel = idx >= 0
        out[sel] += v * psi[idx[sel]]
    nanl = d["nanl"]
    a = d["anlga"][:mtxd, :nanl]
    out += a @ (d["xnlkb"][:nanl, None] * (a.conj().T @ psi))
    return out


def ref_ditsp_c16(a):
    """The reference ditsp_c16 (Src/ditsp/ditsp_c16.f90:20-27, flang build
    with the vendored BLAS / LAPACK) on the argument dict ``a`` (the names of
    the Fortran dummies; psi, hpsi, ei written in place).  Returns ifail."""
    lib = ref_lib()
    i = lambda v: C.byref(C.c_int(int(v)))
    ifail = C.c_int(0)
    lnew = C.c_int(1)
    neig = C.c_int(int(a["neig"]))
    eps = C.c_double(float(a["epspsi"]))
    lib.ditsp_c16_(i(a["ipr"]), C.byref(ifail), i(a["icmax"]), i(a["iguess"]), C.byref(eps), C.byref(lnew),
                   C.byref(neig), i(a["mtxd"]), i(a["mtxds"]), _ptr(a["psi"]), _ptr(a["hpsi"]), _ptr(a["ei"]),
                   _ptr(a["ekpg"]), _ptr(a["isort"]), _ptr(a["vscr"]), _ptr(a["kmscr"]), i(a["ng"]), _ptr(a["kgv"]),
                   _ptr(a["anlga"]), _ptr(a["xnlkb"]), i(a["nanl"]), _ptr(a["hamsm"]), _ptr(a["hdiag"]),
                   i(a["mxddim"]), i(a["mxdsml"]), i(a["mxdbnd"]), i(a["mxdgve"]), i(a["mxdscr"]), i(a["mxdanl"]))
    return ifail.value
