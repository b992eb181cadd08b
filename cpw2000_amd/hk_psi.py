"""Python mirror of the cpw2000 ``hk_psi_c16`` interface over the C ABI.

``hk_psi_c16`` takes the Fortran subroutine's arguments, in its order and
meaning (Src/hk_psi/reference/hk_psi_c16.f90:28-78), as numpy arrays in
Fortran layout, and fills ``hpsi`` in place.  Where the reference stops the
program (ng > mxdgve, FFT work size > mxdscr, sphere wrap-around) this raises
``CpwGpuError`` with the library's status code.

``Plan`` is the device-resident path: index maps, V(r) and projectors are set
once per k-point and H is applied to torch tensors already in HBM.

The library is loaded from the package directory; there is no fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libcpwgpu.so")
FLIB_PATH = os.path.join(PKG, "libcpwgpu_fortran.so")

OK = 0
ERR_NG = 1
ERR_MXDSCR = 2
ERR_WRAP = 3
ERR_MTXD = 4
ERR_MODE = 5
ERR_SIZE = 6
ERR_ARG = 7
ERR_NOTPD = 8
ERR_HIP = 10
ERR_NOMEM = 11

PREC_FP64 = 0
PREC_FP32 = 1

# every symbol declared in include/cpwgpu.h
EXPORTS = (
    "cpwgpu_init", "cpwgpu_finalize", "cpwgpu_last_error", "cpwgpu_set_precision",
    "cpwgpu_set_overlap", "cpwgpu_set_ditsp_method",
    "cpwgpu_hk_psi_c16", "cpwgpu_plan_create", "cpwgpu_plan_destroy",
    "cpwgpu_plan_set_potential", "cpwgpu_plan_set_projectors",
    "cpwgpu_plan_set_projectors_device", "cpwgpu_plan_apply", "cpwgpu_plan_info",
    "cpwgpu_profile", "cpwgpu_profile_read", "cpwgpu_rayleigh_ritz", "cpwgpu_plan_forces",
    "cpwgpu_graph_create", "cpwgpu_graph_launch", "cpwgpu_graph_destroy",
    "cpwgpu_plan_set_variant", "cpwgpu_plan_get_variant", "cpwgpu_plan_tune", "cpwgpu_plan_tune_spin",
    "cpwgpu_plan_ditsp", "cpwgpu_ditsp_c16", "cpwgpu_plan_density",
    "cpwgpu_hk_psi_spin_c16", "cpwgpu_plan_set_potential_spin", "cpwgpu_plan_set_projectors_spin",
    "cpwgpu_plan_apply_spin", "cpwgpu_hk_psi_c8", "cpwgpu_hk_psi_allk_c16",
    "cpwgpu_proj_nl_kb_c16", "cpwgpu_plan_build_projectors_kb", "cpwgpu_plan_apply_noalloc",
)

# tunables of a plan (CPWGPU_TUNE_* in include/cpwgpu.h)
TUNE_PLANE, TUNE_SPLITK, TUNE_PROJ, TUNE_BACK, TUNE_KIN, TUNE_SPIN, TUNE_COL, TUNE_SCHED = range(8)
# plane-stage kernels (values of TUNE_PLANE): 0 LDS two-pass, 1-3 register
# lines, 4-6 LDS two-pass exchanged within one wavefront, 7 register x rows
# with wave-exchanged LDS y lines, 8-11 register shapes that skip the slots
# outside the sphere box (meshes with the box inside N/4)
NPLANE_VAR = 12
# projection kernels (TUNE_PROJ): 0 LDS-staged, 1-4 direct-load shapes, 5
# LDS-staged with half-size wave tiles; back-projection kernels (TUNE_BACK):
# 0 LDS-staged, 1-3 direct-load shapes, 4 LDS-staged with half-size chunks
NPROJ_VAR = 6
NBACK_VAR = 5

# kernel ids of the measurement hook (CPWGPU_K_* in include/cpwgpu.h)
KERNELS = ("k_col_bwd", "k_plane", "k_col_fwd", "k_proj", "k_proj_reduce", "k_back", "k_proj_kb")


class CpwGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cpwgpu status {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded libcpwgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # PyTorch bundles its own libamdhip64.so.7 (same soname as the system
        # one this library links).  Loading torch first makes the whole
        # process use a single HIP runtime, shared with torch tensors/streams.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m cpw2000_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.c_void_p
        L.cpwgpu_last_error.restype = C.c_char_p
        for name in EXPORTS:
            if name != "cpwgpu_last_error":
                getattr(L, name).restype = C.c_int
        L.cpwgpu_plan_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int]
        L.cpwgpu_plan_destroy.argtypes = [vp]
        L.cpwgpu_plan_set_potential.argtypes = [vp, dp, C.c_int]
        L.cpwgpu_plan_set_projectors.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        L.cpwgpu_plan_set_projectors_device.argtypes = [vp, dp, C.c_long, dp, C.c_int]
        L.cpwgpu_plan_build_projectors_kb.argtypes = [vp, dp, C.c_int, vp, dp, dp, vp, C.c_int, vp, dp, dp,
                                                      C.c_int, C.c_int, C.c_int]
        L.cpwgpu_plan_apply.argtypes = [vp, C.c_int, dp, C.c_long, dp, C.c_long, C.c_int, vp]
        L.cpwgpu_plan_apply_noalloc.argtypes = [vp, C.c_int, dp, C.c_long, dp, C.c_long, C.c_int, vp]
        L.cpwgpu_plan_info.argtypes = [vp, ip]
        L.cpwgpu_profile.argtypes = [C.c_int, C.c_uint]
        L.cpwgpu_rayleigh_ritz.argtypes = [C.c_int, C.c_int, dp, C.c_long, dp, C.c_long, C.c_int,
                                           dp, dp, vp]
        L.cpwgpu_plan_forces.argtypes = [vp, C.c_int, dp, C.c_long, dp, C.c_int, dp, dp, vp, C.c_int,
                                         vp, C.c_int, vp, dp, dp, dp]
        L.cpwgpu_graph_create.argtypes = [C.POINTER(vp), C.c_int, vp, vp, vp, vp, vp, vp, C.c_int]
        L.cpwgpu_graph_launch.argtypes = [vp, vp]
        L.cpwgpu_graph_destroy.argtypes = [vp]
        L.cpwgpu_plan_set_variant.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.cpwgpu_plan_get_variant.argtypes = [vp, C.c_int, C.c_int, ip]
        L.cpwgpu_plan_tune.argtypes = [vp, C.c_int, dp, C.c_long, dp, C.c_long, C.c_int, vp]
        L.cpwgpu_profile_read.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_long), C.c_int]
        L.cpwgpu_plan_density.argtypes = [vp, C.c_int, dp, C.c_long, dp, dp, vp]
        L.cpwgpu_plan_set_potential_spin.argtypes = [vp, dp, C.c_int, C.c_int]
        L.cpwgpu_plan_set_projectors_spin.argtypes = [vp, dp, dp, C.c_int, C.c_int]
        L.cpwgpu_plan_apply_spin.argtypes = [vp, C.c_int, dp, C.c_long, dp, C.c_long, C.c_int, vp]
        L.cpwgpu_plan_ditsp.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int,
                                        dp, C.c_long, dp, dp, C.c_long, dp, C.c_long, dp, ip, ip, vp]
        _lib = L
    return _lib


def _check(st):
    if st != OK:
        raise CpwGpuError(st, lib().cpwgpu_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _iref(v):
    return C.byref(C.c_int(int(v)))


def set_precision(prec):
    _check(lib().cpwgpu_set_precision(int(prec)))


DITSP_BLOCK, DITSP_CHEB = 0, 1


def set_ditsp_method(method):
    """Iteration of Plan.ditsp / ditsp_c16: DITSP_CHEB (Chebyshev-filtered
    subspace iteration, default) or DITSP_BLOCK (locally optimal block
    iteration)."""
    _check(lib().cpwgpu_set_ditsp_method(int(method)))


def set_overlap(on):
    """Non-local GEMMs on a second stream beside the FFT kernels (default) or
    every kernel in sequence on the caller's stream."""
    _check(lib().cpwgpu_set_overlap(1 if on else 0))


def hk_psi_c16(mtxd, neig, psi, hpsi, lnewanl, ng, kgv, ekpg, isort, vscr, kmscr, anlga, xnlkb,
               nanl, mxddim, mxdbnd, mxdanl, mxdgve, mxdscr):
    """H|psi> with the arguments of Src/hk_psi/reference/hk_psi_c16.f90:28-32.

    Arrays are numpy, Fortran layout: psi/hpsi complex128 (mxddim, >=neig),
    kgv int32 (3, mxdgve), anlga complex128 (mxddim, mxdanl).  hpsi is
    written in place (rows 0..mtxd-1 of the first neig columns)."""
    for name, arr, dt in (("psi", psi, np.complex128), ("hpsi", hpsi, np.complex128),
                          ("kgv", kgv, np.int32), ("ekpg", ekpg, np.float64),
                          ("isort", isort, np.int32), ("vscr", vscr, np.float64),
                          ("kmscr", kmscr, np.int32), ("anlga", anlga, np.complex128),
                          ("xnlkb", xnlkb, np.float64)):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{name} must be a Fortran-contiguous {np.dtype(dt).name} array")
    lnew = C.c_int(1 if lnewanl else 0)
    st = lib().cpwgpu_hk_psi_c16(
        _iref(mtxd), _iref(neig), _p(psi), _p(hpsi), C.byref(lnew), _iref(ng), _p(kgv), _p(ekpg),
        _p(isort), _p(vscr), _p(kmscr), _p(anlga), _p(xnlkb), _iref(nanl), _iref(mxddim),
        _iref(mxdbnd), _iref(mxdanl), _iref(mxdgve), _iref(mxdscr))
    _check(st)
    return hpsi


def _kb_tables(nqnl, delqnl, vkb, nkb, natom, rat, adot):
    """The pseudopotential arrays of proj_nl_kb_c16 as Fortran-ordered numpy
    arrays of the reference's kinds; vkb (mxdlqp + 3, 4, 3, mxdtyp) with its
    first index shifted by 2, nkb (4, 3, mxdtyp), rat (3, mxdatm, mxdtyp)."""
    out = []
    for name, arr, dt, nd in (("nqnl", nqnl, np.int32, 1), ("delqnl", delqnl, np.float64, 1),
                              ("vkb", vkb, np.float64, 4), ("nkb", nkb, np.int32, 3),
                              ("natom", natom, np.int32, 1), ("rat", rat, np.float64, 3),
                              ("adot", adot, np.float64, 2)):
        a = np.asfortranarray(arr, dt)
        if a.ndim != nd:
            raise ValueError(f"{name} must have {nd} dimensions")
        out.append(a)
    nq, dq, v, nk, na, r, ad = out
    if v.shape[1:3] != (4, 3) or nk.shape[:2] != (4, 3) or ad.shape != (3, 3) or r.shape[0] != 3:
        raise ValueError("vkb (mxdlqp+3, 4, 3, mxdtyp), nkb (4, 3, mxdtyp), rat (3, mxdatm, mxdtyp), adot (3, 3)")
    mxdtyp = v.shape[3]
    if nk.shape[2] != mxdtyp or r.shape[2] != mxdtyp or nq.shape[0] < mxdtyp or dq.shape[0] < mxdtyp \
            or na.shape[0] < mxdtyp:
        raise ValueError("per-type arrays must cover mxdtyp types")
    return nq, dq, v, nk, na, r, ad, mxdtyp, r.shape[1], v.shape[0] - 3


def proj_nl_kb_c16(rkpt, mtxd, isort, nanl, ng, kgv, nqnl, delqnl, vkb, nkb, ntype, natom, rat, adot,
                   anlga, xnlkb, mxddim, mxdanl, mxdgve):
    """KB projectors with the arguments of Src/omp/proj_nl_kb_c16.f90:23-28,
    computed on the GPU (cpwgpu_proj_nl_kb_c16).  anlga complex128 (mxddim,
    mxdanl) and xnlkb float64 (mxdanl,) are written in place (rows 0..mtxd-1,
    columns 0..nanl-1); mxdtyp, mxdatm and mxdlqp come from the shapes of
    vkb (mxdlqp + 3, 4, 3, mxdtyp) and rat (3, mxdatm, mxdtyp)."""
    nq, dq, v, nk, na, r, ad, mxdtyp, mxdatm, mxdlqp = _kb_tables(nqnl, delqnl, vkb, nkb, natom, rat, adot)
    for name, arr, dt in (("isort", isort, np.int32), ("kgv", kgv, np.int32),
                          ("anlga", anlga, np.complex128), ("xnlkb", xnlkb, np.float64)):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{name} must be a Fortran-contiguous {np.dtype(dt).name} array")
    if anlga.ndim != 2 or anlga.shape[0] != mxddim or anlga.shape[1] < min(nanl, mxdanl) \
            or xnlkb.size < min(nanl, mxdanl) or isort.size < mtxd or kgv.shape[0] != 3 \
            or kgv.shape[1] < min(ng, mxdgve):
        raise ValueError("array shapes do not match mxddim, mxdanl, nanl, mtxd, ng")
    rk = np.ascontiguousarray(rkpt, np.float64)
    if rk.size != 3:
        raise ValueError("rkpt must have 3 components")
    st = lib().cpwgpu_proj_nl_kb_c16(
        _p(rk), _iref(mtxd), _p(isort), _iref(nanl), _iref(ng), _p(kgv), _p(nq), _p(dq), _p(v), _p(nk),
        _iref(ntype), _p(na), _p(r), _p(ad), _p(anlga), _p(xnlkb), _iref(mxdtyp), _iref(mxdatm),
        _iref(mxdlqp), _iref(mxddim), _iref(mxdanl), _iref(mxdgve))
    _check(st)
    return anlga, xnlkb


def hk_psi_allk_c16(nrk, mtxd_allk, nband, psi_allk, hpsi_allk, ng, kgv, ekpg_allk, isort_allk, vscr,
                    kmscr, anlga_allk, xnlkb_allk, nanl_allk, mxddim, mxdbnd, mxdanl, mxdgve, mxdscr, mxdnrk):
    """H|psi> for every k-point of a self-consistency step (cpwgpu_hk_psi_allk_c16),
    with the all-k arrays of the k-point loop of Src/cpw/cpw_scf.f90:488-544:
    psi_allk/hpsi_allk complex128 (mxddim, mxdbnd, mxdnrk), isort_allk int32 and
    ekpg_allk float64 (mxddim, mxdnrk), anlga_allk complex128 (mxddim, mxdanl,
    mxdnrk), xnlkb_allk float64 (mxdanl, mxdnrk), mtxd_allk/nband/nanl_allk int32
    (mxdnrk,); numpy, Fortran layout.  hpsi_allk is written in place (rows
    0..mtxd_allk[k]-1 of the first nband[k] columns of each k-point)."""
    for name, arr, dt in (("mtxd_allk", mtxd_allk, np.int32), ("nband", nband, np.int32),
                          ("psi_allk", psi_allk, np.complex128), ("hpsi_allk", hpsi_allk, np.complex128),
                          ("kgv", kgv, np.int32), ("ekpg_allk", ekpg_allk, np.float64),
                          ("isort_allk", isort_allk, np.int32), ("vscr", vscr, np.float64),
                          ("kmscr", kmscr, np.int32), ("anlga_allk", anlga_allk, np.complex128),
                          ("xnlkb_allk", xnlkb_allk, np.float64), ("nanl_allk", nanl_allk, np.int32)):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{name} must be a Fortran-contiguous {np.dtype(dt).name} array")
    _check(lib().cpwgpu_hk_psi_allk_c16(
        _iref(nrk), _p(mtxd_allk), _p(nband), _p(psi_allk), _p(hpsi_allk), _iref(ng), _p(kgv),
        _p(ekpg_allk), _p(isort_allk), _p(vscr), _p(kmscr), _p(anlga_allk), _p(xnlkb_allk),
        _p(nanl_allk), _iref(mxddim), _iref(mxdbnd), _iref(mxdanl), _iref(mxdgve), _iref(mxdscr),
        _iref(mxdnrk)))
    return hpsi_allk


def hk_psi_c8(mtxd, neig, psi, hpsi, lnewanl, ng, kgv, ekpg, isort, vscr, kmscr, anlga, xnlkb,
              nanl, mxddim, mxdbnd, mxdanl, mxdgve, mxdscr):
    """Single-precision H|psi> (cpwgpu_hk_psi_c8): the arguments of hk_psi_c16
    with complex64 psi/hpsi/anlga and float32 ekpg/vscr/xnlkb (numpy, Fortran
    layout).  hpsi is written in place."""
    for name, arr, dt in (("psi", psi, np.complex64), ("hpsi", hpsi, np.complex64),
                          ("kgv", kgv, np.int32), ("ekpg", ekpg, np.float32),
                          ("isort", isort, np.int32), ("vscr", vscr, np.float32),
                          ("kmscr", kmscr, np.int32), ("anlga", anlga, np.complex64),
                          ("xnlkb", xnlkb, np.float32)):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{name} must be a Fortran-contiguous {np.dtype(dt).name} array")
    lnew = C.c_int(1 if lnewanl else 0)
    _check(lib().cpwgpu_hk_psi_c8(
        _iref(mtxd), _iref(neig), _p(psi), _p(hpsi), C.byref(lnew), _iref(ng), _p(kgv), _p(ekpg),
        _p(isort), _p(vscr), _p(kmscr), _p(anlga), _p(xnlkb), _iref(nanl), _iref(mxddim),
        _iref(mxdbnd), _iref(mxdanl), _iref(mxdgve), _iref(mxdscr)))
    return hpsi


def hk_psi_spin_c16(mtxd, neig, psi_sp, hpsi_sp, lnewanl, ng, kgv, ekpg, isort, vscr_sp, kmscr,
                    nsp, anlsp, xnlkbsp, nanlsp, mxddim, mxdbnd, mxdasp, mxdgve, mxdscr, mxdnsp):
    """Two-component H|psi> with the arguments of
    Src/hk_psi/reference/hk_psi_spin_c16.f90:28-32 (numpy, Fortran layout):
    psi_sp/hpsi_sp complex128 (2*mxddim, >=neig) with up/down rows
    interleaved, vscr_sp float64 (mxdscr, mxdnsp), anlsp complex128
    (2*mxddim, mxdasp).  hpsi_sp is written in place."""
    for name, arr, dt in (("psi_sp", psi_sp, np.complex128), ("hpsi_sp", hpsi_sp, np.complex128),
                          ("kgv", kgv, np.int32), ("ekpg", ekpg, np.float64),
                          ("isort", isort, np.int32), ("vscr_sp", vscr_sp, np.float64),
                          ("kmscr", kmscr, np.int32), ("anlsp", anlsp, np.complex128),
                          ("xnlkbsp", xnlkbsp, np.float64)):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{name} must be a Fortran-contiguous {np.dtype(dt).name} array")
    lnew = C.c_int(1 if lnewanl else 0)
    _check(lib().cpwgpu_hk_psi_spin_c16(
        _iref(mtxd), _iref(neig), _p(psi_sp), _p(hpsi_sp), C.byref(lnew), _iref(ng), _p(kgv),
        _p(ekpg), _p(isort), _p(vscr_sp), _p(kmscr), _iref(nsp), _p(anlsp), _p(xnlkbsp),
        _iref(nanlsp), _iref(mxddim), _iref(mxdbnd), _iref(mxdasp), _iref(mxdgve), _iref(mxdscr),
        _iref(mxdnsp)))
    return hpsi_sp


def ditsp_c16(ipr, icmax, iguess, epspsi, lnewanl, neig, mtxd, mtxds, psi, hpsi, ei, ekpg,
              isort, vscr, kmscr, ng, kgv, anlga, xnlkb, nanl, hamsm, hdiag, mxddim, mxdsml,
              mxdbnd, mxdgve, mxdscr, mxdanl):
    """Iterative diagonalisation with the arguments of Src/ditsp/ditsp_c16.f90:20-27
    (ifail, the reference's first output, is returned instead of passed).

    Arrays are numpy, Fortran layout: psi/hpsi complex128 (mxddim, >=neig)
    (psi holds guesses when iguess == 1), ei float64 (>=neig), hamsm
    complex128 (mxdsml, mxdsml) holding the leading mtxds x mtxds block of H,
    hdiag float64 (mxddim).  psi, hpsi and ei are written in place."""
    for name, arr, dt in (("psi", psi, np.complex128), ("hpsi", hpsi, np.complex128),
                          ("ei", ei, np.float64), ("kgv", kgv, np.int32),
                          ("ekpg", ekpg, np.float64), ("isort", isort, np.int32),
                          ("vscr", vscr, np.float64), ("kmscr", kmscr, np.int32),
                          ("anlga", anlga, np.complex128), ("xnlkb", xnlkb, np.float64),
                          ("hamsm", hamsm, np.complex128), ("hdiag", hdiag, np.float64)):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{name} must be a Fortran-contiguous {np.dtype(dt).name} array")
    ifail = C.c_int(0)
    lnew = C.c_int(1 if lnewanl else 0)
    nref = C.c_int(int(neig))
    eps = C.c_double(float(epspsi))
    _check(lib().cpwgpu_ditsp_c16(
        _iref(ipr), C.byref(ifail), _iref(icmax), _iref(iguess), C.byref(eps), C.byref(lnew),
        C.byref(nref), _iref(mtxd), _iref(mtxds), _p(psi), _p(hpsi), _p(ei), _p(ekpg), _p(isort),
        _p(vscr), _p(kmscr), _iref(ng), _p(kgv), _p(anlga), _p(xnlkb), _iref(nanl), _p(hamsm),
        _p(hdiag), _iref(mxddim), _iref(mxdsml), _iref(mxdbnd), _iref(mxdgve), _iref(mxdscr),
        _iref(mxdanl)))
    return ifail.value


def hk_psi_c16_dict(d, psi=None, hpsi=None, neig=None):
    """hk_psi_c16 on a dictionary of inputs (keys named as the Fortran dummies)."""
    psi = np.asfortranarray(d["psi"] if psi is None else psi, np.complex128)
    neig = psi.shape[1] if neig is None else neig
    if hpsi is None:
        hpsi = np.zeros_like(psi, order="F")
    return hk_psi_c16(d["mtxd"], neig, psi, hpsi, True, d["ng"],
                      np.asfortranarray(d["kgv"], np.int32), np.asfortranarray(d["ekpg"]),
                      np.asfortranarray(d["isort"], np.int32), np.asfortranarray(d["vscr"]),
                      np.asfortranarray(d["kmscr"], np.int32),
                      np.asfortranarray(d["anlga"], np.complex128), np.asfortranarray(d["xnlkb"]),
                      d["nanl"], d["mxddim"], max(psi.shape[1], 1), max(d["mxdanl"], 1),
                      d["mxdgve"], d["mxdscr"])


def profile(on, kernels=None):
    """Per-kernel HIP-event timing of later applies (kernels: names of KERNELS)."""
    mask = 0
    for k in (KERNELS if kernels is None else kernels):
        mask |= 1 << KERNELS.index(k)
    _check(lib().cpwgpu_profile(1 if on else 0, mask))


def profile_read():
    """{kernel: (summed ms, launches)} since the previous read."""
    n = len(KERNELS)
    ms = (C.c_double * n)()
    cnt = (C.c_long * n)()
    _check(lib().cpwgpu_profile_read(ms, cnt, n))
    return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNELS)}


def _ld(t):
    """Leading dimension of a (bands, ld) tensor whose rows are contiguous."""
    if t.dim() != 2 or (t.shape[1] > 1 and t.stride(1) != 1):
        raise ValueError("expected a 2-D tensor with contiguous rows (one band per row)")
    return t.shape[1] if t.shape[0] == 1 else t.stride(0)


def rayleigh_ritz(psi_t, hpsi_t, rotate=True, residuals=True, stream=None):
    """FP64 Rayleigh-Ritz of the bands in psi_t with hpsi_t = H psi_t (complex128
    device tensors shaped (nb, ld), rows 0..mtxd-1 used with mtxd = the size
    of dimension 1).  Cholesky orthonormalisation, Jacobi eigensolve; with
    rotate the tensors become the Ritz vectors and H applied to them.
    Returns (eig, res) numpy arrays (res = ||H psi_n - eig_n psi_n||^2)."""
    import torch

    nb, mtxd = psi_t.shape
    eig = np.zeros(nb)
    res = np.zeros(nb) if residuals else None
    if stream is None:
        stream = torch.cuda.current_stream(psi_t.device).cuda_stream
    _check(lib().cpwgpu_rayleigh_ritz(
        int(mtxd), int(nb), C.c_void_p(psi_t.data_ptr()), _ld(psi_t),
        C.c_void_p(hpsi_t.data_ptr()), _ld(hpsi_t), 1 if rotate else 0,
        eig.ctypes.data_as(C.c_void_p), res.ctypes.data_as(C.c_void_p) if residuals else None,
        C.c_void_p(stream)))
    return eig, res


class Plan:
    """One k-point resident on the device (see cpwgpu_plan_* in include/cpwgpu.h)."""

    def __init__(self, mtxd, ng, kgv, ekpg, isort, kmscr, mxddim=None, mxdgve=None):
        self._h = C.c_void_p()
        kgv = np.asfortranarray(kgv, np.int32)
        ekpg = np.ascontiguousarray(ekpg, np.float64)
        isort = np.ascontiguousarray(isort, np.int32)
        kmscr = np.ascontiguousarray(kmscr, np.int32)
        mxddim = mtxd if mxddim is None else mxddim
        mxdgve = ng if mxdgve is None else mxdgve
        _check(lib().cpwgpu_plan_create(C.byref(self._h), int(mtxd), int(ng), _p(kgv), _p(ekpg),
                                        _p(isort), _p(kmscr), int(mxddim), int(mxdgve)))
        self.mtxd = int(mtxd)

    @classmethod
    def from_inputs(cls, d, projectors=True):
        p = cls(d["mtxd"], d["ng"], d["kgv"], d["ekpg"], d["isort"], d["kmscr"], d["mxddim"],
                d["mxdgve"])
        p.set_potential(d["vscr"])
        if projectors:
            p.set_projectors(d["anlga"], d["xnlkb"], d["nanl"])
        return p

    def info(self):
        out = (C.c_int * 10)()
        _check(lib().cpwgpu_plan_info(self._h, out))
        keys = ("mtxd", "n1", "n2", "n3", "id", "kd1", "kd2", "kd3", "ncol", "nanl")
        return dict(zip(keys, list(out)))

    def set_variant(self, prec, what, value):
        """Kernel choice for a precision: what = TUNE_PLANE (plane-stage kernel,
        0 .. NPLANE_VAR-1 as listed at NPLANE_VAR, -1 default), TUNE_PROJ
        (projection GEMM, 0 LDS-staged, 1..4 direct-load shapes, 5 LDS-staged
        with half-size wave tiles), TUNE_BACK (back-projection GEMM, 0
        LDS-staged, 1..3 direct-load shapes, 4 LDS-staged with half-size chunks),
        TUNE_SPLITK (projection GEMM split-K target blocks, -1 default),
        TUNE_KIN (kinetic term in the back-projection epilogue 0 or in the
        column gather 1), TUNE_SPIN (plane kernel of apply_spin with nsp = 4:
        0 LDS two-pass, 1 register lines where built, -1 default), TUNE_COL
        (z-column kernels: 0 LDS two-pass, 1 register lines where built) or
        TUNE_SCHED (stream schedule: 0 the whole non-local term beside the
        local one, 1 projection beside it and back-projection after it, 2 the
        non-local term on the caller's stream with the column and plane
        kernels beside it)."""
        _check(lib().cpwgpu_plan_set_variant(self._h, int(prec), int(what), int(value)))

    def get_variant(self, prec, what):
        out = C.c_int(0)
        _check(lib().cpwgpu_plan_get_variant(self._h, int(prec), int(what), C.byref(out)))
        return out.value

    def tune(self, psi_t, hpsi_t, prec=PREC_FP32, stream=None):
        """Autotune the plan's kernels on these device arrays (hpsi_t receives
        H psi); returns the choice {"plane", "proj", "back", "splitk", "kin", "col",
        "sched"}."""
        import torch

        neig = psi_t.shape[0]
        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        _check(lib().cpwgpu_plan_tune(self._h, int(neig), C.c_void_p(psi_t.data_ptr()), _ld(psi_t),
                                      C.c_void_p(hpsi_t.data_ptr()), _ld(hpsi_t), int(prec),
                                      C.c_void_p(stream)))
        return {"plane": self.get_variant(prec, TUNE_PLANE),
                "proj": self.get_variant(prec, TUNE_PROJ),
                "back": self.get_variant(prec, TUNE_BACK),
                "splitk": self.get_variant(prec, TUNE_SPLITK),
                "kin": self.get_variant(prec, TUNE_KIN),
                "col": self.get_variant(prec, TUNE_COL),
                "sched": self.get_variant(prec, TUNE_SCHED)}

    def set_potential(self, vscr):
        v = np.ascontiguousarray(vscr, np.float64)
        _check(lib().cpwgpu_plan_set_potential(self._h, _p(v), int(v.shape[0])))

    def set_projectors(self, anlga, xnlkb, nanl=None):
        a = np.asfortranarray(anlga, np.complex128)
        x = np.ascontiguousarray(xnlkb, np.float64)
        nanl = a.shape[1] if nanl is None else nanl
        _check(lib().cpwgpu_plan_set_projectors(self._h, _p(a), _p(x), int(nanl), int(a.shape[0])))

    def build_projectors_kb(self, rkpt, nanl, nqnl, delqnl, vkb, nkb, ntype, natom, rat, adot):
        """The projectors of proj_nl_kb_c16 (Src/omp/proj_nl_kb_c16.f90) built on
        the device for this plan's plane waves (cpwgpu_plan_build_projectors_kb);
        the arrays as for proj_nl_kb_c16 above."""
        nq, dq, v, nk, na, r, ad, mxdtyp, mxdatm, mxdlqp = _kb_tables(nqnl, delqnl, vkb, nkb, natom, rat, adot)
        rk = np.ascontiguousarray(rkpt, np.float64)
        if rk.size != 3:
            raise ValueError("rkpt must have 3 components")
        _check(lib().cpwgpu_plan_build_projectors_kb(
            self._h, _p(rk), int(nanl), _p(nq), _p(dq), _p(v), _p(nk), int(ntype), _p(na), _p(r), _p(ad),
            int(mxdtyp), int(mxdatm), int(mxdlqp)))

    def set_projectors_device(self, anlga_t, xnlkb_t):
        """anlga_t: complex128 device tensor (nanl, lda) (each projector contiguous)."""
        nanl, lda = anlga_t.shape
        _check(lib().cpwgpu_plan_set_projectors_device(
            self._h, C.c_void_p(anlga_t.data_ptr()), int(anlga_t.stride(0)),
            C.c_void_p(xnlkb_t.data_ptr()), int(nanl)))

    def apply(self, psi_t, hpsi_t, prec=PREC_FP64, stream=None):
        """H|psi> on complex128 device tensors shaped (neig, ld) (band-major,
        each band contiguous: the column-major psi(ld, neig) of the reference)."""
        import torch

        neig = psi_t.shape[0]
        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        _check(lib().cpwgpu_plan_apply(self._h, int(neig), C.c_void_p(psi_t.data_ptr()),
                                       _ld(psi_t), C.c_void_p(hpsi_t.data_ptr()), _ld(hpsi_t),
                                       int(prec), C.c_void_p(stream)))
        return hpsi_t

    def apply_noalloc(self, psi_t, hpsi_t, prec=PREC_FP64, stream=None):
        """apply() with the library's fixed per-device workspace instead of the
        plan's buffers: no device allocation (cpwgpu_plan_apply_noalloc)."""
        import torch

        neig = psi_t.shape[0]
        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        _check(lib().cpwgpu_plan_apply_noalloc(self._h, int(neig), C.c_void_p(psi_t.data_ptr()),
                                               _ld(psi_t), C.c_void_p(hpsi_t.data_ptr()), _ld(hpsi_t),
                                               int(prec), C.c_void_p(stream)))
        return hpsi_t

    def set_potential_spin(self, vscr_sp, nsp):
        """Potential components vscr_sp (mxdscr, >=nsp), nsp = 1, 2 or 4."""
        v = np.asfortranarray(vscr_sp, np.float64)
        _check(lib().cpwgpu_plan_set_potential_spin(self._h, _p(v), int(v.shape[0]), int(nsp)))

    def set_projectors_spin(self, anlsp, xnlkbsp, nanlsp=None):
        """Projectors on the interleaved rows: anlsp (2*mxddim, >=nanlsp)."""
        a = np.asfortranarray(anlsp, np.complex128)
        x = np.ascontiguousarray(xnlkbsp, np.float64)
        nanlsp = a.shape[1] if nanlsp is None else nanlsp
        _check(lib().cpwgpu_plan_set_projectors_spin(self._h, _p(a), _p(x), int(nanlsp),
                                                     int(a.shape[0] // 2)))

    def apply_spin(self, psi_t, hpsi_t, prec=PREC_FP64, stream=None):
        """H psi for two-component bands: complex128 device tensors (neig, ld),
        ld >= 2*mtxd, up/down rows interleaved."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        _check(lib().cpwgpu_plan_apply_spin(self._h, int(psi_t.shape[0]), C.c_void_p(psi_t.data_ptr()),
                                            _ld(psi_t), C.c_void_p(hpsi_t.data_ptr()), _ld(hpsi_t),
                                            int(prec), C.c_void_p(stream)))
        return hpsi_t

    def tune_spin(self, psi_t, hpsi_t, prec=PREC_FP32, stream=None):
        """Autotune the two-component path on these device arrays (as
        apply_spin; hpsi_t receives H psi); returns the choice {"spin",
        "proj", "back", "splitk"}."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        _check(lib().cpwgpu_plan_tune_spin(self._h, int(psi_t.shape[0]), C.c_void_p(psi_t.data_ptr()),
                                           _ld(psi_t), C.c_void_p(hpsi_t.data_ptr()), _ld(hpsi_t), int(prec),
                                           C.c_void_p(stream)))
        return {"spin": self.get_variant(prec, TUNE_SPIN),
                "proj": self.get_variant(prec, TUNE_PROJ),
                "back": self.get_variant(prec, TUNE_BACK),
                "splitk": self.get_variant(prec, TUNE_SPLITK)}

    def density(self, psi_t, occ=None, stream=None):
        """rho(r) = sum_n occ_n |psi_n(r)|^2 on the mesh (cpwgpu_plan_density):
        a float64 device tensor of id*n2*n3 values in the vscr layout."""
        import torch

        info = self.info()
        rho = torch.empty(info["id"] * info["n2"] * info["n3"], dtype=torch.float64,
                          device=psi_t.device)
        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        occ_a = None if occ is None else np.ascontiguousarray(occ, np.float64)
        _check(lib().cpwgpu_plan_density(
            self._h, int(psi_t.shape[0]), C.c_void_p(psi_t.data_ptr()), _ld(psi_t),
            None if occ_a is None else _p(occ_a), C.c_void_p(rho.data_ptr()), C.c_void_p(stream)))
        return rho

    def ditsp(self, psi_t, hpsi_t, prec=PREC_FP64, icmax=100, iguess=0, epspsi=1e-6, mtxds=0,
              hamsm_t=None, hdiag_t=None, stream=None):
        """Lowest eigenpairs of H (cpwgpu_plan_ditsp): psi_t / hpsi_t are
        complex128 device tensors (neig, ld); psi_t holds guesses when iguess
        == 1 and gets the eigenvectors, hpsi_t H times them.  Returns (ei,
        ifail, stats) with stats = dict(iters, hpsi_products, nvec, mtxds,
        fp64_from)."""
        import torch

        neig = psi_t.shape[0]
        ei = np.zeros(neig)
        ifail = C.c_int(0)
        stats = (C.c_int * 5)()
        if stream is None:
            stream = torch.cuda.current_stream(psi_t.device).cuda_stream
        hs = None if hamsm_t is None else C.c_void_p(hamsm_t.data_ptr())
        ldhs = 0 if hamsm_t is None else _ld(hamsm_t)
        hd = None if hdiag_t is None else C.c_void_p(hdiag_t.data_ptr())
        _check(lib().cpwgpu_plan_ditsp(
            self._h, int(prec), int(icmax), int(iguess), float(epspsi), int(neig), int(mtxds), hs,
            ldhs, hd, C.c_void_p(psi_t.data_ptr()), _ld(psi_t), C.c_void_p(hpsi_t.data_ptr()),
            _ld(hpsi_t), ei.ctypes.data_as(C.c_void_p), C.byref(ifail),
            C.cast(stats, C.POINTER(C.c_int)), C.c_void_p(stream)))
        keys = ("iters", "hpsi_products", "nvec", "mtxds", "fp64_from")
        return ei, ifail.value, dict(zip(keys, list(stats)))

    def forces(self, psi_t, rat, avec, proj_atom, floc_q=None, floc_f=None, species=None,
               occ=None):
        """Cartesian forces (natom, 3) Ha/bohr and E_NL for E = sum_n occ_n
        <psi_n|V_loc + V_NL|psi_n> at fixed psi (psi_t: complex128 (nb, ld) on
        cuda).  rat (natom, 3) fractional, avec (3, 3) lattice vectors as
        columns, proj_atom (nanl,) atom of each projector, floc_q (nq, 3) and
        floc_f (nq, nspec) the position-dependent local potential
        V(q) = sum_s f_s(q) sum_{atoms of s} exp(-2 pi i q.rat), species (natom,)."""
        import torch

        torch.cuda.synchronize()
        nb = psi_t.shape[0]
        rat = np.ascontiguousarray(rat, np.float64)
        natom = rat.shape[0]
        av = np.asfortranarray(avec, np.float64)
        pa = np.ascontiguousarray(proj_atom, np.int32)
        force = np.zeros((natom, 3))
        enl = C.c_double(0.0)
        if floc_q is not None and len(floc_q):
            q = np.ascontiguousarray(floc_q, np.int32)
            f = np.asfortranarray(floc_f, np.float64)
            sp = np.ascontiguousarray(species, np.int32)
            nq, nspec = q.shape[0], f.shape[1]
            qp, fp, spp = _p(q), _p(f), _p(sp)
        else:
            nq, nspec, qp, fp, spp = 0, 0, None, None, None
        oc = None if occ is None else np.ascontiguousarray(occ, np.float64)
        _check(lib().cpwgpu_plan_forces(
            self._h, int(nb), C.c_void_p(psi_t.data_ptr()), _ld(psi_t),
            None if oc is None else _p(oc), int(natom), _p(rat), _p(av), _p(pa), int(nq), qp,
            int(nspec), spp, fp, _p(force), C.byref(enl)))
        return force, enl.value

    def close(self):
        if self._h:
            lib().cpwgpu_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """H|psi> of several k-points (each its own Plan, plane-wave count and
    bands) recorded once into a HIP graph and replayed with one launch."""

    def __init__(self, plans, psis, hpsis, prec=PREC_FP64):
        import torch

        torch.cuda.synchronize()
        nk = len(plans)
        self._keep = (list(plans), list(psis), list(hpsis))
        P = (C.c_void_p * nk)(*[p._h.value for p in plans])
        ne = (C.c_int * nk)(*[t.shape[0] for t in psis])
        ps = (C.c_void_p * nk)(*[t.data_ptr() for t in psis])
        lp = (C.c_long * nk)(*[_ld(t) for t in psis])
        hs = (C.c_void_p * nk)(*[t.data_ptr() for t in hpsis])
        lh = (C.c_long * nk)(*[_ld(t) for t in hpsis])
        self._h = C.c_void_p()
        _check(lib().cpwgpu_graph_create(C.byref(self._h), nk, P, ne, ps, lp, hs, lh, int(prec)))

    def launch(self, stream=None):
        import torch

        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _check(lib().cpwgpu_graph_launch(self._h, C.c_void_p(stream)))

    def close(self):
        if self._h:
            lib().cpwgpu_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
