"""ORACLE -- test infrastructure only.

Kleinman-Bylander projectors of one k-point, restating in numpy the
reference's projector construction proj_nl_kb_c16
(Src/omp/proj_nl_kb_c16.f90:23-275) with the pieces it calls:

* cell: avec from the metric adot in the reference's canonical orientation
  (a1 along x, a2 in the xy plane; Src/crystal/adot_to_avec.f90), bvec =
  2 pi inv(avec)^T, vcell = sqrt(det adot) (Src/crystal/adot_to_bdot.f90);
* angular part: rylm_lm(q) = |q|^l sqrt(4 pi / (2l+1)) Y_lm, the real
  harmonics of Src/numeric/ylm_rc.f90:82-107 for l <= 3;
* radial part: the quadratic interpolation of the table vkb(-2:mxdlqp, l, 0,
  type) at |q| / delqnl (proj_nl_kb_c16.f90:227-240), zero from nqnl - 1 on,
  times 1/sqrt(vcell) (:128) and i^l (:243);
* phase exp(-2 pi i g . rat) of the G-vector alone, not k + G (:214-223);
* column order type -> l (nkb(l, 0, type) /= 0) -> m = -l..l -> atom (:169-186);
  xnlkb = nkb as a real number (:265-267).

Argument names, meanings and 1-based index conventions follow the
reference's dummy list; arrays are numpy arrays in Fortran order, vkb with
its first index shifted by 2 (vkb[j + 2, l, s + 1, type] holds vkb(j, l, s, type)).
"""

from __future__ import annotations

import math

import numpy as np

LMAX = 3


class ProjError(ValueError):
    """A condition on which the reference writes a message and stops."""


def adot_to_avec(adot):
    """avec (columns a_j, upper triangular) and bvec (columns b_j, lower
    triangular) of the metric adot (Src/crystal/adot_to_avec.f90)."""
    a = np.zeros((3, 3))
    eps = 1e-10
    if adot[0, 0] < eps:
        raise ProjError("adot_to_avec: avec(1,1)")
    a[0, 0] = math.sqrt(adot[0, 0])
    a[0, 1] = adot[0, 1] / a[0, 0]
    t = adot[1, 1] - a[0, 1] * a[0, 1]
    if t < eps:
        raise ProjError("adot_to_avec: avec(2,2)")
    a[1, 1] = math.sqrt(t)
    a[0, 2] = adot[0, 2] / a[0, 0]
    a[1, 2] = (adot[1, 2] - a[0, 1] * a[0, 2]) / a[1, 1]
    t = adot[2, 2] - a[1, 2] * a[1, 2] - a[0, 2] * a[0, 2]
    if t < eps:
        raise ProjError("adot_to_avec: avec(3,3)")
    a[2, 2] = math.sqrt(t)
    b = 2.0 * math.pi * np.linalg.inv(a).T
    return a, b


def vcell_of(adot):
    """Cell volume sqrt(det adot) (Src/crystal/adot_to_bdot.f90: cofactor
    expansion, then the square root)."""
    d = float(np.linalg.det(adot))
    if d < 1e-10:
        raise ProjError("adot_to_bdot: vcell")
    return math.sqrt(d)


def rylm(q):
    """rylm[l][m + l] for l = 0..3 at the Cartesian points q (..., 3)
    (Src/numeric/ylm_rc.f90:82-107)."""
    x, y, z = q[..., 0], q[..., 1], q[..., 2]
    s3 = math.sqrt(3.0)
    r = [[np.ones_like(x)], [y, z, x],
         [s3 * x * y, s3 * z * y, 0.5 * (2 * z * z - x * x - y * y), s3 * z * x,
          math.sqrt(0.75) * (x * x - y * y)],
         [math.sqrt(5.0 / 8.0) * (3 * x * x - y * y) * y, math.sqrt(15.0) * z * x * y,
          math.sqrt(3.0 / 8.0) * (4 * z * z - x * x - y * y) * y, 0.5 * (2 * z * z - 3 * x * x - 3 * y * y) * z,
          math.sqrt(3.0 / 8.0) * (4 * z * z - x * x - y * y) * x, math.sqrt(15.0 / 4.0) * (x * x - y * y) * z,
          -math.sqrt(5.0 / 8.0) * (3 * y * y - x * x) * x]]
    return r


def interp_vkb(line, nq, delq, qi):
    """Quadratic interpolation of one table line (line[j + 2] = vkb(j)) at
    |q| = qi, zero where nint(qi / delq) >= nq - 1
    (Src/omp/proj_nl_kb_c16.f90:227-240)."""
    xni = qi / delq
    ni = np.floor(xni + 0.5).astype(np.int64)          # nint of a non-negative number
    ni = np.maximum(ni, 0)
    ok = ni < nq - 1
    nis = np.where(ok, ni, 0)
    xi = xni - nis
    v0, vp, vm = line[nis + 2], line[nis + 3], line[nis + 1]
    v = v0 * (1.0 + xi) * (1.0 - xi) + 0.5 * (vp * (1.0 + xi) - vm * (1.0 - xi)) * xi
    return np.where(ok, v, 0.0)


def projector_columns(ntype, natom, nkb):
    """(type, atom, l, m) of each projector column, 0-based type and atom
    (Src/omp/proj_nl_kb_c16.f90:169-186), and the count check of :141-151."""
    cols = []
    for k in range(ntype):
        for l in range(LMAX + 1):
            if nkb[l, 1, k] != 0:
                for m in range(-l, l + 1):
                    for kk in range(natom[k]):
                        cols.append((k, kk, l, m))
    return cols


def proj_nl_kb_c16(rkpt, mtxd, isort, nanl, ng, kgv, nqnl, delqnl, vkb, nkb, ntype, natom, rat, adot,
                   mxddim, mxdanl, mxdgve):
    """anlga (mxddim, mxdanl) complex, xnlkb (mxdanl,) of
    Src/omp/proj_nl_kb_c16.f90:23-275.  Rows beyond mtxd and columns beyond
    nanl are zero.  The reference's stops raise ProjError."""
    if ng > mxdgve:
        raise ProjError("ng > mxdgve")
    cols = projector_columns(ntype, natom, nkb)
    if len(cols) != nanl:
        raise ProjError(f"proj_nl_kb: inconsistent value of nanl {nanl} should be {len(cols)}")
    if nanl > mxdanl:
        raise ProjError(f"proj_nl_kb: increase mxdanl from {mxdanl} to {nanl}")
    _, bvec = adot_to_avec(adot)
    fac = 1.0 / math.sqrt(vcell_of(adot))
    g = kgv[:, isort[:mtxd] - 1].T.astype(np.float64)          # (mtxd, 3)
    qk = g + np.asarray(rkpt, float)[None, :]
    qcar = qk @ bvec.T
    qi = np.sqrt(np.einsum("ij,ij->i", qcar, qcar))
    ry = rylm(qcar)
    anlga = np.zeros((mxddim, mxdanl), complex, order="F")
    xnlkb = np.zeros(mxdanl)
    vq = {}
    for ind, (k, kk, l, m) in enumerate(cols):
        if (k, l) not in vq:
            vq[(k, l)] = fac * interp_vkb(vkb[:, l, 1, k], nqnl[k], delqnl[k], qi) * (1j ** l)
        fi = -2.0 * math.pi * (g @ rat[:, kk, k])
        st = np.cos(fi) + 1j * np.sin(fi)
        anlga[:mtxd, ind] = st * ry[l][m + l] * vq[(k, l)]
        xnlkb[ind] = float(nkb[l, 1, k])
    return anlga, xnlkb
