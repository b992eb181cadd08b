"""ORACLE -- test infrastructure only.

CPU restatement of the two-component Hamiltonian application
Src/hk_psi/reference/hk_psi_spin_c16.f90 (subroutine hk_psi_spin_c16), and an
FFT-free dense form of the same operator.

Pinned by (tests/test_spin_oracle.py):
* nsp = 1 and 2: per component it must equal the scalar C restatement
  (liborc, itself pinned to the reference's golden vectors) with V1 and
  V1 +/- V2, with spin-diagonal projectors;
* nsp = 4: the dense spinor H built in G space (no FFT) applied to psi;
* constant fields V1..V4 and no projectors: each plane wave i is an
  eigenspace with eigenvalues ekpg_i + V1 +/- sqrt(V2^2 + V3^2 + V4^2) (the
  Pauli matrices with sigma_y = [[0, -i], [i, 0]]), which fixes the sign of
  the sigma_y coupling;
* Hermiticity <phi|H psi> = conj <psi|H phi>.

Only tests use this module.
"""

from __future__ import annotations

import numpy as np

from oracle.oracle import pw_vectors


def _mesh(d):
    n1, n2, n3, idd = [int(x) for x in d["kmscr"][3:7]]
    return n1, n2, n3, idd


def _component(vcol, d):
    n1, n2, n3, idd = _mesh(d)
    return vcol[: idd * n2 * n3].reshape(n3, n2, idd)[:, :, :n1]


def hk_psi_spin(d, psi_sp, vscr_sp, nsp, anlsp=None, xnlkbsp=None, nanlsp=0):
    """H psi_sp (rows 2i / 2i+1 = up / down of plane wave i).

    Local part per band (hk_psi_spin_c16.f90:197-289): scatter both components
    onto the mesh, transform G -> r (unnormalised, e^{+iG.r}), multiply by
    nsp = 1: V1;  nsp = 2: V1 + V2 (up), V1 - V2 (down);
    nsp = 4: up' = (V1+V2) up + (V3 - i V4) down, down' = (V1-V2) down + (V3 + i V4) up
    (:232-262), transform back with 1/N, gather; plus ekpg psi per component
    (:282-287).  Non-local part (:302-303): hk_psi_nl_c16 on the 2*mtxd
    interleaved rows, H psi += anlsp xnlkbsp anlsp^H psi."""
    n1, n2, n3, idd = _mesh(d)
    mtxd = d["mtxd"]
    g = pw_vectors(d)
    idx = (g[:, 2] % n3, g[:, 1] % n2, g[:, 0] % n1)
    nn = n1 * n2 * n3
    v = [_component(vscr_sp[:, c], d) for c in range(nsp)]
    psi = np.asarray(psi_sp)
    out = np.zeros((2 * mtxd, psi.shape[1]), np.complex128)
    ek = d["ekpg"][:mtxd]
    for n in range(psi.shape[1]):
        f = []
        for c in range(2):
            box = np.zeros((n3, n2, n1), np.complex128)
            box[idx] = psi[c:2 * mtxd:2, n]
            f.append(np.fft.ifftn(box) * nn)
        up, dn = f
        if nsp == 1:
            tu, td = v[0] * up, v[0] * dn
        elif nsp == 2:
            tu, td = (v[0] + v[1]) * up, (v[0] - v[1]) * dn
        else:
            tu = (v[0] + v[1]) * up + (v[2] - 1j * v[3]) * dn
            td = (v[0] - v[1]) * dn + (v[2] + 1j * v[3]) * up
        for c, t in enumerate((tu, td)):
            out[c::2, n] = np.fft.fftn(t)[idx] / nn + ek * psi[c:2 * mtxd:2, n]
    if nanlsp:
        a = anlsp[: 2 * mtxd, :nanlsp]
        out += (a * xnlkbsp[None, :nanlsp]) @ (a.conj().T @ psi[: 2 * mtxd])
    return out


def dense_h_spin(d, vscr_sp, nsp, anlsp=None, xnlkbsp=None, nanlsp=0):
    """Dense (2 mtxd)^2 spinor H with the same interleaved ordering, built in
    G space: H[(i,s),(j,t)] = ekpg_i d_ij d_st + W_st(G_i - G_j) + NL, with
    W_uu = V1 + V2, W_dd = V1 - V2, W_ud = V3 - i V4, W_du = V3 + i V4 and
    V_c(q) = (1/N) sum_r V_c(r) e^{-i q.r}."""
    n1, n2, n3, idd = _mesh(d)
    mtxd = d["mtxd"]
    g = pw_vectors(d)
    dg = g[:, None, :] - g[None, :, :]
    ix = (dg[..., 2] % n3, dg[..., 1] % n2, dg[..., 0] % n1)
    nn = n1 * n2 * n3
    vq = [np.fft.fftn(_component(vscr_sp[:, c], d))[ix] / nn for c in range(nsp)]
    z = np.zeros((mtxd, mtxd), np.complex128)
    if nsp == 1:
        w = {(0, 0): vq[0], (1, 1): vq[0], (0, 1): z, (1, 0): z}
    elif nsp == 2:
        w = {(0, 0): vq[0] + vq[1], (1, 1): vq[0] - vq[1], (0, 1): z, (1, 0): z}
    else:
        w = {(0, 0): vq[0] + vq[1], (1, 1): vq[0] - vq[1],
             (0, 1): vq[2] - 1j * vq[3], (1, 0): vq[2] + 1j * vq[3]}
    h = np.zeros((2 * mtxd, 2 * mtxd), np.complex128)
    for (s, t), blk in w.items():
        h[s::2, t::2] = blk
    h[np.diag_indices_from(h)] += np.repeat(d["ekpg"][:mtxd], 2)
    if nanlsp:
        a = anlsp[: 2 * mtxd, :nanlsp]
        h += (a * xnlkbsp[None, :nanlsp]) @ a.conj().T
    return h
