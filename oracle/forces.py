"""ORACLE -- test infrastructure only.

Atomic forces of E = sum_n occ_n <psi_n| V_loc + V_NL |psi_n> at fixed psi,
restating the reference force routines in numpy:

* non-local: Src/ALFA/for_str_nl_kb.f90:286-362 and :392-398 -- projections
  <beta|psi> and <(-i g) beta|psi> (the dhd rows 1 and 2-4 of its zgemm at
  :346), f_j = 2 pi * sum_n 2 occ xnlkb Re[dhd_j conj(dhd_1)], covariant
  (per unit of the fractional coordinate rat_j).
* local: Src/ALFA/for_str_local_force.f90:102-145 -- f_j = 2 pi * sum_q g_j
  Re[i exp(-2 pi i g.rat) v_s(q) conj(rho(q))], with rho(q) the band density
  in reciprocal space (Src/omp/charge_by_fft.f90:191-214).
* Cartesian forces F = inv(avec)^T f (Src/cpw_lib/force_stress_kb.f90:165-169
  uses bdot/4pi^2 for the contravariant components; the Cartesian form is
  the same vector).

Conventions of this build's inputs: psi_n(r) = sum_G psi_n(G) exp(iG.r) on the
mesh, rho_hat(q) = (1/N) sum_r rho(r) exp(-iq.r), so that
E_loc = sum_q V(q) conj(rho_hat(q)), V(q) = sum_s f_s(q) sum_tau exp(-2 pi i q.rat).
"""

from __future__ import annotations

import numpy as np

TWOPI = 2.0 * np.pi


def band_density_q(d, psi, occ=None):
    """rho_hat on the whole mesh (n3, n2, n1), indexed by q mod n."""
    n1, n2, n3 = [int(x) for x in d["kmscr"][3:6]]
    mtxd = d["mtxd"]
    g = d["kgv"].T[d["isort"][:mtxd] - 1].astype(np.int64)
    idx = (g[:, 2] % n3, g[:, 1] % n2, g[:, 0] % n1)
    nb = psi.shape[1]
    occ = np.ones(nb) if occ is None else np.asarray(occ, float)
    rho = np.zeros((n3, n2, n1))
    N = n1 * n2 * n3
    for n in range(nb):
        grid = np.zeros((n3, n2, n1), complex)
        grid[idx] = psi[:mtxd, n]
        pr = np.fft.ifftn(grid) * N
        rho += occ[n] * (pr.real ** 2 + pr.imag ** 2)
    return np.fft.fftn(rho) / N


def local_forces_cov(d, psi, occ=None):
    """Covariant local forces (natom, 3) (for_str_local_force.f90:102-145)."""
    n1, n2, n3 = [int(x) for x in d["kmscr"][3:6]]
    rq = band_density_q(d, psi, occ)
    q = d["floc_q"].astype(np.int64)
    rho = rq[q[:, 2] % n3, q[:, 1] % n2, q[:, 0] % n1]
    rat, sp = d["rat"], d["species"]
    f = np.zeros((rat.shape[0], 3))
    for a in range(rat.shape[0]):
        ph = np.exp(-1j * TWOPI * (q @ rat[a]))
        e1 = np.real(1j * ph * d["floc_f"][:, sp[a]] * np.conj(rho))
        f[a] = TWOPI * (q * e1[:, None]).sum(axis=0)
    return f


def projections(d, psi):
    """P0 = <beta|psi>, Pj = sum_G g_j conj(beta) psi  (nanl, nb) and (3, nanl, nb)."""
    mtxd, nanl = d["mtxd"], d["nanl"]
    a = d["anlga"][:mtxd, :nanl]
    g = d["kgv"].T[d["isort"][:mtxd] - 1].astype(np.float64)
    p = psi[:mtxd]
    p0 = a.conj().T @ p
    pj = np.stack([(a * g[:, j:j + 1]).conj().T @ p for j in range(3)])
    return p0, pj


def nl_forces_cov(d, psi, occ=None):
    """Covariant non-local forces (natom, 3) and E_NL (for_str_nl_kb.f90:346-398).

    dhd_1 = P0 and dhd_{1+j} = -i Pj up to a common per-projector phase."""
    nb = psi.shape[1]
    occ = np.ones(nb) if occ is None else np.asarray(occ, float)
    p0, pj = projections(d, psi)
    x = d["xnlkb"][: d["nanl"]]
    natom = d["rat"].shape[0]
    f = np.zeros((natom, 3))
    for j in range(3):
        t = 2.0 * occ[None, :] * x[:, None] * np.real((-1j * pj[j]) * np.conj(p0))
        np.add.at(f[:, j], d["proj_atom"][: d["nanl"]], TWOPI * t.sum(axis=1))
    enl = float(np.sum(occ[None, :] * x[:, None] * np.abs(p0) ** 2))
    return f, enl


def cartesian(f_cov, avec):
    """Cartesian forces from covariant components: F = inv(avec)^T f."""
    return f_cov @ np.linalg.inv(avec)


def forces(d, psi, occ=None):
    """Cartesian forces (natom, 3) in Ha/bohr and E_NL."""
    fnl, enl = nl_forces_cov(d, psi, occ)
    fl = local_forces_cov(d, psi, occ)
    return cartesian(fnl + fl, d["avec"]), enl


def potential_energy(d, psi, apply_h):
    """sum_n <psi_n|H - T|psi_n> with apply_h(d, psi) -> H psi."""
    mtxd = d["mtxd"]
    h = apply_h(d, psi)[:mtxd]
    p = psi[:mtxd]
    return float(np.real(np.sum(p.conj() * h)) - np.sum(d["ekpg"][:mtxd, None] * np.abs(p) ** 2))
