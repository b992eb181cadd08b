"""The force oracle (restatement of for_str_nl_kb / for_str_local_force)
against central finite differences of the energy E = sum_n <psi_n|V_loc +
V_NL|psi_n> computed by the hk_psi restatement with regenerated V(r) and
projectors at displaced atoms (h = 1e-4 bohr, psi fixed)."""

import numpy as np

from cpw2000_amd import synth


def test_force_oracle_matches_finite_differences(orc):
    from oracle import forces as Fo

    d = synth.make_config("si2f")
    psi = d["psi"]
    F, enl = Fo.forces(d, psi)
    assert np.abs(F).max() > 1e-3                       # displaced atoms: real forces
    apply = lambda dd, p: orc.orc_hk_psi(dd, psi=p)[1]
    h = 1e-4
    for a in range(2):
        for j in range(3):
            dfrac = np.linalg.solve(d["avec"], h * np.eye(3)[j])
            es = []
            for sgn in (1, -1):
                r = d["rat"].copy()
                r[a] += sgn * dfrac
                es.append(Fo.potential_energy(synth.make_config("si2f", positions=[r]), psi, apply))
            assert abs(F[a, j] + (es[0] - es[1]) / (2 * h)) < 1e-8


def test_nonlocal_energy_matches_operator(orc):
    """E_NL from the projections equals <psi|V_NL|psi> from the operator."""
    from oracle import forces as Fo

    d = dict(synth.make_config("si2f"))
    psi = d["psi"]
    _, enl = Fo.nl_forces_cov(d, psi)
    full = Fo.potential_energy(d, psi, lambda dd, p: orc.orc_hk_psi(dd, psi=p)[1])
    d0 = dict(d, nanl=0)
    loc = Fo.potential_energy(d0, psi, lambda dd, p: orc.orc_hk_psi(dd, psi=p)[1])
    assert abs((full - loc) - enl) < 1e-12 * max(1.0, abs(enl))
