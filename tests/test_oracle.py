"""The oracle pinned: the C restatement of hk_psi_c16 against the reference's
own output (golden vectors), against FFT-free forms of the operator, and
against closed forms and invariants of the 3-D transforms."""

import numpy as np
import pytest

from tests.conftest import GOLDEN, GOLDEN_CASES, load_golden, rel

from cpw2000_amd import synth


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_restatement_matches_reference(orc, name):
    d, psi, href = load_golden(name)
    st, h = orc.orc_hk_psi(d, psi=psi)
    assert st == 0
    assert rel(h[: d["mtxd"]], href) < 1e-13


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_reference_matches_convolution(orc, name):
    """The reference output equals sum_q V(q) psi(G-q) + kinetic + KB, with V(q)
    taken from the generator's list (no mesh, no FFT)."""
    d, psi, href = load_golden(name)
    assert rel(orc.conv_apply(d, psi), href) < 1e-13


def test_dense_hamiltonian_cfg0(orc):
    d, psi, href = load_golden("cfg0")
    h = orc.dense_h(d)
    assert np.abs(h - h.conj().T).max() < 1e-13
    assert rel(h @ psi[: d["mtxd"]], href) < 1e-13


@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 12, 16, 18, 20, 24, 32, 36, 40, 48, 50, 54, 64, 72,
                               80, 96, 100, 108, 128, 144, 160, 216])
def test_fft1d_is_the_dft(orc, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    np.testing.assert_allclose(orc.orc_fft1d(x, 1), np.fft.fft(x), atol=1e-12 * np.sqrt(n))
    np.testing.assert_allclose(orc.orc_fft1d(x, -1), np.fft.ifft(x) * n, atol=1e-12 * np.sqrt(n))


def _fft_fixture():
    import os

    return np.load(os.path.join(GOLDEN, "fft_wf_16x20x24.npz"), allow_pickle=False)


def test_pruned_fft_matches_reference_fft(orc):
    z = _fft_fixture()
    ns, kd = [int(v) for v in z["ns"]], [int(v) for v in z["kd"]]
    st, out = orc.orc_cfft_wf(z["chd_in_bwd"].reshape(-1), ns, kd, -1)
    assert st == 0
    assert rel(out, z["chd_out_bwd"].reshape(-1)) < 1e-14
    st, out = orc.orc_cfft_wf(z["chd_in_fwd"].reshape(-1), ns, kd, 1)
    n1, n2, n3 = ns
    box = lambda n, k: np.r_[0:k + 1, n - k:n]
    sel = np.ix_(box(n3, kd[2]), box(n2, kd[1]), box(n1, kd[0]))
    a = out.reshape(n3, n2, n1)[sel]
    b = z["chd_out_fwd"][sel]
    assert rel(a, b) < 1e-14


def test_pruned_fft_is_numpy_fft(orc):
    """G->r is the unnormalised inverse DFT; r->G the 1/N forward DFT on the box."""
    z = _fft_fixture()
    ns, kd = [int(v) for v in z["ns"]], [int(v) for v in z["kd"]]
    n1, n2, n3 = ns
    x = z["chd_in_bwd"]
    st, out = orc.orc_cfft_wf(x.reshape(-1), ns, kd, -1)
    assert rel(out.reshape(n3, n2, n1), np.fft.ifftn(x) * x.size) < 1e-14
    y = z["chd_in_fwd"]
    st, out = orc.orc_cfft_wf(y.reshape(-1), ns, kd, 1)
    box = lambda n, k: np.r_[0:k + 1, n - k:n]
    sel = np.ix_(box(n3, kd[2]), box(n2, kd[1]), box(n1, kd[0]))
    assert rel(out.reshape(n3, n2, n1)[sel], (np.fft.fftn(y) / y.size)[sel]) < 1e-14


def test_delta_gives_plane_wave(orc):
    ns, kd = (16, 20, 24), (3, 4, 5)
    n1, n2, n3 = ns
    g0 = (2, -3, 4)
    x = np.zeros((n3, n2, n1), complex)
    x[g0[2] % n3, g0[1] % n2, g0[0] % n1] = 1.0
    st, out = orc.orc_cfft_wf(x.reshape(-1), ns, kd, -1)
    k3, k2, k1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    pw = np.exp(2j * np.pi * (g0[0] * k1 / n1 + g0[1] * k2 / n2 + g0[2] * k3 / n3))
    np.testing.assert_allclose(out.reshape(n3, n2, n1), pw, atol=1e-13)


def test_roundtrip_and_parseval(orc):
    ns, kd = (24, 24, 24), (5, 5, 5)
    rng = np.random.default_rng(7)
    x = np.zeros((24, 24, 24), complex)
    box = np.r_[0:6, 19:24]
    sel = np.ix_(box, box, box)
    x[sel] = rng.normal(size=(11, 11, 11)) + 1j * rng.normal(size=(11, 11, 11))
    st, r = orc.orc_cfft_wf(x.reshape(-1), ns, kd, -1)
    # Parseval: sum |f(r)|^2 = N sum |c_G|^2
    np.testing.assert_allclose(np.sum(np.abs(r) ** 2), x.size * np.sum(np.abs(x) ** 2), rtol=1e-13)
    st, g = orc.orc_cfft_wf(r, ns, kd, 1)
    assert rel(g.reshape(24, 24, 24)[sel], x[sel]) < 1e-14


def test_constant_potential_eigenpairs(orc):
    """V(r) = V0, no projectors: H e_i = (ekpg_i + V0) e_i."""
    d = dict(synth.make_config(0, nbands=4))
    v0 = -0.731
    d["vscr"] = np.full_like(d["vscr"], v0)
    d["nanl"] = 0
    psi = np.asfortranarray(np.eye(d["mtxd"], 6, k=0, dtype=complex)[:, :6])
    st, h = orc.orc_hk_psi(d, psi=psi)
    expect = psi * (d["ekpg"][:, None] + v0)
    np.testing.assert_allclose(h, expect, atol=1e-13)


def test_linearity_and_hermiticity(orc):
    d = synth.make_config(0)
    p = d["psi"]
    a, b = 0.3 - 1.1j, -0.7 + 0.2j
    _, h = orc.orc_hk_psi(d, psi=p)
    comb = np.asfortranarray(a * p[:, :4] + b * p[:, 4:8])
    _, hc = orc.orc_hk_psi(d, psi=comb)
    assert rel(hc, a * h[:, :4] + b * h[:, 4:8]) < 1e-14
    m = p.conj().T @ h
    assert np.abs(m - m.conj().T).max() < 1e-13


def test_kb_term_rank(orc):
    """H - H(no projectors) has rank <= nanl."""
    d = dict(synth.make_config(0))
    h1 = orc.dense_h(d)
    d["anlga"] = np.zeros_like(d["anlga"])
    h0 = orc.dense_h(d)
    s = np.linalg.svd(h1 - h0, compute_uv=False)
    assert np.sum(s > 1e-10 * s[0]) <= d["nanl"]


def test_error_codes(orc):
    d = dict(synth.make_config(0, nbands=2))
    # ng > mxdgve  (hk_psi_c16.f90:108)
    e = dict(d, mxdgve=d["ng"] - 1)
    assert orc.orc_hk_psi(e)[0] == 1
    # FFT work size > mxdscr  (:112-119)
    e = dict(d, mxdscr=100)
    assert orc.orc_hk_psi(e)[0] == 2
    # sphere wraps the mesh  (:162-172): kmscr asks for a 10-point mesh, kd = 5
    km = d["kmscr"].copy()
    km[3:7] = (10, 10, 10, 10)
    e = dict(d, kmscr=km)
    assert orc.orc_hk_psi(e)[0] == 3


def test_nan_in_one_band_stays_in_that_band(orc):
    """The restatement, like the reference (no check on psi), returns status 0
    for a NaN in band 5; that band's H psi is non-finite, the others equal
    the clean call bit for bit."""
    import numpy as np

    from cpw2000_amd import synth

    d = synth.make_config(0)
    st, clean = orc.orc_hk_psi(d)
    assert st == 0
    psi = d["psi"].copy(order="F")
    psi[17, 5] = np.nan
    st, out = orc.orc_hk_psi(d, psi=psi)
    assert st == 0
    m = d["mtxd"]
    assert not np.isfinite(out[:m, 5]).any()
    others = [b for b in range(psi.shape[1]) if b != 5]
    assert np.array_equal(out[:m, others], clean[:m, others])
