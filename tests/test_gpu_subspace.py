"""FP64 subspace step on the GPU (Rayleigh-Ritz with Cholesky, Jacobi
eigensolver, rotation, residuals) against LAPACK, and a block iteration built
from it converging to the exact spectrum of the 2-atom cell.

Tolerances: Ritz values 1e-11 relative to the spectral scale, rotated
vectors orthonormal to 1e-12, residual norms 1e-10 relative; converged
eigenvalues 1e-9 Ha (FP64 iteration) and 1e-6 Ha (FP32-mixed iteration
with one FP64 refinement pass)."""

import numpy as np
import pytest
import scipy.linalg as sla

from cpw2000_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.fixture(scope="module")
def hk():
    from cpw2000_amd import hk_psi

    return hk_psi


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to("cuda")


def _apply(torch, plan, psi_t, prec):
    h = torch.zeros_like(psi_t)
    plan.apply(psi_t, h, prec)
    torch.cuda.synchronize()
    return h


@pytest.mark.parametrize("cfg,nb", [(0, 16), (0, 7), (0, 1), (1, 33), (1, 128)])
def test_rayleigh_ritz_matches_lapack(torch_dev, hk, cfg, nb):
    torch = torch_dev
    d = synth.make_config(cfg, nbands=max(nb, 2))
    plan = hk.Plan.from_inputs(d)
    mtxd = d["mtxd"]
    # a non-orthonormal basis: mix the bands
    rng = np.random.default_rng(3)
    mix = np.eye(nb) + 0.3 * (rng.normal(size=(nb, nb)) + 1j * rng.normal(size=(nb, nb)))
    x = d["psi"][:mtxd, :nb] @ mix
    psi = _dev(torch, x)
    hpsi = _apply(torch, plan, psi, hk.PREC_FP64)
    hx = hpsi.cpu().numpy().T
    s = x.conj().T @ x
    m = x.conj().T @ hx
    ref = sla.eigh(0.5 * (m + m.conj().T), s, eigvals_only=True)
    eig, res = hk.rayleigh_ritz(psi, hpsi, rotate=True)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(eig - ref).max() < 1e-11 * scale
    y = psi.cpu().numpy().T
    hy = hpsi.cpu().numpy().T
    assert np.abs(y.conj().T @ y - np.eye(nb)).max() < 1e-12
    # rotated H psi equals H applied to the rotated vectors
    h2 = _apply(torch, plan, psi, hk.PREC_FP64).cpu().numpy().T
    assert np.linalg.norm(h2 - hy) < 1e-12 * np.linalg.norm(hy)
    r = np.sum(np.abs(hy - y * eig[None, :]) ** 2, axis=0)
    assert np.allclose(res, r, rtol=1e-10, atol=1e-24)
    plan.close()


def test_singular_overlap_is_reported(torch_dev, hk):
    torch = torch_dev
    d = synth.make_config(0, nbands=4)
    x = d["psi"][: d["mtxd"], :4].copy()
    x[:, 3] = x[:, 2]
    psi = _dev(torch, x)
    with pytest.raises(hk.CpwGpuError) as e:
        hk.rayleigh_ritz(psi, psi.clone(), rotate=False)
    assert e.value.code == hk.ERR_NOTPD


def _block_davidson(torch, hk, plan, d, nb, prec, iters, tol, nbuf=8):
    """Lowest nb eigenpairs: nb + nbuf vectors (the buffer keeps degenerate
    levels of the cell from being split at the edge of the block), basis
    [psi, P r], FP64 Rayleigh-Ritz on the GPU, Jacobi preconditioner
    1/(ekpg - eps + 1) on the residuals (the relaxation of
    Src/omp/diag_rq_jac_c16.f90 with the kinetic diagonal)."""
    mtxd = d["mtxd"]
    want = nb
    nb = nb + nbuf
    rng = np.random.default_rng(11)
    x0 = (rng.normal(size=(mtxd, nb)) + 1j * rng.normal(size=(mtxd, nb))) * np.exp(
        -d["ekpg"][:mtxd] / 2)[:, None]
    psi = _dev(torch, x0)
    hpsi = _apply(torch, plan, psi, prec)
    eig, res = hk.rayleigh_ritz(psi, hpsi)
    ek = torch.from_numpy(d["ekpg"][:mtxd]).to("cuda")
    for _ in range(iters):
        e = torch.from_numpy(eig).to("cuda")
        r = hpsi - psi * e[:, None]
        t = -r / (ek[None, :] - e[:, None] + 1.0)
        ht = _apply(torch, plan, t.contiguous(), prec)
        basis = torch.cat([psi, t], 0).contiguous()
        hbasis = torch.cat([hpsi, ht], 0).contiguous()
        e2, _ = hk.rayleigh_ritz(basis, hbasis, residuals=False)
        psi = basis[:nb].contiguous()
        hpsi = hbasis[:nb].contiguous()
        eig, res = hk.rayleigh_ritz(psi, hpsi)
        if np.sqrt(res[:want].max()) < tol:
            break
    return psi, eig[:want], res[:want]


def test_block_iteration_converges_fp64(torch_dev, hk, orc):
    torch = torch_dev
    d = synth.make_config(0, nbands=2)
    plan = hk.Plan.from_inputs(d)
    exact = np.linalg.eigvalsh(orc.dense_h(d))[:8]
    psi, eig, res = _block_davidson(torch, hk, plan, d, 8, hk.PREC_FP64, 60, 1e-7)
    assert np.abs(eig - exact).max() < 1e-9
    plan.close()


def test_block_iteration_fp32_mixed_with_fp64_refinement(torch_dev, hk, orc):
    torch = torch_dev
    d = synth.make_config(0, nbands=2)
    plan = hk.Plan.from_inputs(d)
    exact = np.linalg.eigvalsh(orc.dense_h(d))[:8]
    psi, eig, res = _block_davidson(torch, hk, plan, d, 8, hk.PREC_FP32, 60, 1e-5)
    # FP64 refinement pass: H in FP64 on the converged vectors, FP64 Rayleigh-Ritz
    hpsi = _apply(torch, plan, psi, hk.PREC_FP64)
    eig64, res64 = hk.rayleigh_ritz(psi, hpsi)
    assert np.abs(eig64[:8] - exact).max() < 1e-6
    plan.close()


@pytest.mark.parametrize("nb", [64, 96, 128])
def test_eigensolver_on_dense_hermitian_subspace(torch_dev, hk, nb):
    """Orthonormal psi and H psi = psi A (A dense Hermitian): the Ritz values
    are the eigenvalues of A.  The Jacobi sweeps run past convergence here
    (off-diagonal elements decay into the denormal range); every rotation must
    stay unitary (regression: |a_pq| computed by squaring lost its digits
    there and the phase factor rescaled rows and columns)."""
    torch = torch_dev
    d = synth.make_config(1)
    x = _dev(torch, d["psi"][: d["mtxd"], :nb])
    rng = np.random.default_rng(nb)
    a = rng.normal(size=(nb, nb)) + 1j * rng.normal(size=(nb, nb))
    a = 0.5 * (a + a.conj().T)
    hx = (torch.from_numpy(a).to("cuda").T @ x).contiguous()
    eig, _ = hk.rayleigh_ritz(x.clone(), hx, rotate=False, residuals=False)
    assert np.abs(eig - np.linalg.eigvalsh(a)).max() < 1e-11 * np.abs(a).sum() / nb
