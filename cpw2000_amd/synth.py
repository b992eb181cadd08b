"""Deterministic synthetic inputs for the ``hk_psi_c16`` hot path.

Every array is produced in the exact layout the Fortran subroutine
``hk_psi_c16`` (Src/hk_psi/reference/hk_psi_c16.f90:28-78) receives:

* ``kgv``   int32  (3, ng)   Fortran order, G in reciprocal-lattice coordinates,
                             ordered by increasing |G| (stars), as produced by the
                             reference g-space set-up.
* ``isort`` int32  (mxddim,) 1-based index into ``kgv`` for plane wave i
                             (Src/ALFA/hamilt_struct.f90:87-153: identity order at
                             Gamma, sorted by |k+G|^2 otherwise).
* ``ekpg``  float64(mxddim,) 1/2 |k+G|^2 in Hartree (hamilt_struct.f90:88-95).
* ``vscr``  float64(mxdscr,) V(r) on the mesh, linear index (k3*n2+k2)*id+k1
                             (Src/cpw_lib/pot_local.f90:182-184).
* ``kmscr`` int32  (7,)      (1:3) max |G_j| of the potential sphere, (4:6) n1,n2,n3,
                             (7) leading dimension id (pot_local.f90:122-128).
* ``anlga`` complex128 (mxddim, nanl) Fortran order, Kleinman-Bylander projectors
                             built as in Src/omp/proj_nl_kb_c16.f90:190-259 with a
                             Gaussian radial form instead of an interpolated table.
* ``xnlkb`` float64 (nanl,)  projector strengths D_l.
* ``psi``   complex128 (mxddim, nbands) Fortran order, Loewdin-orthonormal.

Random draws use numpy's PCG64 with ``seed = 20260000 + cfg``; draw order is
fixed (V(G) picks, V(G) amplitudes, psi) so the inputs are reproducible.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TWOPI = 2.0 * math.pi

# Src/ALFA/size_fft.f90:46-49
_INPRIM = (2, 4, 6, 8, 10, 12, 16, 18, 20, 24, 32)
_INSEC = (36, 40, 48, 50, 54, 64)


def size_fft(kmscr3):
    """FFT mesh sizes from the max G indices (Src/ALFA/size_fft.f90:53-76).

    Returns (nsfft[3], mfft, mwrk)."""
    ns = []
    for km in kmscr3:
        m = 2 * int(km) + 1
        if m < 32:
            n = _INPRIM[-1]
            for c in _INPRIM:
                n = c
                if m <= c:
                    break
        else:
            # int(log(real(m))/log(real(2))) in single precision, as the reference
            k = int(np.float32(np.log(np.float32(m))) / np.float32(np.log(np.float32(2.0))))
            n = None
            for c in _INSEC:
                n = (2 ** (k - 5)) * c
                if m <= n:
                    break
        ns.append(n)
    mfft = (ns[0] + 1) * ns[1] * ns[2]
    mwrk = 4 * max(ns[0] * ns[1], ns[0] * ns[2], ns[1] * ns[2])
    return ns, mfft, mwrk


# --------------------------------------------------------------------------
# configurations (sizes and physical parameters of the five workloads)
# --------------------------------------------------------------------------

A_SI = 10.26


def _diamond_conventional():
    """Fractional positions of the 8-atom conventional diamond cell."""
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    return np.concatenate([fcc, fcc + 0.25])


@dataclass
class CellSpec:
    avec: np.ndarray                    # (3,3) columns are lattice vectors (bohr)
    types: list                         # list of (name, frac positions (nat,3), rc)
    ecut: float
    nbands: int
    nvg: int                            # number of random V(G) components (+-G counted)
    gv: float                           # max |G| (bohr^-1) of the random V(G)
    kpts: np.ndarray = field(default_factory=lambda: np.zeros((1, 3)))
    vloc_gauss: dict = None             # species -> (A, r_s): tau-dependent local part
    name: str = ""


def config_spec(cfg: int) -> CellSpec:
    if cfg == 0:
        a = A_SI
        avec = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], float).T
        # Cartesian (0,0,0) and (1/4,1/4,1/4)a -> fractional of the primitive cell
        cart = np.array([[0, 0, 0], [.25 * a, .25 * a, .25 * a]])
        frac = np.linalg.solve(avec, cart.T).T
        return CellSpec(avec, [("Si", frac, 1.8)], 10.0, 16, 40, 2.5,
                        name="Si2 fcc primitive, Ecut 10 Ha, 24^3, 16 bands")
    if cfg in (1, 2, 3):
        rep = 3 if cfg == 2 else 2
        L = A_SI * rep
        conv = _diamond_conventional()
        pos = np.concatenate([(conv + np.array([i, j, k])) / rep
                              for i in range(rep) for j in range(rep) for k in range(rep)])
        spec = CellSpec(L * np.eye(3), [("Si", pos, 1.8)], 15.0,
                        512 if cfg == 2 else 128, 200, 3.0)
        if cfg == 3:
            # shifted 2x2x4 Monkhorst-Pack grid (fractions of the reciprocal vectors)
            ks = []
            for i in range(2):
                for j in range(2):
                    for k in range(4):
                        ks.append([(2 * i - 1) / 4.0, (2 * j - 1) / 4.0, (2 * k - 3) / 8.0])
            spec.kpts = np.array(ks)
            spec.name = "Si64 L=20.52, 16 k (shifted 2x2x4 MP), Ecut 15 Ha, 128 bands/k"
        elif cfg == 2:
            spec.name = "Si216 L=30.78, Ecut 15 Ha, 108^3, 512 bands"
        else:
            spec.name = "Si64 L=20.52, Ecut 15 Ha, 72^3, 128 bands"
        return spec
    if cfg == 4:
        L = 21.0
        conv = _diamond_conventional()
        pos = np.concatenate([(conv + np.array([i, j, k])) / 2.0
                              for i in range(2) for j in range(2) for k in range(2)])
        # zincblende-like: the fcc sublattice is Si, the shifted one Ge
        sub = np.tile(np.array([0, 0, 0, 0, 1, 1, 1, 1]), 8)
        rng = np.random.Generator(np.random.PCG64(20260000 + 4 + 1000))
        disp = rng.uniform(-0.05, 0.05, size=pos.shape) / L
        pos = pos + disp
        si = pos[sub == 0]
        ge = pos[sub == 1]
        spec = CellSpec(L * np.eye(3), [("Si", si, 1.8), ("Ge", ge, 2.0)], 15.0, 128, 200, 3.0,
                        name="Si32Ge32 L=21.0 displaced, Ecut 15 Ha, 128 bands")
        # Gaussian local form factors v_s(q) = A_s exp(-q^2 r_s^2/2) (Ha bohr^3)
        spec.vloc_gauss = {"Si": (-8.0, 1.2), "Ge": (-9.0, 1.3)}
        return spec
    if cfg == "si2f":
        # the 2-atom cell with a tau-dependent local potential and moved atoms:
        # nonzero forces, small enough for dense checks
        spec = config_spec(0)
        frac = spec.types[0][1] + np.array([[0.011, -0.007, 0.004], [-0.006, 0.009, 0.013]])
        spec.types = [("Si", frac, 1.8)]
        spec.vloc_gauss = {"Si": (-8.0, 1.2)}
        spec.name = "Si2 fcc primitive, displaced, local Gaussian form factor"
        return spec
    if cfg == "ortho":
        # small orthorhombic cell: n1, n2, n3 all different (layout checks)
        avec = np.diag([9.0, 11.5, 7.2])
        frac = np.array([[0.0, 0.0, 0.0], [0.3, 0.45, 0.55]])
        return CellSpec(avec, [("Si", frac, 1.8)], 6.0, 8, 30, 2.0,
                        name="Si2 orthorhombic 9.0x11.5x7.2, Ecut 6 Ha, 8 bands")
    raise ValueError(f"unknown config {cfg}")


# --------------------------------------------------------------------------
# g-space
# --------------------------------------------------------------------------

def _gvectors(bvec, gcut):
    """All integer G with |G| <= gcut, ordered by |G|^2 then lexicographically."""
    avec = TWOPI * np.linalg.inv(bvec).T
    lim = [int(math.floor(gcut * np.linalg.norm(avec[:, j]) / TWOPI)) + 1 for j in range(3)]
    r = [np.arange(-m, m + 1) for m in lim]
    g = np.stack(np.meshgrid(*r, indexing="ij"), -1).reshape(-1, 3)
    gc = g @ bvec.T
    g2 = np.einsum("ij,ij->i", gc, gc)
    keep = g2 <= gcut * gcut * (1 + 1e-12)
    g = g[keep]
    g2 = g2[keep]
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], np.round(g2, 9)))
    return g[order].astype(np.int32), g2[order]


def build_kspace(spec: CellSpec, kpt=(0.0, 0.0, 0.0), kmax_override=None, idshift=0):
    """G list, isort, ekpg and mesh for one k-point (hamilt_struct.f90:87-153).

    ``kmax_override`` replaces kmscr(1:3) (a smaller potential mesh, as the
    reference's dual-grid option produces); ``idshift`` pads the leading
    dimension id = n1 + idshift."""
    avec = spec.avec
    bvec = TWOPI * np.linalg.inv(avec).T
    bdot = bvec.T @ bvec
    gmax = math.sqrt(2.0 * spec.ecut)
    kgv, _ = _gvectors(bvec, 2.0 * gmax + 1e-9)           # density / potential sphere
    ng = kgv.shape[0]
    kpt = np.asarray(kpt, float)
    qk = kgv + kpt
    ekin = 0.5 * np.einsum("ij,jk,ik->i", qk, bdot, qk)
    if np.dot(kpt, bdot @ kpt) < 1e-20:
        irow = np.arange(ng)
    else:
        irow = np.argsort(ekin, kind="stable")
    mtxd = 0
    for i in range(ng):
        if ekin[irow[i]] > spec.ecut:
            break
        mtxd = i + 1
    isort = (irow[:mtxd] + 1).astype(np.int32)
    ekpg = ekin[irow[:mtxd]].copy()
    kmax = np.abs(kgv).max(axis=0) if kmax_override is None else np.asarray(kmax_override)
    ns, mfft, mwrk = size_fft(kmax)
    kmscr = np.array([kmax[0], kmax[1], kmax[2], ns[0], ns[1], ns[2], ns[0] + idshift], np.int32)
    return dict(kgv=np.asfortranarray(kgv.T), ng=ng, isort=isort, ekpg=ekpg, mtxd=mtxd,
                kmscr=kmscr, bvec=bvec, bdot=bdot, kpt=kpt, mwrk=mwrk)


# --------------------------------------------------------------------------
# potential, wave functions, projectors
# --------------------------------------------------------------------------

def _random_potential(rng, spec, kgv, bvec, ns):
    """V(G) on nvg random components (nvg/2 +-G pairs) with |G|<=gv, V(-G)=V(G)*, V(0)=0."""
    g = kgv.T
    gc = g @ bvec.T
    gl = np.sqrt(np.einsum("ij,ij->i", gc, gc))
    # one representative per +-G pair: first nonzero component positive
    first = np.where(g[:, 0] != 0, g[:, 0], np.where(g[:, 1] != 0, g[:, 1], g[:, 2]))
    cand = np.nonzero((gl <= spec.gv) & (gl > 0) & (first > 0))[0]
    npair = spec.nvg // 2
    pick = np.sort(rng.choice(cand, size=npair, replace=False))
    amp = rng.normal(0.0, 0.3, npair) + 1j * rng.normal(0.0, 0.3, npair)
    gq = np.concatenate([g[pick], -g[pick]]).astype(np.int64)
    vq = np.concatenate([amp, np.conj(amp)])
    return gq, vq


def vloc_from_species(spec, kgv, bvec, ns, omega):
    """Tau-dependent local potential V(G) = sum_s v_s(|G|) S_s(G) / Omega with
    S_s(G) = sum_tau exp(-2 pi i G.tau) (Src/cpw_lib/structure_factor.f90:93)."""
    g = kgv.T.astype(np.int64)
    gc = g @ bvec.T
    g2 = np.einsum("ij,ij->i", gc, gc)
    gl = np.sqrt(g2)
    sel = (gl > 0) & (gl <= spec.gv)
    g = g[sel]
    g2 = g2[sel]
    v = np.zeros(g.shape[0], complex)
    for name, frac, _rc in spec.types:
        A, rs = spec.vloc_gauss[name]
        sf = np.exp(-1j * TWOPI * (g @ frac.T)).sum(axis=1)
        v += A * np.exp(-0.5 * g2 * rs * rs) * sf / omega
    return g, v


def local_form_factors(spec, kgv, bvec, omega):
    """q list (|q| <= gv, q != 0) and f_s(q) = A_s exp(-q^2 r_s^2/2)/Omega per species,
    so that the tau-dependent local potential is
    V(q) = sum_s f_s(q) sum_{atoms of s} exp(-2 pi i q.rat)."""
    g = kgv.T.astype(np.int64)
    gc = g @ bvec.T
    g2 = np.einsum("ij,ij->i", gc, gc)
    gl = np.sqrt(g2)
    sel = (gl > 0) & (gl <= spec.gv)
    g = g[sel]
    g2 = g2[sel]
    f = np.zeros((g.shape[0], len(spec.types)))
    for k, (name, _frac, _rc) in enumerate(spec.types):
        A, rs = spec.vloc_gauss[name]
        f[:, k] = A * np.exp(-0.5 * g2 * rs * rs) / omega
    return g.astype(np.int32), f


def projector_atoms(spec):
    """Atom (global index, species-major) of each projector column, in the
    ordering of ``projectors`` (type -> l -> m -> atom)."""
    out, base = [], 0
    for name, frac, rc in spec.types:
        for l in (0, 1):
            for _m in range(2 * l + 1):
                out.extend(base + np.arange(frac.shape[0]))
        base += frac.shape[0]
    return np.array(out, np.int32)


def potential_on_mesh(gq, vq, ns, idd=None):
    """V(r_j) = sum_q V(q) exp(+2 pi i q.j/n), stored as (k3*n2+k2)*id+k1."""
    n1, n2, n3 = ns
    idd = n1 if idd is None else idd
    grid = np.zeros((n3, n2, n1), complex)
    np.add.at(grid, (gq[:, 2] % n3, gq[:, 1] % n2, gq[:, 0] % n1), vq)
    vr = np.fft.ifftn(grid) * (n1 * n2 * n3)
    vr = vr.real.copy()
    if idd == n1:
        return vr.reshape(-1)
    out = np.zeros((n3, n2, idd))
    out[:, :, :n1] = vr
    return out.reshape(-1)


def projectors(spec, kgv, isort, kpt, bvec, omega, mxddim=None):
    """anlga(i,(type,l,m,atom)) = Omega^-1/2 e^{-2 pi i g.tau} rylm_lm(q) i^l exp(-q^2 rc^2/2)
    (Src/omp/proj_nl_kb_c16.f90:169-186 ordering, :194-259 values, rylm convention
    Src/numeric/ylm_rc.f90:82-88: l=1, m=-1,0,1 -> y,z,x)."""
    mtxd = isort.shape[0]
    mxddim = mtxd if mxddim is None else mxddim
    g = kgv.T[isort - 1].astype(np.float64)
    q = (g + kpt) @ bvec.T
    q2 = np.einsum("ij,ij->i", q, q)
    fac = 1.0 / math.sqrt(omega)
    cols, xnl = [], []
    dl = {0: 1.5, 1: 0.8}
    for name, frac, rc in spec.types:
        st = np.exp(-1j * TWOPI * (g @ frac.T))           # (mtxd, natom)
        rad = fac * np.exp(-0.5 * q2 * rc * rc)
        for l in (0, 1):
            ylm = [np.ones(mtxd)] if l == 0 else [q[:, 1], q[:, 2], q[:, 0]]
            for y in ylm:
                base = (1j ** l) * y * rad
                for a in range(frac.shape[0]):
                    cols.append(st[:, a] * base)
                    xnl.append(dl[l])
    anl = np.zeros((mxddim, len(cols)), complex, order="F")
    if cols:
        anl[:mtxd, :] = np.stack(cols, axis=1)
    return anl, np.array(xnl, float)


def kb_pseudo(spec, delq=0.002, lsets=None, tables=None, nqnl=None, nkb_sign=None, mxdatm_pad=0):
    """The pseudopotential inputs of the projector construction
    Src/omp/proj_nl_kb_c16.f90:23-73 in the reference's layouts (numpy,
    Fortran order): adot(3,3), natom(ntype), rat(3, mxdatm, ntype),
    nkb(0:3, -1:1, ntype) as (4, 3, ntype), vkb(-2:mxdlqp, 0:3, -1:1, ntype)
    as (mxdlqp + 3, 4, 3, ntype), nqnl and delqnl per type, nanl.

    Default: s and p projectors with the Gaussian radial part of
    ``projectors`` (vkb_l(q) = exp(-q^2 rc^2 / 2); rylm carries q^l),
    tabulated every ``delq`` up to |k+G| = sqrt(2 Ecut) + 3 bohr^-1.
    ``lsets[t]`` (the l values of type t), ``tables(t, l, q)`` (the table at
    |q|, an array), ``nqnl`` and ``nkb_sign[t][l]`` replace the defaults."""
    ntype = len(spec.types)
    natom = np.array([frac.shape[0] for _, frac, _ in spec.types], np.int32)
    mxdatm = int(natom.max()) + mxdatm_pad
    rat = np.zeros((3, mxdatm, ntype), order="F")
    for t, (_, frac, _) in enumerate(spec.types):
        rat[:, :frac.shape[0], t] = frac.T
    delqnl = np.broadcast_to(np.asarray(delq, float), (ntype,)).copy()
    if nqnl is None:
        qmax = math.sqrt(2.0 * spec.ecut) + 3.0
        nqnl = [int(qmax / dq) + 4 for dq in delqnl]
    nqnl = np.asarray(nqnl, np.int32)
    mxdlqp = max(int(nqnl.max()) - 1, 0)    # the last table point nqnl - 1 is index mxdlqp
    if lsets is None:
        lsets = [(0, 1)] * ntype
    if tables is None:
        rcs = [rc for _, _, rc in spec.types]
        tables = lambda t, l, q: np.exp(-0.5 * q * q * rcs[t] ** 2)
    nkb = np.zeros((4, 3, ntype), np.int32, order="F")
    vkb = np.zeros((mxdlqp + 3, 4, 3, ntype), order="F")
    j = np.arange(-2, mxdlqp + 1)
    for t in range(ntype):
        for l in lsets[t]:
            nkb[l, 1, t] = 1 if nkb_sign is None else nkb_sign[t][l]
            vkb[:, l, 1, t] = tables(t, l, np.abs(j) * delqnl[t])
    nanl = sum((2 * l + 1) * int(natom[t]) for t in range(ntype) for l in lsets[t])
    return dict(nqnl=nqnl, delqnl=delqnl, vkb=vkb, nkb=nkb, ntype=ntype, natom=natom, rat=rat,
                adot=np.asfortranarray(spec.avec.T @ spec.avec), nanl=nanl)


def wavefunctions(rng, ekpg, nb, mxddim=None):
    """psi = (N(0,1)+iN(0,1))/sqrt2 * exp(-|k+G|^2/8), Loewdin-orthonormalised."""
    mtxd = ekpg.shape[0]
    mxddim = mtxd if mxddim is None else mxddim
    z = (rng.standard_normal((mtxd, nb)) + 1j * rng.standard_normal((mtxd, nb))) / math.sqrt(2.0)
    z *= np.exp(-2.0 * ekpg / 8.0)[:, None]
    s = z.conj().T @ z
    w, u = np.linalg.eigh(s)
    z = z @ (u * (1.0 / np.sqrt(w))) @ u.conj().T
    psi = np.zeros((mxddim, nb), complex, order="F")
    psi[:mtxd] = z
    return psi


def make_config(cfg, nbands: int | None = None, kindex: int = 0, with_psi: bool = True,
                mxddim_pad: int = 0, kpt=None, kmax_override=None, idshift: int = 0,
                positions=None):
    """All hk_psi_c16 inputs for workload ``cfg`` (k-point ``kindex`` for cfg 3).

    Variants: ``kpt`` (fractional k, overrides the config's), ``kmax_override``
    (kmscr(1:3), smaller mesh), ``idshift`` (id > n1), ``mxddim_pad`` (rows
    beyond mtxd in every column), ``positions`` (list of fractional position
    arrays, one per species, replacing the config's: the random draws, hence
    psi and the tau-independent part of V, stay the same)."""
    spec = config_spec(cfg)
    if positions is not None:
        spec.types = [(name, np.asarray(pos, float), rc)
                      for (name, _, rc), pos in zip(spec.types, positions)]
    nb = spec.nbands if nbands is None else nbands
    k = spec.kpts[kindex] if kpt is None else kpt
    ks = build_kspace(spec, k, kmax_override, idshift)
    mtxd = ks["mtxd"]
    mxddim = mtxd + mxddim_pad
    kmscr = ks["kmscr"]
    ns = [int(x) for x in kmscr[3:6]]
    omega = abs(np.linalg.det(spec.avec))
    seed = 20260000 + (cfg if isinstance(cfg, int) else 99)
    rng = np.random.Generator(np.random.PCG64(seed))
    gq, vq = _random_potential(rng, spec, ks["kgv"], ks["bvec"], ns)
    floc_q, floc_f = None, None
    if spec.vloc_gauss is not None:
        gl, vl = vloc_from_species(spec, ks["kgv"], ks["bvec"], ns, omega)
        gq = np.concatenate([gq, gl])
        vq = np.concatenate([vq, vl])
        floc_q, floc_f = local_form_factors(spec, ks["kgv"], ks["bvec"], omega)
    vscr = potential_on_mesh(gq, vq, ns, int(kmscr[6]))
    anlga, xnlkb = projectors(spec, ks["kgv"], ks["isort"], ks["kpt"], ks["bvec"], omega, mxddim)
    psi = None
    if with_psi:
        krng = rng if kindex == 0 else np.random.Generator(
            np.random.PCG64(seed + 7919 * kindex))
        psi = wavefunctions(krng, ks["ekpg"], nb, mxddim)
    isort = np.zeros(mxddim, np.int32)
    isort[:mtxd] = ks["isort"]
    ekpg = np.zeros(mxddim)
    ekpg[:mtxd] = ks["ekpg"]
    return dict(cfg=cfg, spec=spec, mtxd=mtxd, neig=nb, psi=psi, ng=ks["ng"], kgv=ks["kgv"],
                ekpg=ekpg, isort=isort, vscr=vscr, kmscr=kmscr, anlga=anlga, xnlkb=xnlkb,
                nanl=anlga.shape[1], mxddim=mxddim, mxdbnd=nb, mxdanl=anlga.shape[1],
                mxdgve=ks["ng"], mxdscr=vscr.shape[0], vq_g=gq, vq=vq, omega=omega,
                bvec=ks["bvec"], bdot=ks["bdot"], kpt=ks["kpt"], avec=spec.avec,
                rat=np.concatenate([t[1] for t in spec.types]),
                species=np.concatenate([np.full(t[1].shape[0], k, np.int32)
                                        for k, t in enumerate(spec.types)]),
                proj_atom=projector_atoms(spec), floc_q=floc_q, floc_f=floc_f)


# --------------------------------------------------------------------------
# two-component (spinor) inputs of hk_psi_spin_c16
# --------------------------------------------------------------------------

def spin_inputs(d, nbands, seed=11):
    """Inputs of Src/hk_psi/reference/hk_psi_spin_c16.f90 built on a scalar
    config ``d``:

    * ``vscr_sp`` (mxdscr, 4): component 1 is d's V(r); components 2-4
      (sigma_z, sigma_x, sigma_y parts) are further band-limited random
      potentials at 0.3 of its amplitude, so every coupling is exercised;
    * ``anlsp`` (2*mxddim, 2*nanl), ``xnlkbsp``: each scalar projector beta_a
      gives the pair beta_a (x) U|up>, beta_a (x) U|down> (interleaved rows,
      U a fixed SU(2) rotation) with strengths xnlkb_a and 0.5*xnlkb_a, a
      spin-dependent non-local term like a spin-orbit one;
    * ``psi_sp`` (2*mxddim, nbands): two random wave-function sets
      interleaved, normalised per band."""
    spec = d["spec"]
    ns = [int(x) for x in d["kmscr"][3:6]]
    idd = int(d["kmscr"][6])
    rng = np.random.Generator(np.random.PCG64(seed))
    mxdscr = d["mxdscr"]
    vsp = np.zeros((mxdscr, 4), order="F")
    vsp[:, 0] = d["vscr"]
    for c in range(1, 4):
        gq, vq = _random_potential(rng, spec, d["kgv"], d["bvec"], ns)
        v = potential_on_mesh(gq, 0.3 * vq, ns, idd)
        vsp[: v.shape[0], c] = v
    mtxd, mxddim, nanl = d["mtxd"], d["mxddim"], d["nanl"]
    th, ph = 0.7, 1.3
    U = np.array([[np.cos(th / 2), -np.exp(-1j * ph) * np.sin(th / 2)],
                  [np.exp(1j * ph) * np.sin(th / 2), np.cos(th / 2)]])
    anlsp = np.zeros((2 * mxddim, 2 * nanl), np.complex128, order="F")
    xsp = np.zeros(2 * nanl)
    for a in range(nanl):
        b = d["anlga"][:mtxd, a]
        for s in range(2):
            col = 2 * a + s
            anlsp[0:2 * mtxd:2, col] = b * U[0, s]
            anlsp[1:2 * mtxd:2, col] = b * U[1, s]
            xsp[col] = d["xnlkb"][a] * (1.0 if s == 0 else 0.5)
    w = np.exp(-d["ekpg"][:mtxd] / 4)
    psi = np.zeros((2 * mxddim, nbands), np.complex128, order="F")
    for c in range(2):
        x = rng.normal(size=(mtxd, nbands)) + 1j * rng.normal(size=(mtxd, nbands))
        psi[c:2 * mtxd:2] = x * w[:, None]
    psi /= np.linalg.norm(psi, axis=0)
    return dict(vscr_sp=vsp, anlsp=anlsp, xnlkbsp=xsp, nanlsp=2 * nanl, psi_sp=psi,
                mxdasp=2 * nanl, mxdnsp=4)
