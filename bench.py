#!/usr/bin/env python3
"""H|psi> throughput on one MI355X: 64-atom Si, 128 bands, FP32-mixed vs FP64.

Workload (synthetic, seeded): Si64 cubic supercell L = 20.52 bohr, Ecut 15 Ha,
Gamma point, 23,847 plane waves, 72^3 mesh, 256 s+p Kleinman-Bylander
projectors, 128 bands.  One step = one application of H (kinetic + local
potential by FFT + non-local projectors) to all 128 bands, with psi, V(r) and
the projectors already resident in HBM.

``value`` is the FP32-mixed throughput (complex64 FFTs, V(r) product and
projector GEMMs with FP32 accumulation; psi, H|psi> and the kinetic term in
FP64).  It is reported only together with the eigenvalue error of the band
subspace (Rayleigh-Ritz in FP64) against the FP64 path of the same run; the
FP64 throughput is reported beside it.  Both Rayleigh-Ritz steps run in FP64 on
the GPU (cpwgpu_rayleigh_ritz); the refined error repeats the step with one
FP64 application of H to the FP32-mixed Ritz vectors.

Before timing, the library's autotuner picks the plane-stage FFT kernel and the
split-K shape of the projection GEMM for each precision on this workload
(reported under "tuned").  Per-kernel durations come from HIP start/stop
events carried in each launch's dispatch on the launch stream inside the timed
loops (library measurement hook): the kernels' own durations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H|ψ> bands/sec (64-atom Si, 128 bands): FP32-mixed vs FP64 at ≤1e-6 Ha"

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_work(info, nanl, prec):
    """Algorithmic bytes (HBM) and flops per band of each kernel.

    s_c, s_r: bytes of a complex / real in the kernel's arithmetic; psi and
    H|psi> cross the boundary as complex128 (16 B) in both modes."""
    s_c = 16 if prec == "fp64" else 8
    s_r = 8 if prec == "fp64" else 4
    mtxd, ncol = info["mtxd"], info["ncol"]
    n1, n2, n3 = info["n1"], info["n2"], info["n3"]
    col = s_c * ncol * n3                 # one pass over the z-column buffer
    return {
        "k_col_bwd": dict(bytes=16 * mtxd + col),
        "k_plane": dict(bytes=2 * col + s_r * n1 * n2 * n3),
        "k_col_fwd": dict(bytes=col + 16 * mtxd + 16 * mtxd),
        "k_proj": dict(flops=8 * mtxd * nanl),
        "k_back": dict(flops=8 * mtxd * nanl),
    }


def roofline_entry(name, work, avg_ms, nb, prec):
    w = work[name]
    if "bytes" in w:
        ach = w["bytes"] * nb / (avg_ms * 1e-3) / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": None,
                "avg_ms": round(avg_ms, 5), "algorithmic_bytes": w["bytes"] * nb}
    ach = w["flops"] * nb / (avg_ms * 1e-3) / 1e12
    pk = PEAK_TFLOPS[prec]
    return {"kernel": name, "bound": "mfma", "achieved": round(ach, 2), "peak": pk,
            "unit": "TFLOP/s", "frac": round(ach / pk, 4), "traffic": None,
            "avg_ms": round(avg_ms, 5), "algorithmic_flops": w["flops"] * nb}


def call_roofline(info, nanl, nb, prec, ms):
    """Whole-call roofline of one H|psi> over nb bands: the local stage at the
    HBM bound (per band s_c*2*mtxd for psi in and H psi out, 4*s_c*ncol*n3 for
    the z-column round trips of the two 3-D transforms, s_r*N for V(r)) plus the
    two projector GEMMs at the matrix-core bound (16*mtxd*nanl flop per band);
    frac = t_roof / t_measured."""
    s_c = 16 if prec == "fp64" else 8
    s_r = 8 if prec == "fp64" else 4
    mtxd, ncol = info["mtxd"], info["ncol"]
    n = info["n1"] * info["n2"] * info["n3"]
    b_loc = s_c * 2 * mtxd + 4 * s_c * ncol * info["n3"] + s_r * n
    f_nl = 16 * mtxd * nanl
    t_loc = b_loc * nb / (PEAK_HBM_GBS * 1e9) * 1e6
    t_nl = f_nl * nb / (PEAK_TFLOPS[prec] * 1e12) * 1e6
    return {"t_roof_us": round(t_loc + t_nl, 1), "t_local_us": round(t_loc, 1),
            "t_nonlocal_us": round(t_nl, 1), "t_measured_us": round(ms * 1e3, 1),
            "frac": round((t_loc + t_nl) / (ms * 1e3), 4),
            "roofline_bands_per_s": round(nb / ((t_loc + t_nl) * 1e-6), 1)}


def local_roofline(info, nb, prec, per):
    """The local stage alone (z-column inverse, plane stage, z-column forward
    and gather, one stream) against its HBM bound: the same per-band bytes as
    the local part of call_roofline; measured = sum of the three kernels'
    average HIP-event durations (ms) in `per`."""
    s_c = 16 if prec == "fp64" else 8
    s_r = 8 if prec == "fp64" else 4
    n = info["n1"] * info["n2"] * info["n3"]
    b_loc = s_c * 2 * info["mtxd"] + 4 * s_c * info["ncol"] * info["n3"] + s_r * n
    t_roof = b_loc * nb / (PEAK_HBM_GBS * 1e9) * 1e6
    parts = {k: round(per[k] * 1e3, 1) for k in ("k_col_bwd", "k_plane", "k_col_fwd")}
    t_meas = sum(parts.values())
    return {"t_roof_us": round(t_roof, 1), "t_measured_us": round(t_meas, 1), "kernels_us": parts,
            "achieved_gbs": round(b_loc * nb / (t_meas * 1e-6) / 1e9, 1), "frac": round(t_roof / t_meas, 4)}


def measured_traffic(prec, kernel, variant, n1):
    """HBM bytes per launch of a plane-stage, column or projector kernel
    measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated by a copy of
    known size) on this workload, recorded in profiles/roofline_traffic.json
    under "<prec>/<plane|col_bwd|col_fwd|proj|back><variant>/<n1>"; None if
    that file has no entry."""
    fam = {"k_plane": "plane", "k_proj": "proj", "k_back": "back", "k_col_bwd": "col_bwd",
           "k_col_fwd": "col_fwd"}.get(kernel)
    if fam is None:
        return None, None
    path = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None, None
    e = rec.get("kernels", {}).get(f"{prec}/{fam}{variant}/{n1}")
    if not e:
        return None, None
    return e["bytes"], f"rocprofv3 FETCH_SIZE+WRITE_SIZE of {e['kernel']} (profiles/roofline_traffic.json)"


def time_path(torch, hk, plan, psi, hpsi, prec, steps, warmup, kernels):
    """Times `steps` applies; returns ms/step and per-kernel average ms."""
    for _ in range(warmup):
        plan.apply(psi, hpsi, prec)
    torch.cuda.synchronize()
    hk.profile_read()
    hk.profile(True, kernels)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        plan.apply(psi, hpsi, prec)
    e1.record(stream)
    torch.cuda.synchronize()
    hk.profile(False)
    ms = e0.elapsed_time(e1) / steps
    per = {k: (v[0] / v[1]) for k, v in hk.profile_read().items() if v[1] > 0}
    return ms, per


def ritz(hk, psi_t, hpsi_t):
    """Ritz values of the band subspace: FP64 Rayleigh-Ritz on the GPU
    (Gram matrices on the matrix cores, Cholesky, Jacobi), inputs untouched."""
    eig, _ = hk.rayleigh_ritz(psi_t.clone(), hpsi_t.clone(), rotate=False, residuals=False)
    return eig


def cpu_baseline(d, seconds, nb=8):
    """The reference hk_psi_c16 (Fortran, NCAR FFT, netlib zgemm) when built,
    else the C restatement; bounded sample of nb bands per call."""
    import numpy as np

    from oracle import oracle as O

    psi = np.asfortranarray(d["psi"][:, :nb])
    if O.have_ref():
        kind, cores = "reference", 1
        fn = lambda: O.ref_hk_psi(d, psi=psi)
    else:
        kind, cores = "port", O.orc_lib().orc_num_threads()
        fn = lambda: O.orc_hk_psi(d, psi=psi)
    fn()
    t0 = time.perf_counter()
    calls = 0
    while True:
        fn()
        calls += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    out = {"value": round(calls * nb / dt, 2), "unit": "bands/s", "cores": cores, "kind": kind,
           "sample": f"{calls} calls x {nb} bands of the same 64-atom Si workload ({dt:.1f} s)"}

    def port_rate(threads):
        lib = O.orc_lib()
        prev = lib.orc_num_threads()
        lib.orc_set_num_threads(threads)
        try:
            O.orc_hk_psi(d, psi=psi)
            t0 = time.perf_counter()
            n = 0
            while time.perf_counter() - t0 < seconds / 2:
                O.orc_hk_psi(d, psi=psi)
                n += 1
            return round(n * nb / (time.perf_counter() - t0), 2), lib.orc_num_threads()
        finally:
            lib.orc_set_num_threads(prev)

    # the C restatement on one core (calibrates it against the reference on
    # the same core) and with OpenMP on the host cores
    v1, _ = port_rate(1)
    out["port_1core"] = {"value": v1, "cores": 1}
    if kind == "reference":
        out["reference_over_port_1core"] = round(out["value"] / v1, 3)
    vn, nth = port_rate(O.orc_lib().orc_num_threads())
    out["port_openmp"] = {"value": vn, "cores": nth}
    return out


def projector_cpu_baseline(d, pp, seconds):
    """The reference proj_nl_kb_c16 (flang build, one core) when built, else the
    numpy restatement, on the same inputs as the device build; bounded sample.
    Beside it ("openmp"), the reference's OpenMP build on every host core."""
    import ctypes as C

    import numpy as np

    from oracle import oracle as O
    from oracle import proj_kb as PK

    m, nanl = int(d["mtxd"]), int(pp["nanl"])
    kgv = np.asfortranarray(d["kgv"], np.int32)
    isort = np.ascontiguousarray(d["isort"], np.int32)
    rk = np.ascontiguousarray(d["kpt"], np.float64)
    anl = np.zeros((m, nanl), complex, order="F")
    xnl = np.zeros(nanl)
    vkb = pp["vkb"]
    if O.have_ref():
        lib = O.ref_lib()
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        i = lambda v: C.byref(C.c_int(int(v)))
        kind = "reference"
        fn = lambda: lib.proj_nl_kb_c16_(
            p(rk), i(m), p(isort), i(nanl), i(d["ng"]), p(kgv), p(pp["nqnl"]), p(pp["delqnl"]), p(vkb),
            p(pp["nkb"]), i(pp["ntype"]), p(pp["natom"]), p(pp["rat"]), p(pp["adot"]), p(anl), p(xnl),
            i(vkb.shape[3]), i(pp["rat"].shape[1]), i(vkb.shape[0] - 3), i(m), i(nanl), i(kgv.shape[1]))
    else:
        kind = "port"
        fn = lambda: PK.proj_nl_kb_c16(rk, m, isort, nanl, d["ng"], kgv, pp["nqnl"], pp["delqnl"], vkb,
                                       pp["nkb"], pp["ntype"], pp["natom"], pp["rat"], pp["adot"], m, nanl,
                                       kgv.shape[1])
    def timed(f):
        f()
        t0 = time.perf_counter()
        calls = 0
        while True:
            f()
            calls += 1
            dt = time.perf_counter() - t0
            if dt >= seconds:
                return dt, calls

    dt, calls = timed(fn)
    out = {"ms_per_call": round(dt / calls * 1e3, 2), "cores": 1, "kind": kind,
           "sample": f"{calls} calls on the same inputs ({dt:.1f} s)"}
    lomp = O.ref_omp_lib()
    if kind == "reference" and lomp is not None:
        fo = lambda: lomp.proj_nl_kb_c16_(
            p(rk), i(m), p(isort), i(nanl), i(d["ng"]), p(kgv), p(pp["nqnl"]), p(pp["delqnl"]), p(vkb),
            p(pp["nkb"]), i(pp["ntype"]), p(pp["natom"]), p(pp["rat"]), p(pp["adot"]), p(anl), p(xnl),
            i(vkb.shape[3]), i(pp["rat"].shape[1]), i(vkb.shape[0] - 3), i(m), i(nanl), i(kgv.shape[1]))
        dt, calls = timed(fo)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        out["openmp"] = {"ms_per_call": round(dt / calls * 1e3, 2), "cores": threads,
                         "sample": f"{calls} calls ({dt:.1f} s)"}
    return out


def bench_kpoints(torch, hk, synth, np, steps):
    """16 k-points of the 64-atom cell on a shifted 2x2x4 grid (uneven
    plane-wave counts), 128 bands each: the eager k loop against one launch of
    the HIP graph that holds all of them.  Kernel choices come from tuning the
    first k-point (same mesh and sizes for all)."""
    spec = synth.config_spec(3)
    nk = len(spec.kpts)
    plans, psis, outs, mt = [], [], [], []
    for k in range(nk):
        d = synth.make_config(3, kindex=k)
        p = hk.Plan.from_inputs(d)
        x = torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")
        plans.append(p)
        psis.append(x)
        outs.append(torch.zeros_like(x))
        mt.append(int(d["mtxd"]))
    nb = psis[0].shape[0]
    out = {"k_points": nk, "bands_per_k": nb, "plane_waves": [min(mt), max(mt)]}
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        c = plans[0].tune(psis[0], outs[0], pid)
        for p in plans[1:]:
            for what, key in ((hk.TUNE_PLANE, "plane"), (hk.TUNE_PROJ, "proj"),
                              (hk.TUNE_BACK, "back"), (hk.TUNE_SPLITK, "splitk"), (hk.TUNE_KIN, "kin"),
                              (hk.TUNE_COL, "col"), (hk.TUNE_SCHED, "sched")):
                p.set_variant(pid, what, c[key])
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        for _ in range(2):
            for p, x, h in zip(plans, psis, outs):
                p.apply(x, h, pid)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            for p, x, h in zip(plans, psis, outs):
                p.apply(x, h, pid)
        e1.record()
        torch.cuda.synchronize()
        eager = e0.elapsed_time(e1) / steps
        g = hk.Graph(plans, psis, outs, pid)
        for _ in range(2):
            g.launch()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            g.launch()
        e1.record()
        torch.cuda.synchronize()
        graph = e0.elapsed_time(e1) / steps
        g.close()
        out[prec] = {"eager_ms": round(eager, 4), "graph_ms": round(graph, 4),
                     "eager_bands_per_s": round(nk * nb / (eager * 1e-3), 1),
                     "graph_bands_per_s": round(nk * nb / (graph * 1e-3), 1), "tuned": c}
    for p in plans:
        p.close()
    return out


def bench_ditsp(torch, hk, plan, d, nb, np):
    """Device-resident iterative diagonalisation of the 64-atom cell: the
    lowest nb bands to max residual^2 < epspsi^2 (epspsi = 1e-6), FP64 and
    FP32-mixed; wall time of the whole call.  "cold": from the small-matrix
    start.  "scf_step": from the converged bands of a potential 2% weaker,
    as in a self-consistency step.  Default iteration (Chebyshev-filtered
    subspace iteration); "block_fp64_cold": the locally optimal block
    iteration on the same cold start, for comparison."""
    mtxd = d["mtxd"]
    out = {}
    eis = {}
    # the default iteration (Chebyshev-filtered), then the block iteration on the same start
    hk.set_ditsp_method(hk.DITSP_BLOCK)
    psi = torch.zeros((nb, mtxd), dtype=torch.complex128, device="cuda")
    hpsi = torch.zeros_like(psi)
    plan.ditsp(psi, hpsi, prec=hk.PREC_FP64, icmax=2, epspsi=1e-6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eib, ifail, st = plan.ditsp(psi, hpsi, prec=hk.PREC_FP64, icmax=200, epspsi=1e-6)
    dt = time.perf_counter() - t0
    out["block_fp64_cold"] = {"seconds": round(dt, 4), "ifail": ifail, **st}
    log(f"bench: ditsp block iteration fp64 cold: {dt * 1e3:.1f} ms, {st['iters']} iterations, ifail {ifail}")
    hk.set_ditsp_method(hk.DITSP_CHEB)
    del psi, hpsi
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        psi = torch.zeros((nb, mtxd), dtype=torch.complex128, device="cuda")
        hpsi = torch.zeros_like(psi)
        plan.ditsp(psi, hpsi, prec=pid, icmax=2, epspsi=1e-6)  # warm-up (handles, buffers)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ei, ifail, st = plan.ditsp(psi, hpsi, prec=pid, icmax=200, epspsi=1e-6)
        dt = time.perf_counter() - t0
        eis[prec] = ei
        out[prec] = {"cold": {"seconds": round(dt, 4), "ifail": ifail, **st}}
        log(f"bench: ditsp {prec} cold: {dt * 1e3:.1f} ms, {st['iters']} iterations, "
            f"{st['hpsi_products']} H|psi> products, ifail {ifail}")
        # self-consistency step: previous bands of V, now V*1.02
        plan.set_potential(np.asfortranarray(d["vscr"] * 0.98))
        plan.ditsp(psi, hpsi, prec=pid, icmax=200, iguess=1, epspsi=1e-6)
        plan.set_potential(d["vscr"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ei2, ifail, st = plan.ditsp(psi, hpsi, prec=pid, icmax=200, iguess=1, epspsi=1e-6)
        dt = time.perf_counter() - t0
        out[prec]["scf_step"] = {"seconds": round(dt, 4), "ifail": ifail, **st,
                                 "eig_vs_cold_ha": float(np.abs(ei2 - ei).max())}
        log(f"bench: ditsp {prec} scf step: {dt * 1e3:.1f} ms, {st['iters']} iterations, "
            f"{st['hpsi_products']} H|psi> products, ifail {ifail}")
        del psi, hpsi
    out["eig_fp32_vs_fp64_ha"] = float(np.abs(eis["fp64"] - eis["fp32"]).max())
    out["eig_vs_block_ha"] = float(np.abs(eis["fp64"] - eib).max())
    out["method"] = "Chebyshev-filtered subspace iteration"
    out["bands"] = nb
    out["epspsi"] = 1e-6
    try:
        from cpw2000_amd import synth

        out["cpu_baseline"] = ditsp_cpu_baseline(hk, synth, np)
    except (OSError, RuntimeError) as exc:
        log(f"bench: reference ditsp unavailable ({exc})")
        out["cpu_baseline"] = None
    return out


def ditsp_cpu_baseline(hk, synth, np, cfg=0, neig=8, epspsi=1e-6):
    """The reference ditsp_c16 (flang build of Src/ditsp with its vendored
    BLAS / LAPACK and the reference hk_psi_c16, one core) against the device
    drop-in cpwgpu_ditsp_c16 on the same host arrays: neig bands of the
    2-atom cell to max residual^2 < epspsi^2 from the same small matrix
    (leading block of H, built here by applying H to unit vectors) and diag(H).
    Wall time of each call and the largest eigenvalue difference."""
    from oracle import oracle as O

    if not O.have_ref():
        return None
    d = synth.make_config(cfg, nbands=2)
    m, mxddim = d["mtxd"], d["mxddim"]
    ms = 3 * neig + 20

    def args():
        psi = np.zeros((mxddim, neig), np.complex128, order="F")
        return dict(ipr=0, icmax=1000, iguess=0, epspsi=epspsi, lnewanl=True, neig=neig, mtxd=m, mtxds=ms,
                    mxdsml=ms, mxdbnd=neig, psi=psi, hpsi=np.zeros_like(psi, order="F"), ei=np.zeros(neig),
                    ekpg=np.asfortranarray(d["ekpg"]), isort=np.asfortranarray(d["isort"], np.int32),
                    vscr=np.asfortranarray(d["vscr"]), kmscr=np.asfortranarray(d["kmscr"], np.int32), ng=d["ng"],
                    kgv=np.asfortranarray(d["kgv"], np.int32), anlga=np.asfortranarray(d["anlga"], np.complex128),
                    xnlkb=np.asfortranarray(d["xnlkb"]), nanl=d["nanl"], hamsm=hs, hdiag=hd, mxddim=mxddim,
                    mxdgve=d["mxdgve"], mxdscr=d["mxdscr"], mxdanl=max(d["mxdanl"], 1))

    unit = np.zeros((mxddim, ms), np.complex128, order="F")
    unit[np.arange(ms), np.arange(ms)] = 1.0
    _, he = O.orc_hk_psi(d, psi=unit)
    hs = np.asfortranarray(he[:ms, :ms])
    n = int(np.prod(d["kmscr"][3:6]))
    hd = np.zeros(mxddim, order="F")
    hd[:m] = (d["ekpg"][:m] + np.mean(d["vscr"][:n])
              + (np.abs(d["anlga"][:m, :d["nanl"]]) ** 2) @ d["xnlkb"][:d["nanl"]])
    a = args()
    t0 = time.perf_counter()
    ifail_ref = O.ref_ditsp_c16(a)
    t_ref = time.perf_counter() - t0
    hk.set_precision(hk.PREC_FP64)
    b = args()
    hk.ditsp_c16(**{k: v for k, v in b.items()})  # warm-up (plan, handles)
    b = args()
    t0 = time.perf_counter()
    ifail_dev = hk.ditsp_c16(**b)
    t_dev = time.perf_counter() - t0
    return {"workload": f"2-atom Si, {m} plane waves, {neig} bands, epspsi {epspsi}",
            "reference_s": round(t_ref, 4), "reference_ifail": ifail_ref, "cores": 1, "kind": "reference",
            "device_dropin_s": round(t_dev, 4), "device_ifail": ifail_dev,
            "eig_diff_ha": float(np.abs(a["ei"] - b["ei"]).max())}


def bench_spin_density(torch, hk, synth, np, d, steps):
    """Two-component H|psi> (non-collinear potential, nsp = 4, spin-mixing
    projectors) on 64 spinor bands of the 64-atom cell, tuned by the library's
    two-component tuner (then each coupled plane kernel timed: 0 LDS two-pass,
    1 register lines), and the electron density of its 128 scalar bands;
    device arrays, per-call time."""
    mtxd = d["mtxd"]
    s = synth.spin_inputs(d, 64)
    plan = hk.Plan.from_inputs(d, projectors=False)
    plan.set_potential_spin(s["vscr_sp"], 4)
    plan.set_projectors_spin(s["anlsp"], s["xnlkbsp"], s["nanlsp"])
    pt = torch.from_numpy(np.ascontiguousarray(s["psi_sp"][: 2 * mtxd].T)).to("cuda")
    ht = torch.zeros_like(pt)
    out = {"spinor_bands": 64, "nsp": 4}
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        tuned = plan.tune_spin(pt, ht, pid)
        per = {}
        for kern in (0, 1):
            plan.set_variant(pid, hk.TUNE_SPIN, kern)
            for _ in range(3):
                plan.apply_spin(pt, ht, pid)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                plan.apply_spin(pt, ht, pid)
            torch.cuda.synchronize()
            per[kern] = (time.perf_counter() - t0) / steps * 1e3
        plan.set_variant(pid, hk.TUNE_SPIN, tuned["spin"])
        ms = per[tuned["spin"]]
        out[prec] = {"ms_per_call": round(ms, 4), "spinor_bands_per_s": round(64 / (ms * 1e-3), 1),
                     "tuned": tuned, "ms_per_plane_kernel": {str(k): round(v, 4) for k, v in per.items()}}
    plan.close()
    plan = hk.Plan.from_inputs(d)
    psi = torch.from_numpy(np.ascontiguousarray(d["psi"][:mtxd].T)).to("cuda")
    occ = np.full(psi.shape[0], 2.0)
    for _ in range(3):
        plan.density(psi, occ)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.density(psi, occ)
    torch.cuda.synchronize()
    out["density_128_bands_ms"] = round((time.perf_counter() - t0) / steps * 1e3, 4)
    plan.close()
    return out


def bench_forces(torch, hk, synth, np, steps):
    """Atomic forces at fixed psi (adjoint of E_loc + E_NL, cpwgpu_plan_forces)
    on the two-species 64-atom Si/Ge cell with 128 occupied bands: per-call
    time including the band density, its transform to the potential's q list,
    the structure-factor sums and the weighted projection GEMMs; host results."""
    d = synth.make_config(4)
    plan = hk.Plan.from_inputs(d)
    psi = torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")
    args = (d["rat"], d["avec"], d["proj_atom"], d["floc_q"], d["floc_f"], d["species"])
    for _ in range(2):
        f, _ = plan.forces(psi, *args)
    t0 = time.perf_counter()
    for _ in range(steps):
        f, enl = plan.forces(psi, *args)
    ms = (time.perf_counter() - t0) / steps * 1e3
    plan.close()
    return {"workload": "Si/Ge 64 atoms L=21.0 bohr, Ecut 15 Ha, 128 bands", "atoms": int(d["rat"].shape[0]),
            "bands": int(psi.shape[0]), "ms_per_call": round(ms, 3),
            "max_force_ha_bohr": round(float(np.abs(f).max()), 6)}


def bench_projectors(torch, hk, synth, np, d, steps, cfg=1,
                     workload="Si64 L=20.52, Ecut 15 Ha: 256 s+p KB projectors on 23,847 plane waves",
                     cpu_seconds=None):
    """KB projectors built on the device (cpwgpu_plan_build_projectors_kb, the
    construction of Src/omp/proj_nl_kb_c16.f90) for the cell of ``d``
    (default the 64-atom Si cell: 256 s and p projectors on 23,847 plane
    waves) from interpolation tables of 4,242 points per (species, l).  Kernel
    time from HIP events in each launch's dispatch; per-call time includes the table
    upload and the synchronisation.  Algorithmic HBM bytes per launch: the
    projectors written once (16 B per element) plus g of each plane wave read
    once (24 B)."""
    spec = synth.config_spec(cfg)
    pp = synth.kb_pseudo(spec)
    plan = hk.Plan.from_inputs(d, projectors=False)
    args = (np.zeros(3), pp["nanl"], pp["nqnl"], pp["delqnl"], pp["vkb"], pp["nkb"], pp["ntype"], pp["natom"],
            pp["rat"], pp["adot"])
    for _ in range(2):
        plan.build_projectors_kb(*args)
    hk.profile_read()
    hk.profile(True, ["k_proj_kb"])
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.build_projectors_kb(*args)
    ms_call = (time.perf_counter() - t0) / steps * 1e3
    hk.profile(False)
    ms, cnt = hk.profile_read()["k_proj_kb"]
    plan.close()
    mtxd, nanl = int(d["mtxd"]), int(pp["nanl"])
    us = ms / max(cnt, 1) * 1e3
    nbytes = 16 * mtxd * nanl + 24 * mtxd
    gbs = nbytes / (us * 1e-6) / 1e9
    # HBM bytes per launch measured with rocprofv3 on this shape, when recorded
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "roofline_traffic.json")) as f:
            e = json.load(f).get("kernels", {}).get(f"fp64/proj_kb/{mtxd}x{nanl}")
        traffic = e["bytes"] if e else None
    except (OSError, ValueError):
        pass
    cpu = projector_cpu_baseline(d, pp, cpu_seconds) if cpu_seconds else None
    return {"workload": workload, "nanl": nanl, "cpu_baseline": cpu, "mtxd": mtxd, "table_points": int(pp["nqnl"][0]), "launches": int(cnt),
            "us_per_kernel": round(us, 2), "ms_per_call": round(ms_call, 3),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(gbs / PEAK_HBM_GBS, 3), "traffic": traffic, "bytes_per_launch": nbytes}}


def bench_cfg2(torch, hk, synth, np, steps):
    """216-atom Si (108^3 mesh, 864 projectors, 512 bands): throughput of both
    paths and the FP32-mixed error against FP64 (H psi norm and Ritz values of
    the band subspace; subspace matrices in FP64 on the GPU, eigenvalues of the
    512 x 512 problem by LAPACK on the host)."""
    import scipy.linalg as sla

    d = synth.make_config(2)
    plan = hk.Plan.from_inputs(d)
    x = torch.from_numpy(np.ascontiguousarray(d["psi"][: d["mtxd"]].T)).to("cuda")
    nb = x.shape[0]
    h = {}
    out = {"mtxd": int(d["mtxd"]), "bands": nb, "nanl": int(d["nanl"]), "mesh": 108}
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        h[prec] = torch.zeros_like(x)
        c = plan.tune(x, h[prec], pid)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(steps):
            plan.apply(x, h[prec], pid)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        out[prec] = {"ms_per_step": round(ms, 4), "bands_per_s": round(nb / (ms * 1e-3), 1),
                     "tuned": c}
    out["fp32_over_fp64"] = round(out["fp64"]["ms_per_step"] / out["fp32"]["ms_per_step"], 3)
    out["projectors_kb"] = bench_projectors(
        torch, hk, synth, np, d, 10, cfg=2,
        workload="Si216 L=30.78, Ecut 15 Ha: 864 s+p KB projectors on 80,725 plane waves")
    out["hpsi_rel_err_fp32"] = float(torch.linalg.norm(h["fp32"] - h["fp64"]) /
                                     torch.linalg.norm(h["fp64"]))
    s_mat = (x.conj() @ x.T).cpu().numpy()
    eig = {}
    for prec in ("fp64", "fp32"):
        m = (x.conj() @ h[prec].T).cpu().numpy()
        eig[prec] = sla.eigh(0.5 * (m + m.conj().T), s_mat, eigvals_only=True)
    out["eig_err_ha"] = float(np.abs(eig["fp32"] - eig["fp64"]).max())
    out["eig_tol_ha"] = 1e-6
    out["eig_ok"] = out["eig_err_ha"] <= 1e-6
    plan.close()
    return out


def issue_bound(prec, kernel, variant, n1, avg_ms, rates):
    """Instruction-issue floor of a plane-stage kernel: its wave-level VALU
    instructions per launch by kind (rocprofv3 SQ_INSTS_VALU_* counters,
    profiles/roofline_traffic.json "valu") times the issue cost of each kind
    measured on this GPU (cycles per wave instruction per SIMD), over all SIMDs
    at the measured shader clock; frac = floor / measured duration."""
    if kernel != "k_plane" or not rates or "v_add_f32" not in rates:
        return None
    try:
        with open(os.path.join(ROOT, "profiles", "roofline_traffic.json")) as f:
            e = json.load(f).get("valu", {}).get(f"{prec}/plane{variant}/{n1}")
    except (OSError, ValueError):
        return None
    if not e:
        return None
    # DPP lane exchanges (counted by the hardware as FP32 adds or as moves,
    # "dpp_in") issue at the DPP rate; the remaining 32-bit work at the plain rates
    dpp = e.get("dpp", 0)
    f32 = e["f32"] - (dpp if e.get("dpp_in") == "f32" else 0)
    other = e["other"] - (dpp if e.get("dpp_in") == "other" else 0)
    r64 = 0.5 * (rates["v_fma_f64"] + rates["v_add_f64"])
    cyc = (e["f64"] * r64 + dpp * rates["v_mov_b32_dpp"] + f32 * rates["v_fma_f32"]
           + e["int32"] * rates["v_add_u32"] + other * rates["v_add_f32"]) / rates["simds"]
    floor_us = cyc / (rates["shader_clock_ghz"] * 1e3)
    return {"kernel": e["kernel"], "valu_instr": e["valu"], "dpp_instr": dpp, "cycles_per_simd": round(cyc),
            "clock_ghz": rates["shader_clock_ghz"], "floor_us": round(floor_us, 1),
            "measured_us": round(avg_ms * 1e3, 1), "frac": round(floor_us / (avg_ms * 1e3), 4),
            "source": "SQ_INSTS_VALU_* of profiles/roofline_traffic.json x bench.py --issue-rates"}


ISSUE_KINDS = ("v_fma_f64", "v_add_f64", "v_fma_f32", "v_add_u32", "v_mov_b32_dpp", "v_fma_f64+v_mov_b32_dpp",
               "v_add_f32", "v_add_f32_dpp", "v_permlane32_swap", "v_permlane16_swap", "v_fmac_f32_dpp")


def issue_rates():
    """Cycles per wave instruction per SIMD of the VALU instruction kinds the
    FFT kernels are made of, measured on this GPU (libcpwprobe.so), and the
    shader clock during the measurement."""
    import ctypes as C

    import torch  # noqa: F401  (one HIP runtime in the process: torch's)

    lib = C.CDLL(os.path.join(ROOT, "cpw2000_amd", "libcpwprobe.so"))
    nk = len(ISSUE_KINDS)
    res = (C.c_double * (nk + 2))()
    st = lib.cpwprobe_issue_rates(res, nk + 2)
    if st != 0:
        raise RuntimeError(f"cpwprobe_issue_rates: HIP error {st}")
    out = {k: round(res[i], 3) for i, k in enumerate(ISSUE_KINDS)}
    out["shader_clock_ghz"] = round(res[nk], 3)
    out["simds"] = int(res[nk + 1])
    return out


def variant_sweep(torch, hk, plan, psi, info, nanl, nb, reps):
    """Standalone (one stream) duration of every plane-stage, projection and
    back-projection kernel built, both precisions, by HIP events in each
    launch's dispatch; then the 512 MiB calibration copy.  For profiling these kernels one
    by one."""
    hk.set_overlap(False)
    h = torch.zeros_like(psi)
    out = {}
    for prec, pid in (("fp32", hk.PREC_FP32), ("fp64", hk.PREC_FP64)):
        work = kernel_work(info, nanl, prec)
        for what, fam, nv in ((hk.TUNE_PLANE, "plane", hk.NPLANE_VAR), (hk.TUNE_PROJ, "proj", 6),
                              (hk.TUNE_BACK, "back", 5)):
            for v in range(nv):
                try:
                    plan.set_variant(pid, what, v)
                except hk.CpwGpuError:
                    continue  # plane shape not built for this mesh
                for _ in range(2):
                    plan.apply(psi, h, pid)
                torch.cuda.synchronize()
                hk.profile(True, ["k_" + fam])
                hk.profile_read()
                for _ in range(reps):
                    plan.apply(psi, h, pid)
                torch.cuda.synchronize()
                ms, cnt = hk.profile_read()["k_" + fam]
                hk.profile(False)
                e = roofline_entry("k_" + fam, work, ms / max(cnt, 1), nb, prec)
                out[f"{prec}/{fam}{v}"] = {"us": round(1e3 * ms / max(cnt, 1), 1), "achieved": e["achieved"],
                                           "unit": e["unit"], "frac": e["frac"]}
                log(f"bench: {prec} {fam} variant {v}: {out[f'{prec}/{fam}{v}']}")
            plan.set_variant(pid, what, -1)
        for v in (0, 1):  # column kernels: LDS lines, register lines
            try:
                plan.set_variant(pid, hk.TUNE_COL, v)
            except hk.CpwGpuError:
                continue
            for _ in range(2):
                plan.apply(psi, h, pid)
            torch.cuda.synchronize()
            hk.profile(True, ["k_col_bwd", "k_col_fwd"])
            hk.profile_read()
            for _ in range(reps):
                plan.apply(psi, h, pid)
            torch.cuda.synchronize()
            r = hk.profile_read()
            hk.profile(False)
            for k in ("k_col_bwd", "k_col_fwd"):
                ms, cnt = r[k]
                e = roofline_entry(k, work, ms / max(cnt, 1), nb, prec)
                out[f"{prec}/{k}{v}"] = {"us": round(1e3 * ms / max(cnt, 1), 1), "achieved": e["achieved"],
                                         "unit": e["unit"], "frac": e["frac"]}
                log(f"bench: {prec} {k} variant {v}: {out[f'{prec}/{k}{v}']}")
        plan.set_variant(pid, hk.TUNE_COL, -1)
    hk.set_overlap(True)
    x = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    x.fill_(1.0)
    y = x.clone()  # 512 MiB read + 512 MiB written
    torch.cuda.synchronize()
    del x, y
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kpoints", action="store_true", help="skip the 16 k-point graph section")
    ap.add_argument("--no-ditsp", action="store_true",
                    help="skip the iterative diagonalisation section")
    ap.add_argument("--no-si216", action="store_true",
                    help="skip the 216-atom cell (108^3 mesh, 512 bands)")
    ap.add_argument("--variant-sweep", action="store_true",
                    help="only time every projection / back-projection kernel alone (one "
                         "stream), then the calibration copy; prints their JSON and exits")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the 64-atom headline (tuning, accuracy, timed loops, rooflines): "
                         "for profiling its kernels")
    ap.add_argument("--only", choices=("ditsp", "spinor", "forces", "kpoints", "projectors", "si216"),
                    help="run only this section (after tuning the 64-atom plan) and print its JSON; "
                         "for profiling one path at a time")
    ap.add_argument("--issue-rates", action="store_true",
                    help="only measure the issue cost of the FFT kernels' VALU instruction kinds "
                         "(cycles per wave instruction per SIMD) and the shader clock; prints JSON")
    ap.add_argument("--set", action="append", default=[], metavar="PREC.KIND=V",
                    help="after tuning, override one choice, e.g. fp32.proj=2 (kinds: plane, "
                         "proj, back, splitk, kin, col, sched): for profiling a configuration "
                         "the tuner may or may not pick")
    ap.add_argument("--calib-copy", action="store_true",
                    help="after the timed loops, one plain 512 MiB device copy (a known-size "
                         "dispatch that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    overrides = []
    for item in args.set:
        try:
            lhs, val = item.split("=")
            prec, kind = lhs.split(".")
            overrides.append((prec, kind, int(val)))
        except ValueError:
            ap.error(f"--set {item!r}: expected PREC.KIND=V")
        if prec not in ("fp32", "fp64") or kind not in ("plane", "proj", "back", "splitk", "kin", "col", "sched"):
            ap.error(f"--set {item!r}: PREC is fp32 or fp64, KIND one of plane proj back splitk kin col sched")
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only (--gpus 1): H|psi> of independent "
              "bands and k-points shards across GPUs with no communication inside "
              "the apply, so the one-GPU rate is the per-GPU figure.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from cpw2000_amd import hk_psi as hk
    from cpw2000_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.issue_rates:
        print(json.dumps({"issue_rates": issue_rates()}), flush=True)
        return
    log("bench: generating 64-atom Si workload")
    d = synth.make_config(1)
    plan = hk.Plan.from_inputs(d)
    info = plan.info()
    mtxd, nb, nanl = d["mtxd"], d["neig"], d["nanl"]
    psi = torch.from_numpy(np.ascontiguousarray(d["psi"][:mtxd].T)).to("cuda")
    h64 = torch.zeros_like(psi)
    h32 = torch.zeros_like(psi)
    if args.variant_sweep:
        print(json.dumps({"variant_sweep": variant_sweep(torch, hk, plan, psi, info, nanl, nb, 10)}), flush=True)
        plan.close()
        return

    # autotuner: plane-stage kernel and split-K target per precision, timed on
    # this workload (part of the warm-up; every later apply uses the choice)
    if args.only in ("spinor", "forces", "kpoints", "projectors", "si216"):
        # these sections build and tune their own plans
        steps = max(5, args.steps // 3)
        sec = {"spinor": lambda: bench_spin_density(torch, hk, synth, np, d, steps),
               "forces": lambda: bench_forces(torch, hk, synth, np, 5),
               "projectors": lambda: bench_projectors(torch, hk, synth, np, d, max(10, args.steps),
                                                      cpu_seconds=None if args.no_cpu else 3.0),
               "kpoints": lambda: bench_kpoints(torch, hk, synth, np, max(3, args.steps // 6)),
               "si216": lambda: bench_cfg2(torch, hk, synth, np, max(3, args.steps // 6))}[args.only]
        print(json.dumps({args.only: sec()}), flush=True)
        plan.close()
        return
    tuned = {}
    for prec, pid, h in (("fp64", hk.PREC_FP64, h64), ("fp32", hk.PREC_FP32, h32)):
        tuned[prec] = plan.tune(psi, h, pid)
    for prec, kind, val in overrides:
        pid = {"fp64": hk.PREC_FP64, "fp32": hk.PREC_FP32}[prec]
        what = {"plane": hk.TUNE_PLANE, "proj": hk.TUNE_PROJ, "back": hk.TUNE_BACK, "splitk": hk.TUNE_SPLITK,
                "kin": hk.TUNE_KIN, "col": hk.TUNE_COL, "sched": hk.TUNE_SCHED}[kind]
        plan.set_variant(pid, what, val)
        tuned[prec][kind] = plan.get_variant(pid, what)
        log(f"bench: {prec} {kind} set to {tuned[prec][kind]}")
    if args.only == "ditsp":
        print(json.dumps({"ditsp": bench_ditsp(torch, hk, plan, d, nb, np), "tuned": tuned}), flush=True)
        plan.close()
        return
    log(f"bench: tuned kernels {tuned}")

    # accuracy of the FP32-mixed path against FP64 (same inputs, same run)
    plan.apply(psi, h64, hk.PREC_FP64)
    plan.apply(psi, h32, hk.PREC_FP32)
    torch.cuda.synchronize()
    eig64 = ritz(hk, psi, h64)
    eig32 = ritz(hk, psi, h32)
    eig_err = float(np.abs(eig32 - eig64).max())
    herr = float(torch.linalg.norm(h32 - h64) / torch.linalg.norm(h64))
    # FP64 refinement: Ritz vectors of the FP32-mixed step, one FP64 H pass on them
    p32, hp32 = psi.clone(), h32.clone()
    hk.rayleigh_ritz(p32, hp32, rotate=True, residuals=False)
    hr = torch.zeros_like(p32)
    plan.apply(p32, hr, hk.PREC_FP64)
    eig_ref = ritz(hk, p32, hr)
    eig_err_ref = float(np.abs(eig_ref - eig64).max())
    del p32, hp32, hr
    log(f"bench: FP32-mixed eigenvalue error {eig_err:.2e} Ha (after FP64 refinement "
        f"{eig_err_ref:.1e}), |dH psi|/|H psi| {herr:.2e}")

    # calibration: every kernel's duration (not the timed loops)
    cal = {}
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        _, cal[prec] = time_path(torch, hk, plan, psi, h64 if pid == 0 else h32, pid, 3, 1, None)
    work = {p: kernel_work(info, nanl, p) for p in ("fp64", "fp32")}
    # the local stage on one stream (no projector GEMMs beside it), and every
    # kernel's duration with nothing beside it (picks the dominant kernel)
    hk.set_overlap(False)
    local, alone = {}, {}
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        _, alone[prec] = time_path(torch, hk, plan, psi, h64 if pid == 0 else h32, pid, 3, 1, None)
        _, per = time_path(torch, hk, plan, psi, h64 if pid == 0 else h32, pid, 5, 1,
                           ["k_col_bwd", "k_plane", "k_col_fwd"])
        local[prec] = local_roofline(info, nb, prec, per)
    hk.set_overlap(True)
    log(f"bench: local stage alone {local}")

    def dominant(prec):
        # the kernel with the most work: the longest one alone (under overlap
        # the kernels beside each other stretch by turns, with the schedule)
        return max((k for k in alone[prec] if k in work[prec]), key=lambda k: alone[prec][k])

    res = {}
    for prec, pid in (("fp64", hk.PREC_FP64), ("fp32", hk.PREC_FP32)):
        dk = dominant(prec)
        log(f"bench: timing {prec} ({args.steps} steps), dominant kernel {dk}")
        ms, per = time_path(torch, hk, plan, psi, h64 if pid == 0 else h32, pid, args.steps,
                            args.warmup, [dk])
        res[prec] = dict(ms=ms, bands_per_s=nb / (ms * 1e-3), dominant=dk, dom_ms=per[dk])
        log(f"bench: {prec}: {ms:.3f} ms/step, {nb / (ms * 1e-3):.0f} bands/s")

    kernels = {p: {k: round(v, 5) for k, v in cal[p].items()} for p in cal}
    out = {
        "metric": METRIC,
        "value": round(res["fp32"]["bands_per_s"], 1),
        "unit": "bands/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["fp32"]["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (FP32 FFT/V/projector GEMMs, FP32 accumulation; FP64 psi, H psi, kinetic)",
        "data": "synthetic (seeded generator, 64-atom Si)",
        "config": {"workload": "Si64 cubic L=20.52 bohr, Ecut 15 Ha, Gamma, 128 bands",
                   "mtxd": mtxd, "mesh": [info["n1"], info["n2"], info["n3"]],
                   "z_columns": info["ncol"], "nanl": nanl, "bands": nb},
        "eig_err_ha": eig_err,
        "eig_err_refined_ha": eig_err_ref,
        "eig_tol_ha": 1e-6,
        "eig_ok": eig_err <= 1e-6,
        "hpsi_rel_err_fp32": herr,
        "fp64": {"bands_per_s": round(res["fp64"]["bands_per_s"], 1),
                 "ms_per_step": round(res["fp64"]["ms"], 4)},
        "fp32_over_fp64": round(res["fp64"]["ms"] / res["fp32"]["ms"], 3),
        "roofline": roofline_entry(res["fp32"]["dominant"], work["fp32"], res["fp32"]["dom_ms"],
                                   nb, "fp32"),
        "roofline_fp64": roofline_entry(res["fp64"]["dominant"], work["fp64"],
                                        res["fp64"]["dom_ms"], nb, "fp64"),
        "call_roofline": {p: call_roofline(info, nanl, nb, p, res[p]["ms"]) for p in ("fp32", "fp64")},
        "local_roofline": local,
        "tuned": tuned,
        "kernel_ms": kernels,
        "kernel_ms_alone": {p: {k: round(v, 5) for k, v in alone[p].items()} for p in alone},
        "kernel_roofline": {p: [roofline_entry(k, work[p], cal[p][k], nb, p)
                                for k in work[p] if k in cal[p]] for p in cal},
    }
    # measured HBM traffic of each kernel in its tuned variant, where recorded
    fams = {"k_plane": "plane", "k_proj": "proj", "k_back": "back", "k_col_bwd": "col", "k_col_fwd": "col"}
    for p in out["kernel_roofline"]:
        for e in out["kernel_roofline"][p]:
            if e["kernel"] in fams:
                e["traffic"], _ = measured_traffic(p, e["kernel"], tuned[p][fams[e["kernel"]]], info["n1"])
    if args.calib_copy:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        src.fill_(1.0)
        dst = src.clone()  # 512 MiB read + 512 MiB written
        torch.cuda.synchronize()
        del src, dst
    if args.headline_only:
        args.no_kpoints = args.no_ditsp = args.no_cpu = True
    if not args.no_kpoints:
        log("bench: 16 k-points, eager loop vs one graph launch")
        out["kpoints_graph"] = bench_kpoints(torch, hk, synth, np, max(3, args.steps // 6))
    if not args.headline_only:
        log("bench: two-component bands and the density")
        out["spinor_density"] = bench_spin_density(torch, hk, synth, np, d, max(5, args.steps // 3))
        log("bench: atomic forces (Si/Ge 64 atoms, 128 bands)")
        out["forces"] = bench_forces(torch, hk, synth, np, 5)
        log("bench: KB projectors built on the device (64-atom Si)")
        out["projectors_kb"] = bench_projectors(torch, hk, synth, np, d, 20,
                                                cpu_seconds=None if args.no_cpu else 3.0)
    if not args.no_ditsp:
        log("bench: iterative diagonalisation on the device (128 bands)")
        out["ditsp"] = bench_ditsp(torch, hk, plan, d, nb, np)
    if not (args.no_si216 or args.headline_only):
        log("bench: 216-atom cell (108^3, 512 bands)")
        out["si216"] = bench_cfg2(torch, hk, synth, np, max(3, args.steps // 6))
    for key, prec in (("roofline", "fp32"), ("roofline_fp64", "fp64")):
        # the timed loop runs the projection beside the local kernels, so the
        # dominant kernel's in-loop duration depends on which of them it shares
        # the GPU with (the tuned projection kernel and stream schedule); the
        # same kernel alone on one stream, from its own dispatch events:
        kern = out[key]["kernel"]
        if alone[prec].get(kern):
            a = roofline_entry(kern, work[prec], alone[prec][kern], nb, prec)
            out[key]["avg_ms_alone"] = a["avg_ms"]
            out[key]["frac_alone"] = a["frac"]
    for key, prec in (("roofline", "fp32"), ("roofline_fp64", "fp64")):
        kern = out[key]["kernel"]
        fam = {"k_plane": "plane", "k_proj": "proj", "k_back": "back"}.get(kern)
        if fam is not None and out[key].get("traffic") is None:
            tb, src = measured_traffic(prec, kern, tuned[prec][fam], info["n1"])
            out[key]["traffic"] = tb
            out[key]["traffic_source"] = src
    # the plane stage's real bound: VALU instruction issue (cost per kind measured here)
    try:
        rates = issue_rates()
    except (OSError, RuntimeError) as exc:
        log(f"bench: issue-rate probe unavailable ({exc})")
        rates = None
    out["issue_rates"] = rates
    for key, prec in (("roofline", "fp32"), ("roofline_fp64", "fp64")):
        iss = issue_bound(prec, out[key]["kernel"], tuned[prec]["plane"], info["n1"],
                          out[key]["avg_ms"], rates)
        if iss is not None:
            # the same kernel alone on one stream (no projector GEMMs beside it)
            alone = local[prec]["kernels_us"]["k_plane"]
            iss["measured_alone_us"] = alone
            iss["frac_alone"] = round(iss["floor_us"] / alone, 4)
        out[key]["issue"] = iss
    if not args.headline_only:
        # the Fortran-facing drop-in: host arrays in and out (PCIe included)
        hk.set_precision(hk.PREC_FP64)
        hpsi_h = np.zeros_like(d["psi"], order="F")
        hk.hk_psi_c16_dict(d, hpsi=hpsi_h)
        t0 = time.perf_counter()
        ncall = 3
        for _ in range(ncall):
            hk.hk_psi_c16_dict(d, hpsi=hpsi_h)
        dt = (time.perf_counter() - t0) / ncall
        out["dropin_host_fp64"] = {"bands_per_s": round(nb / dt, 1), "ms_per_call": round(dt * 1e3, 2),
                                   "note": "psi/H psi/V/projectors copied host<->device every call"}
    if not args.no_cpu:
        log(f"bench: CPU baseline (~{args.cpu_seconds:.0f} s)")
        out["cpu_baseline"] = cpu_baseline(d, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    plan.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
