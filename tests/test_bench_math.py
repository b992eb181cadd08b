"""The arithmetic of bench.py's roofline fields, on the 64-atom workload's
sizes (mtxd 23,847, 997 z columns, 72^3 mesh, 256 projectors, 128 bands):

* the whole-call roofline is the local stage at 8 TB/s plus the two projector
  GEMMs at the matrix-core peak: 133.5 + 159.1 = 292.6 us FP64 and 66.7 + 79.5
  = 146.2 us FP32, figures worked out by hand from the per-band bytes and flops;
* the local-stage fraction uses the same bytes over the sum of the three local
  kernels;
* the issue floor prices every instruction class at its measured cost, DPP
  exchanges at the DPP rate wherever the hardware counts them;
* the DPP counts stored with the instruction mix follow from the kernels'
  structure (wave-level line transforms x stages x values x words)."""

import json
import math
import os

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INFO = {"mtxd": 23847, "ncol": 997, "n1": 72, "n2": 72, "n3": 72}
NANL, NB = 256, 128


@pytest.mark.parametrize("prec,loc,nl", [("fp64", 133.5, 159.1), ("fp32", 66.7, 79.5)])
def test_call_roofline_matches_the_hand_figures(prec, loc, nl):
    r = bench.call_roofline(INFO, NANL, NB, prec, 0.5)
    assert abs(r["t_local_us"] - loc) < 0.1 and abs(r["t_nonlocal_us"] - nl) < 0.1
    assert abs(r["t_roof_us"] - (loc + nl)) < 0.2
    assert abs(r["frac"] - (loc + nl) / 500.0) < 1e-3


def test_local_roofline():
    per = {"k_col_bwd": 0.050, "k_plane": 0.270, "k_col_fwd": 0.060}
    r = bench.local_roofline(INFO, NB, "fp64", per)
    assert abs(r["t_roof_us"] - 133.5) < 0.1
    assert r["t_measured_us"] == 380.0
    assert abs(r["frac"] - 133.5 / 380.0) < 1e-3


RATES = {"v_fma_f64": 4.0, "v_add_f64": 4.2, "v_fma_f32": 2.0, "v_add_u32": 2.5, "v_mov_b32_dpp": 4.0,
         "v_add_f32": 2.1, "shader_clock_ghz": 2.0, "simds": 1000}


def test_issue_bound_prices_each_class(monkeypatch, tmp_path):
    prof = tmp_path / "profiles"
    prof.mkdir()
    entry = {"kernel": "k", "valu": 100, "f64": 40, "f32": 30, "int32": 10, "other": 20, "dpp": 15,
             "dpp_in": "other"}
    (prof / "roofline_traffic.json").write_text(json.dumps({"valu": {"fp64/plane1/72": entry}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    # 40 x 4.1 + 15 x 4.0 + 30 x 2.0 + 10 x 2.5 + (20 - 15) x 2.1 = 319.5 cycles over 1000 SIMDs
    cyc = 40 * 4.1 + 15 * 4.0 + 30 * 2.0 + 10 * 2.5 + 5 * 2.1
    r = bench.issue_bound("fp64", "k_plane", 1, 72, 1e-3, RATES)
    assert abs(r["floor_us"] - round(cyc / 1000 / 2.0e3, 1)) < 0.05
    # the same exchanges counted among the FP32 adds instead
    entry.update({"f32": 45, "other": 5, "dpp_in": "f32"})
    (prof / "roofline_traffic.json").write_text(json.dumps({"valu": {"fp64/plane1/72": entry}}))
    assert bench.issue_bound("fp64", "k_plane", 1, 72, 1e-3, RATES)["floor_us"] == r["floor_us"]
    assert bench.issue_bound("fp64", "k_proj", 1, 72, 1e-3, RATES) is None
    assert bench.issue_bound("fp64", "k_plane", 1, 72, 1e-3, None) is None


def _wave_transforms(ntask, g, threads):
    """Wave-level line transforms of one phase: rounds of threads / g tasks,
    each round issued by the waves its active lanes occupy."""
    per_round, waves, done = threads // g, 0, 0
    while done < ntask:
        act = min(per_round, ntask - done)
        waves += math.ceil(act * g / 64)
        done += act
    return waves


@pytest.mark.parametrize("key,shape,words", [
    ("fp32/plane1/72", (8, 8, 1, 256), 2), ("fp32/plane2/72", (8, 8, 2, 256), 2),
    ("fp32/plane3/72", (4, 4, 1, 256), 2), ("fp64/plane1/72", (8, 8, 1, 256), 4),
    ("fp64/plane2/72", (8, 8, 1, 320), 4), ("fp64/plane3/72", (4, 8, 1, 256), 4)])
def test_stored_dpp_counts_follow_the_kernel_structure(key, shape, words):
    gx, gy, planes, threads = shape
    n, nx, n3 = 72, 35, 72
    per = lambda g: int(math.log2(g)) * (n // g) * words
    wy = 2 * _wave_transforms(planes * nx, gy, threads)   # inverse and forward y
    wx = 2 * _wave_transforms(planes * n, gx, threads)    # inverse and forward x of each row
    want = n3 * math.ceil(NB / planes) * (wy * per(gy) + wx * per(gx))
    with open(os.path.join(ROOT, "profiles", "roofline_traffic.json")) as f:
        stored = json.load(f)["valu"][key]
    assert stored["dpp"] == want
    assert stored["dpp"] <= stored["valu"]


def test_projector_traffic_record_is_the_write_stream():
    """The recorded HBM traffic of k_proj_kb on the 64-atom cell (rocprofv3
    FETCH_SIZE / WRITE_SIZE scaled by the copy calibration) is the projectors
    written once plus little else: within 10% of 16 mtxd nanl + 24 mtxd."""
    with open(os.path.join(ROOT, "profiles", "roofline_traffic.json")) as f:
        rec = json.load(f)
    e = rec["kernels"]["fp64/proj_kb/23847x256"]
    algo = 16 * INFO["mtxd"] * NANL + 24 * INFO["mtxd"]
    assert algo <= e["bytes"] <= 1.1 * algo
    assert abs(e["bytes"] - (e["fetch_kib"] * rec["fetch_scale"] + e["write_kib"] * rec["write_scale"]) * 1024) < 2048


@pytest.mark.parametrize("item,msg", [("fp32.prj=2", "KIND one of"), ("fp16.proj=2", "PREC is fp32 or fp64"),
                                      ("fp32.proj", "expected PREC.KIND=V"), ("fp32.proj=x", "expected PREC.KIND=V")])
def test_set_override_is_checked_before_any_gpu_work(item, msg):
    """bench.py --set rejects a malformed override at argument parsing (exit
    status 2), before it imports the library or touches a GPU."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--set", item],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and msg in r.stderr, r.stderr[-400:]
