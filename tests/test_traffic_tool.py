"""The measured-traffic reduction (dynamical_model_amd.traffic) on synthetic
rocprofv3 counter files: kind grouping, the FP32-run window, the copy
calibration and the three-kernel network launch."""
from __future__ import annotations

import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        for n, (name, val) in enumerate(rows, 1):
            w.writerow([n, name, 131072 if "copyBuffer" in name else 256, counter, val])


def test_traffic_reduction(tmp_path):
    pitch = (31 + 64 + 3 + 31) // 32 * 32
    field = pitch * (64 + 4) * 8 * 4  # one FP32 field allocation at jx = iy = 64, kz = 8
    kib = field / 1024.0
    # FP64 run first (ignored), then the FP32 run; copies: u, v (one field),
    # w (a little more), pi (twice); fetch reads half, write all of a copy
    common = [("void mlc::k_sound_bs<double, 8, 8, 32>(x)", 1e9)]
    fp32 = [("void mlc::k_case_build<float>(x)", 1.0),
            ("void mlc::k_sound_bs<float, 8, 8, 32>(x)", 100.0),
            ("void mlc::k_sound_bs<float, 8, 8, 32>(x)", 300.0),
            ("mlc::k_mlp_fused(x)", 10.0), ("mlc::k_mlp_dense(x)", 2.0), ("mlc::k_mlp_dense(x)", 4.0)]
    copies_f = [("__amd_rocclr_copyBuffer", kib / 2)] * 2 + [("__amd_rocclr_copyBuffer", kib)]
    copies_w = [("__amd_rocclr_copyBuffer", kib)] * 2 + [("__amd_rocclr_copyBuffer", 2 * kib)]
    _write(str(tmp_path / "f"), "FETCH_SIZE", common + fp32 + copies_f)
    _write(str(tmp_path / "w"), "WRITE_SIZE", common + fp32 + copies_w)
    out = subprocess.run([sys.executable, "-m", "dynamical_model_amd.traffic", str(tmp_path / "f"),
                          str(tmp_path / "w"), "64", "64", "8", "--x"], cwd=ROOT, capture_output=True,
                         text=True, check=True).stdout
    d = json.loads(out)
    cal = d["calibration"]
    assert cal["copy_bytes"] == field
    assert abs(cal["fetch_factor"] - 2.0) < 1e-12 and abs(cal["write_factor"] - 1.0) < 1e-12
    sb = d["kinds"]["sound_b"]
    assert sb["dispatches"] == 2
    assert abs(sb["fetch_bytes"] - 200.0 * 1024 * 2.0) < 1e-6  # mean 200 KiB, factor 2
    assert abs(sb["write_bytes"] - 200.0 * 1024) < 1e-6
    mlp = d["kinds"]["lac_mlp"]  # one fused + two dense launches per batch
    assert abs(mlp["write_bytes"] - (10.0 + 2 * 3.0) * 1024) < 1e-6


def test_traffic_counts_the_last_step(tmp_path):
    """Kinds are counted from the FP32 run's last prologue on; the copies of
    the whole FP32 run calibrate."""
    pitch = (31 + 64 + 3 + 31) // 32 * 32
    kib = pitch * (64 + 4) * 8 * 4 / 1024.0
    fp32 = [("void mlc::k_case_build<float>(x)", 1.0), ("void mlc::k_prologue<float>(x)", 5.0),
            ("void mlc::k_lac_features<float>(x)", 100.0)] + [("__amd_rocclr_copyBuffer", kib / 2)] * 2 + [
        ("void mlc::k_prologue<float>(x)", 7.0), ("void mlc::k_lac_features<float>(x)", 40.0)]
    _write(str(tmp_path / "f"), "FETCH_SIZE", fp32)
    _write(str(tmp_path / "w"), "WRITE_SIZE", [(n, v * 2 if "copyBuffer" in n else v) for n, v in fp32])
    out = subprocess.run([sys.executable, "-m", "dynamical_model_amd.traffic", str(tmp_path / "f"),
                          str(tmp_path / "w"), "64", "64", "8", "--x"], cwd=ROOT, capture_output=True,
                         text=True, check=True).stdout
    d = json.loads(out)
    assert abs(d["calibration"]["fetch_factor"] - 2.0) < 1e-12
    g = d["kinds"]["lac_gather"]
    assert g["dispatches"] == 1
    assert abs(g["fetch_bytes"] - 40.0 * 1024 * 2.0) < 1e-6
    assert abs(d["kinds"]["prologue"]["fetch_bytes"] - 7.0 * 1024 * 2.0) < 1e-6


def test_traffic_reduction_tail_kernel(tmp_path):
    """Layers 2 and 3 in one launch: one lac_mlp launch is layer 1 plus that kernel."""
    pitch = (31 + 64 + 3 + 31) // 32 * 32
    kib = pitch * (64 + 4) * 8 * 4 / 1024.0
    fp32 = [("void mlc::k_case_build<float>(x)", 1.0), ("mlc::k_mlp_fused(x)", 10.0), ("mlc::k_mlp_tail(x)", 4.0),
            ("mlc::k_mlp_fused(x)", 10.0), ("mlc::k_mlp_tail(x)", 6.0)]
    _write(str(tmp_path / "f"), "FETCH_SIZE", fp32 + [("__amd_rocclr_copyBuffer", kib / 2)] * 2)
    _write(str(tmp_path / "w"), "WRITE_SIZE", fp32 + [("__amd_rocclr_copyBuffer", kib)] * 2)
    out = subprocess.run([sys.executable, "-m", "dynamical_model_amd.traffic", str(tmp_path / "f"),
                          str(tmp_path / "w"), "64", "64", "8", "--x"], cwd=ROOT, capture_output=True,
                         text=True, check=True).stdout
    mlp = json.loads(out)["kinds"]["lac_mlp"]
    assert mlp["dispatches"] == 2
    assert abs(mlp["write_bytes"] - (10.0 + 5.0) * 1024) < 1e-6
