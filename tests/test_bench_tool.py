"""The benchmark's own arithmetic on the CPU: the roofline objects of the
bench line, the algorithmic bytes of a column-step, the committed traffic
table it reads, the refusal of more than one GPU and the CPU-oracle sample."""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_roofline_objects():
    b = _bench()
    # (total ms, launches, work): two launches of 2.5e11 flops in 0.5 ms
    prof = {"lac_mlp": (0.5, 2, 5.0e11), "sound_b": (10.0, 2, 3.2e10)}
    r = b.roofline_of(prof, "lac_mlp", {"lac_mlp": {"bytes": 3.0e8}})
    assert r["bound"] == "mfma" and r["unit"] == "TFLOP/s"
    assert r["achieved"] == pytest.approx(1000.0) and r["frac"] == pytest.approx(0.4)
    assert r["avg_launch_ms"] == pytest.approx(0.25) and r["traffic"] == 3.0e8
    s = b.roofline_of(prof, "sound_b", {"sound_b": {"bytes": 2.0e10}})
    assert s["bound"] == "hbm" and s["unit"] == "GB/s"
    assert s["achieved"] == pytest.approx(3200.0) and s["frac"] == pytest.approx(0.4)
    assert s["bytes_per_launch"] == pytest.approx(1.6e10)
    assert s["measured_gbs"] == pytest.approx(4000.0)  # 2e10 bytes per 5 ms launch
    assert b.roofline_of(prof, "sound_b", {})["traffic"] is None


def test_step_bytes():
    """470 field passes of kz levels; in the FP32 core the 53 Exner passes stay
    FP64: 225.6 kB and 125.5 kB per column-step at kz = 60."""
    b = _bench()
    assert b.step_bytes(60, True) == 470 * 8 * 60 == 225600
    assert b.step_bytes(60, False) == (470 * 4 + 108 * 4) * 60 == 138720


def test_committed_traffic_table():
    b = _bench()
    t = b.traffic_table(2048, 2048, 60, "fp32")
    for kind in ("sound_a", "sound_b", "sound_c", "waf_v", "waf_h", "lac_mlp", "lac_gather", "lac_apply"):
        assert t[kind]["bytes"] > 0, kind
    assert b.traffic_table(7, 7, 7, "fp32") == {}


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 2
    assert "one GPU" in (out.stdout + out.stderr)


def test_cpu_baseline_sample():
    b = _bench()
    rate, dt = b.cpu_baseline(16, 16, 10, 1)
    assert rate > 0 and dt > 0
    rate_t, _ = b.cpu_baseline(16, 16, 10, 1, threaded=True)
    assert rate_t > 0


def test_failed_learned_variant_leaves_the_line(monkeypatch, capsys):
    """A library error in the learned variant is recorded in the line, and
    the FP64 and FP32 measurements (value from the FP32 core) still stand; an
    error in a core variant still ends the run.  The variants are stand-ins
    (no GPU): fixed timings, kernel profiles and fields."""
    import json
    import types

    import numpy as np

    import dynamical_model_amd as dm

    b = _bench()
    fields = {f: np.full((3, 8, 8), 1.0 + i) for i, f in enumerate(b.PROGNOSTIC)}

    def fake(dm_, hs, case, precision, steps, warmup, lead, learned=None):
        if learned is not None:
            raise dm.MolochError("mlc_lac_create", 4, "out of device memory (stand-in)")
        secs = 0.2 if precision == dm.MLC_FP32 else 0.4
        prof = {"sound_b": (secs * 400.0, 15 * steps, 1.0e10 * 15 * steps), "waf_h": (secs * 100.0, 3 * steps, 1e9)}
        diag = types.SimpleNamespace(ps_max=1.0, ps_min=0.5, u_max=1.0, u_min=0.0, v_max=1.0, v_min=0.0,
                                     cfl_violation=0)
        return secs, prof, {f: a.copy() for f, a in fields.items()}, diag, {}

    monkeypatch.setattr(b, "run_variant", fake)
    monkeypatch.setattr(sys, "argv", ["bench.py", "--steps", "2", "--jx", "16", "--iy", "16", "--kz", "4",
                                      "--cpu-sample", "none", "--cpu-sample-threaded", "none"])
    assert b.main() == 0
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert "learned" in line["failed_variants"] and "learned" not in line["variants"]
    assert line["config"]["variant"] == "fp32"
    assert line["value"] == pytest.approx(16 * 16 * 2 / 0.2)
    assert line["roofline"]["kernel"] == "sound_b" and line["mlp_roofline"] is None

    def fake_core_error(*args, **kwargs):
        raise dm.MolochError("mlc_create", 4, "stand-in")

    monkeypatch.setattr(b, "run_variant", fake_core_error)
    with pytest.raises(dm.MolochError):
        b.main()
