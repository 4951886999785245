#!/usr/bin/env python3
"""Throughput of the MOLOCH timestep on one MI355X.

Metric: column-steps/s per GPU -- one (j,i) column advanced by one
moloch(istep, lprint) call (dt = 30 s: 3 advection cycles x 5 acoustic
substeps, nompi/mod_moloch.F90:272-421) -- for the plain FP64 core, the FP32
core (FP64 Exner function, theta, temperature, hydrostatic residuals and
column solve) and the FP32 core with the learned acoustic corrector (one long
acoustic substep plus a BF16 MFMA network increment, then the last ordinary
substep; the network is trained on the GPU from ordinary steps of the same
run before timing, and a check every 50 steps against the ordinary substeps
with a 1e-3 relative-error budget sends it back to the ordinary loop, with
retraining, when it misses), with each variant's relative L2 and max-abs
error against the FP64 core after the same steps
(dynamical_model_amd.errgrowth follows them over long runs).  `value` is the
fastest of the two reduced-precision variants whose every field stays within
the stated 1e-3 relative L2 error.

Workload (default): configuration 4, 2048 x 2048 x 60, synthetic initial state
(standard atmosphere, 20 seeded Gaussian hills, jet, RH 60 %, theta noise)
built by the library on the device.  Every field is resident in HBM before the timed
region; the timed region covers exactly --steps calls of the step.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "column-steps/s per GPU: plain FP64 vs FP32+learned acoustic at stated L2 error"
PROGNOSTIC = ["u", "v", "w", "pai", "tetav", "qv"]
ERROR_BOUND = 1e-3  # the stated relative L2 error against the FP64 core


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_variant(dm, hs, case, precision: int, steps: int, warmup: int, lead: int, learned: dict | None = None):
    """lead ordinary steps (the corrector's training run when learned), warmup
    untimed steps, then `steps` timed steps."""
    params = dm.case_params(case, precision=precision)
    m = dm.Moloch()
    m.allocate_moloch(case.jx, case.iy, case.kz, params)
    m.load_case_device(case)  # geometry and initial state built in HBM
    extra = {}
    if learned is not None:
        from dynamical_model_amd.lac import Corrector

        corr = Corrector(m, **learned["config"])
        t0 = time.perf_counter()
        corr.collect(lead)
        m.sync()
        t1 = time.perf_counter()
        losses = [corr.train(it, learned["batch"], lr) for it, lr in learned["schedule"]]
        t2 = time.perf_counter()
        corr.mode("on")
        extra = {"collect_s": t1 - t0, "train_s": t2 - t1, "losses": losses,
                 "samples": corr.stats()["samples"]}
        log(f"corrector: collected {extra['samples']} pairs in {t1 - t0:.1f} s, trained in {t2 - t1:.1f} s, "
            f"losses {' '.join(f'{x:.4f}' for x in losses)}")
    else:
        for n in range(1, lead + 1):
            m.moloch(n, False)
    for n in range(lead + 1, lead + warmup + 1):
        m.moloch(n, False)
    m.sync()
    m.profile(True)
    m.sync()
    t0 = time.perf_counter()
    first = lead + warmup + 1
    for n in range(first, first + steps):
        m.moloch(n, False)
    m.sync()
    t1 = time.perf_counter()
    prof = m.profile_report()
    m.profile(False)
    diag = m.moloch(first + steps, True)  # one more step with the lprint diagnostics
    m.sync()
    if learned is not None:
        extra["stats"] = corr.stats()
    finals = {f: dm.interior(m.get_field(f)).copy() for f in PROGNOSTIC}
    m.close()
    return (t1 - t0), prof, finals, diag, extra


MFMA_KINDS = {"lac_mlp"}  # work counted in flops; every other kind in bytes


def traffic_table(jx: int, iy: int, kz: int, precision: str) -> dict:
    """Measured HBM bytes per launch of each kernel kind (rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs, calibrated on device copies of known size), if a table for
    this workload is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", f"traffic_{jx}x{iy}x{kz}_{precision}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("kinds", {})
MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16


def roofline_of(prof: dict, name: str, traffic: dict) -> dict:
    ms, n, work = prof[name]
    measured = traffic.get(name, {}).get("bytes")
    if name in MFMA_KINDS:
        achieved = work / (ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": achieved, "peak": MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / MFMA_PEAK_TFLOPS, "traffic": measured, "kernel": name,
                "flops_per_launch": work / n, "avg_launch_ms": ms / n}
    achieved = work / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": measured, "kernel": name,
           "bytes_per_launch": work / n, "avg_launch_ms": ms / n}
    if measured:
        out["measured_gbs"] = measured / (ms / n * 1e-3) / 1e9
    return out


def step_bytes(kz: int, fp64: bool) -> float:
    """Algorithmic bytes of one column-step of the ordinary step: 470 field
    passes of kz levels (prologue 11, 15 acoustic substeps x 25, 3 advection
    cycles x 24, epilogue 12); 108 of them are passes over the Exner function,
    theta or the temperature (prologue 3, per substep 6, per advection cycle
    4, epilogue 3), which stay FP64 in the FP32 core."""
    s = 8 if fp64 else 4
    return (470 * s + 108 * (8 - s)) * kz


def dominant(prof: dict) -> tuple[str, float, float, int]:
    name = max(prof, key=lambda k: prof[k][0])
    ms, n, b = prof[name]
    return name, ms, b, n


def cpu_baseline(jx: int, iy: int, kz: int, steps: int, threaded: bool = False):
    """The CPU oracle (a C restatement of the reference step; serial, or the
    same source with loop-level OpenMP threads on all host cores, which
    computes the same bits) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import dynamical_model_amd as dm
    import oracle_lib

    case = dm.case_config(4, jx=jx, iy=iy, kz=kz, hill_hwmax=min(40.0, jx / 6.0))
    hs = dm.build_case(case)
    o = oracle_lib.oracle_from_host(hs, dm.case_params(case), threaded=threaded)
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        o.moloch(n, False)
    dt = time.perf_counter() - t0
    return jx * iy * steps / dt, dt


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50,
                    help="timed steps; 50 = one corrector check period, so the timed window carries its "
                         "steady-state share of checks (and of retraining after a failed check)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4, help="benchmark configuration (grid and physics)")
    ap.add_argument("--jx", type=int, default=0)
    ap.add_argument("--iy", type=int, default=0)
    ap.add_argument("--kz", type=int, default=0)
    ap.add_argument("--cpu-sample", default="128x128x10",
                    help="JXxIYxSTEPS of the serial CPU oracle sample, or 'none'")
    ap.add_argument("--cpu-sample-threaded", default="256x256x10",
                    help="JXxIYxSTEPS of the all-core CPU oracle sample, or 'none'")
    ap.add_argument("--lead", type=int, default=10,
                    help="ordinary steps before warm-up (the corrector's training run)")
    ap.add_argument("--train-iters", type=int, default=6000, help="Adam steps of the corrector")
    ap.add_argument("--train-batch", type=int, default=1024, help="rows per Adam step")
    ap.add_argument("--samples", type=int, default=8192, help="training columns sampled per acoustic phase")
    ap.add_argument("--chunk", type=int, default=65536, help="rows per corrector inference launch (multiple of 256)")
    ap.add_argument("--variants", default="fp64,fp32,learned")
    ap.add_argument("--check-tol", type=float, default=1.0,
                    help="corrector check: fall back when a field's error ratio exceeds this")
    ap.add_argument("--rel-tol", type=float, default=1e-3,
                    help="corrector check: fall back when a field's relative L2 error against the ordinary "
                         "substeps exceeds this (0: no error budget)")
    ap.add_argument("--retrain-steps", type=int, default=2,
                    help="after a failed check: ordinary steps collecting fresh pairs before retraining "
                         "(0: stay on the ordinary loop)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU halo exchange is a design, not an implementation)", file=sys.stderr)
        return 2

    import dynamical_model_amd as dm

    over = {k: getattr(args, k) for k in ("jx", "iy", "kz") if getattr(args, k)}
    case = dm.case_config(args.config, **over)
    jx, iy, kz = case.jx, case.iy, case.kz
    hs = None

    learned_cfg = {
        "config": dict(nhid=512, samples_per_phase=args.samples,
                       capacity=max(131072, 3 * args.lead * args.samples), chunk_rows=args.chunk, check_every=50,
                       check_tol=args.check_tol, seed=11, engine=1, rel_tol=args.rel_tol,
                       retrain_steps=args.retrain_steps, retrain_iters=max(1, args.train_iters // 4),
                       retrain_batch=args.train_batch, retrain_lr=3e-4, max_retrains=3),
        "batch": args.train_batch,
        "schedule": [(args.train_iters // 2, 1e-3), (args.train_iters // 4, 3e-4), (args.train_iters // 4, 1e-4)],
    }
    variants = {"fp64": (dm.MLC_FP64, None), "fp32": (dm.MLC_FP32, None), "learned": (dm.MLC_FP32, learned_cfg)}
    wanted = [v for v in args.variants.split(",") if v]
    if "fp64" not in wanted:
        wanted.insert(0, "fp64")  # the reference of the error figures
    results = {}
    finals = {}
    failed = {}
    for tag in wanted:
        prec, lcfg = variants[tag]
        t0 = time.perf_counter()
        try:
            secs, prof, fin, diag, extra = run_variant(dm, hs, case, prec, args.steps, args.warmup, args.lead, lcfg)
        except dm.MolochError as e:
            # a library error in the learned variant (its network and training
            # set need about 22 GB more than the core) leaves the two cores'
            # measurements standing; the cores themselves must run
            if tag != "learned":
                raise
            log(f"{tag}: failed: {e}")
            failed[tag] = str(e)
            continue
        cs = jx * iy * args.steps / secs
        name, ms, b, n = dominant(prof)
        results[tag] = {
            "value": cs, "ms_per_step": 1e3 * secs / args.steps, "dominant_kernel": name,
            "kernels": {k: {"ms_total": v[0], "launches": v[1],
                            "work_per_s": (v[2] / (v[0] * 1e-3)) if v[0] else 0.0}
                        for k, v in sorted(prof.items(), key=lambda kv: -kv[1][0])},
            "diag": {"ps_max": diag.ps_max, "ps_min": diag.ps_min, "u_max": diag.u_max, "u_min": diag.u_min,
                     "v_max": diag.v_max, "v_min": diag.v_min, "cfl_violation": diag.cfl_violation},
            "_prof": prof,
        }
        if extra:
            results[tag]["corrector"] = extra
        else:
            sb = step_bytes(kz, tag == "fp64") * jx * iy
            results[tag]["step_roofline"] = {"bytes_per_step": sb, "achieved_gbs": sb / (secs / args.steps) / 1e9,
                                             "frac": sb / (secs / args.steps) / 1e9 / HBM_PEAK_GBS}
        finals[tag] = fin
        log(f"{tag}: {cs:.4g} column-steps/s, {1e3 * secs / args.steps:.2f} ms/step, dominant {name} "
            f"({time.perf_counter() - t0:.1f} s incl. set-up)")

    errors = {tag: {f: rel_l2(finals[tag][f], finals["fp64"][f]) for f in PROGNOSTIC}
              for tag in results if tag != "fp64"}
    maxabs = {tag: {f: float(np.max(np.abs(finals[tag][f] - finals["fp64"][f]))) for f in PROGNOSTIC}
              for tag in results if tag != "fp64"}
    del finals

    # value: the fastest reduced-precision variant whose every prognostic
    # field is within the stated error (relative L2 <= ERROR_BOUND against
    # the FP64 core after the run); FP64 when none is
    within = {t: max(e.values()) <= ERROR_BOUND for t, e in errors.items()}
    candidates = [t for t in ("learned", "fp32") if t in results and within.get(t)]
    chosen_tag = max(candidates, key=lambda t: results[t]["value"]) if candidates else "fp64"
    chosen = results[chosen_tag]
    traffic = traffic_table(jx, iy, kz, "fp64" if chosen_tag == "fp64" else "fp32")
    roofline = roofline_of(chosen["_prof"], dominant(chosen["_prof"])[0], traffic)
    # the core's dominant stencil kernel (the column solve), and every stencil
    # kind with its algorithmic and measured rates
    stencil = roofline_of(chosen["_prof"], "sound_b", traffic)
    stencils = {k: roofline_of(chosen["_prof"], k, traffic) for k in chosen["_prof"]
                if k not in MFMA_KINDS and chosen["_prof"][k][2] > 0}
    lp = results["learned"]["_prof"] if "learned" in results else {}
    mlp = roofline_of(lp, "lac_mlp", traffic_table(jx, iy, kz, "fp32")) if "lac_mlp" in lp else None
    # the rate of plain device-to-device copies (the corrector's phase-start
    # snapshots, hipMemcpyAsync) in the same timed window: what streaming
    # reaches on this box, next to the 8 TB/s peak
    # (the kind is charged the interior; a copy moves whole padded planes of
    # pj = roundup(jx + 34, 32) by iy + 4 points)
    copy = next((r["_prof"]["lac_copy"] for r in results.values() if "lac_copy" in r["_prof"]), None)
    plane = ((jx + 34 + 31) // 32 * 32) * (iy + 4)
    copy_gbs = copy[2] * plane / (jx * iy) / (copy[0] * 1e-3) / 1e9 if copy and copy[0] > 0 else None
    if copy_gbs:
        for r in [roofline, stencil] + list(stencils.values()):
            if r.get("bound") == "hbm":
                r["frac_of_copy"] = r["achieved"] / copy_gbs

    cpu = None
    serial = None
    if args.cpu_sample != "none":
        sj, si, ss = (int(x) for x in args.cpu_sample.split("x"))
        v, dt = cpu_baseline(sj, si, kz, ss)
        serial = {"value": v, "unit": "column-steps/s", "cores": 1, "kind": "port",
                  "sample": f"CPU oracle (C restatement of nompi/mod_moloch.F90, FP64, 1 thread), configuration-"
                            f"{args.config} physics on {sj}x{si}x{kz} for {ss} steps ({dt:.1f} s)"}
        log(f"cpu baseline, 1 core: {v:.4g} column-steps/s ({dt:.1f} s)")
        cpu = serial
    if args.cpu_sample_threaded != "none":
        sj, si, ss = (int(x) for x in args.cpu_sample_threaded.split("x"))
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        v, dt = cpu_baseline(sj, si, kz, ss, threaded=True)
        cpu = {"value": v, "unit": "column-steps/s", "cores": threads, "kind": "port",
               "sample": f"CPU oracle (C restatement of nompi/mod_moloch.F90, FP64, loop-level OpenMP on "
                         f"{threads} threads, bit-identical to the serial build), configuration-{args.config} "
                         f"physics on {sj}x{si}x{kz} for {ss} steps ({dt:.1f} s)",
               "one_core": serial}
        log(f"cpu baseline, {threads} threads: {v:.4g} column-steps/s ({dt:.1f} s)")

    for r in results.values():
        r.pop("_prof")
    if "learned" in results:
        results["learned"]["corrector"]["stats"] = dict(results["learned"]["corrector"].get("stats", {}))
    line = {
        "metric": METRIC,
        "value": chosen["value"],
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" + ("+bf16" if chosen_tag == "learned" else ""),
        "data": "synthetic",
        "config": {"workload": f"configuration {args.config}: MOLOCH step on {jx}x{iy}x{kz} "
                               "(dt=30 s, 3 advection cycles x 5 acoustic substeps); "
                               + {"learned": "FP32 state, FP64 Exner/solve, learned acoustic corrector "
                                             "(long substep + BF16 MFMA MLP increment)",
                                  "fp32": "FP32 state, FP64 Exner function, hydrostatic residuals and solve",
                                  "fp64": "FP64"}[chosen_tag],
                   "jx": jx, "iy": iy, "kz": kz, "variant": chosen_tag,
                   "lead_steps": args.lead},
        "fp64": {"value": results["fp64"]["value"], "ms_per_step": results["fp64"]["ms_per_step"], "dtype": "f64"},
        "speedup_vs_fp64": {t: r["value"] / results["fp64"]["value"] for t, r in results.items()},
        "rel_l2_vs_fp64": errors,
        "maxabs_vs_fp64": maxabs,
        "rel_l2_after_steps": args.lead + args.warmup + args.steps + 1,
        "value_rule": f"fastest of fp32, learned with every field's relative L2 vs FP64 <= {ERROR_BOUND:g}",
        "within_error_bound": within,
        "roofline": roofline,
        "stencil_roofline": stencil,
        "stencil_kernels": {k: {x: r[x] for x in ("achieved", "frac", "frac_of_copy", "traffic", "avg_launch_ms",
                                                  "measured_gbs")
                                if x in r} for k, r in stencils.items()},
        "mlp_roofline": mlp,
        "copy_gbs": copy_gbs,
        "cpu_baseline": cpu,
        "variants": results,
    }
    if failed:
        line["failed_variants"] = failed
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
