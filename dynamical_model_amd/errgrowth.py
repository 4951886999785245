"""Error growth of the FP32 core and of the FP32 core with the learned
acoustic corrector against the FP64 core, over a long run on one GPU.

    python -m dynamical_model_amd.errgrowth [--config 2] [--steps 2000] [--every 100] [--variants fp32,learned]
                                            [--max-seconds S] [--sub N] [--ref-out F | --ref-in F]
                                            [--report-steps a,b,...]

All handles start from the same device-built case (configurations 0-4) and
take the same steps: ``--lead`` ordinary steps (for the learned variant:
its training-pair collection, bit-identical to ordinary steps), then
``--steps`` steps, the learned variant with its corrector on (a check every
50 steps against the ordinary substeps with the relative-error budget
``--rel-tol``; a failed check recollects pairs and retrains).  Every
``--every`` steps (or at ``--report-steps``) one JSON line gives, per
prognostic field (u, v, w, pi, theta, qv), the relative L2 and the max-abs
difference from the FP64 core over the interior, and the corrector's
statistics; with ``--max-seconds`` the run ends at the first report after
that much wall time (a last line says so).  The FP64 core is itself
bit-identical to the FP64 CPU oracle (the parity tests), so these are the
errors against the reference's FP64 algorithm.

When the FP64 run and the variants do not fit one GPU lease together, the
FP64 run writes its fields at the report steps with ``--ref-out`` and a
second run compares against them with ``--ref-in``.  ``--sub N`` then keeps
every N-th column in j and in i (all levels): the relative L2 error over
that regular sample estimates the one over the whole grid.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import core as dm
from ._lib import MLC_FP32, MLC_FP64

PROGNOSTIC = ["u", "v", "w", "pai", "tetav", "qv"]


def _sample(a: np.ndarray, sub: int) -> np.ndarray:
    b = dm.interior(a)
    return b[..., ::sub, ::sub] if sub > 1 else b


def _errors(ref: dict, m: dm.Moloch, sub: int) -> dict:
    out = {}
    for f in PROGNOSTIC:
        a = _sample(m.get_field(f), sub)
        b = ref[f]
        d = a - b
        nb = float(np.linalg.norm(b))
        out[f] = {"rel_l2": float(np.linalg.norm(d)) / nb if nb > 0 else float(np.linalg.norm(d)),
                  "max_abs": float(np.max(np.abs(d)))}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--jx", type=int, default=0)
    ap.add_argument("--iy", type=int, default=0)
    ap.add_argument("--kz", type=int, default=0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=100)
    ap.add_argument("--report-steps", default="", help="comma-separated steps after the lead to report at")
    ap.add_argument("--lead", type=int, default=10)
    ap.add_argument("--variants", default="fp32,learned", help="fp32, learned, or none (with --ref-out)")
    ap.add_argument("--train-iters", type=int, default=6000)
    ap.add_argument("--samples", type=int, default=8192)
    ap.add_argument("--rel-tol", type=float, default=1e-3)
    ap.add_argument("--retrain-steps", type=int, default=2)
    ap.add_argument("--max-seconds", type=float, default=0.0, help="stop at the first report after this wall time")
    ap.add_argument("--sub", type=int, default=1, help="compare every N-th column in j and i")
    ap.add_argument("--ref-out", default="", help="write the FP64 fields at the report steps (npz)")
    ap.add_argument("--ref-in", default="", help="compare against FP64 fields written by --ref-out")
    args = ap.parse_args(argv)

    over = {k: getattr(args, k) for k in ("jx", "iy", "kz") if getattr(args, k)}
    case = dm.case_config(args.config, **over)
    wanted = [v for v in args.variants.split(",") if v and v != "none"]
    report_at = {int(x) for x in args.report_steps.split(",") if x}
    ref_in = np.load(args.ref_in) if args.ref_in else None
    if ref_in is not None and int(ref_in["sub"]) != args.sub:
        raise SystemExit("--sub differs from the reference file's")

    def handle(prec):
        m = dm.Moloch()
        m.allocate_moloch(case.jx, case.iy, case.kz, dm.case_params(case, precision=prec))
        m.load_case_device(case)
        return m

    t0 = time.perf_counter()
    ref = handle(MLC_FP64) if ref_in is None else None
    runs = {v: handle(MLC_FP32) for v in wanted}
    corr = None
    for n in range(1, args.lead + 1):
        if ref is not None:
            ref.moloch(n, False)
        if "fp32" in runs:
            runs["fp32"].moloch(n, False)
    if "learned" in runs:
        from .lac import Corrector

        corr = Corrector(runs["learned"], nhid=512, samples_per_phase=args.samples,
                         capacity=max(131072, 3 * args.lead * args.samples), chunk_rows=65536, check_every=50,
                         check_tol=1.0, seed=11, engine=1, rel_tol=args.rel_tol, retrain_steps=args.retrain_steps,
                         retrain_iters=max(1, args.train_iters // 4), retrain_batch=1024, retrain_lr=3e-4,
                         max_retrains=3)
        corr.collect(args.lead)
        it = args.train_iters
        for n_it, lr in ((it // 2, 1e-3), (it // 4, 3e-4), (it // 4, 1e-4)):
            corr.train(n_it, 1024, lr)
        corr.mode("on")
    print(json.dumps({"config": args.config, "grid": [case.jx, case.iy, case.kz], "lead": args.lead,
                      "variants": wanted, "sub": args.sub, "reference": args.ref_in or "fp64 handle",
                      "setup_s": time.perf_counter() - t0}), flush=True)
    saved = {}
    for n in range(args.lead + 1, args.lead + args.steps + 1):
        if ref is not None:
            ref.moloch(n, False)
        for m in runs.values():
            m.moloch(n, False)
        k = n - args.lead
        if k % 50 == 0:  # a progress line at least every 50 steps
            print(json.dumps({"progress": n, "elapsed_s": round(time.perf_counter() - t0, 1)}), flush=True)
        due = (k in report_at) if report_at else (k % args.every == 0)
        if due or n == args.lead + args.steps:
            if ref is not None:
                ref.sync()
                base = {f: _sample(ref.get_field(f), args.sub).copy() for f in PROGNOSTIC}
                if args.ref_out:
                    for f, a in base.items():
                        saved[f"{n}/{f}"] = a
            else:
                if f"{n}/u" not in ref_in:
                    continue
                base = {f: ref_in[f"{n}/{f}"] for f in PROGNOSTIC}
            line = {"step": n, "elapsed_s": time.perf_counter() - t0}
            for v, m in runs.items():
                m.sync()
                line[v] = _errors(base, m, args.sub)
            if corr is not None:
                st = corr.stats()
                line["corrector"] = {k2: st[k2] for k2 in ("learned_phases", "plain_phases", "checks", "fallback",
                                                          "worst_ratio", "worst_rel", "retrains", "collect_phases",
                                                          "failed_checks") if k2 in st}
            print(json.dumps(line), flush=True)
            del base
            if args.max_seconds > 0 and time.perf_counter() - t0 > args.max_seconds and n < args.lead + args.steps:
                print(json.dumps({"stopped_at_step": n, "reason": "max-seconds", "elapsed_s": time.perf_counter() - t0}),
                      flush=True)
                break
    if args.ref_out:
        np.savez(args.ref_out, sub=np.int64(args.sub), **saved)
    if ref is not None:
        ref.close()
    for m in runs.values():
        m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
