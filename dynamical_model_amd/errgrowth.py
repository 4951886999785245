"""Error growth of the FP32 core and of the FP32 core with the learned
acoustic corrector against the FP64 core, over a long run on one GPU.

    python -m dynamical_model_amd.errgrowth [--config 2] [--steps 2000] [--every 100] [--variants fp32,learned]
                                            [--max-seconds S]

All handles start from the same device-built case (configurations 0-4) and
take the same steps: ``--lead`` ordinary steps (for the learned variant:
its training-pair collection, bit-identical to ordinary steps), then
``--steps`` steps, the learned variant with its corrector on (checks every
50 steps, fallback to the ordinary loop when a check fails).  Every
``--every`` steps one JSON line gives, per prognostic field (u, v, w, pi,
theta, qv), the relative L2 and the max-abs difference from the FP64 core
over the interior, and the corrector's check statistics; with
``--max-seconds`` the run ends at the first report after that much wall
time (a last line says so).  The FP64 core is
itself bit-identical to the FP64 CPU oracle (the parity tests), so these are
the errors against the reference's FP64 algorithm.
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


def _errors(ref: dict, m: dm.Moloch) -> dict:
    out = {}
    for f in PROGNOSTIC:
        a = dm.interior(m.get_field(f))
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
    ap.add_argument("--lead", type=int, default=10)
    ap.add_argument("--variants", default="fp32,learned")
    ap.add_argument("--train-iters", type=int, default=6000)
    ap.add_argument("--samples", type=int, default=8192)
    ap.add_argument("--max-seconds", type=float, default=0.0, help="stop at the first report after this wall time")
    args = ap.parse_args(argv)

    over = {k: getattr(args, k) for k in ("jx", "iy", "kz") if getattr(args, k)}
    case = dm.case_config(args.config, **over)
    wanted = [v for v in args.variants.split(",") if v]

    def handle(prec):
        m = dm.Moloch()
        m.allocate_moloch(case.jx, case.iy, case.kz, dm.case_params(case, precision=prec))
        m.load_case_device(case)
        return m

    t0 = time.perf_counter()
    ref = handle(MLC_FP64)
    runs = {v: handle(MLC_FP32) for v in wanted}
    corr = None
    for n in range(1, args.lead + 1):
        ref.moloch(n, False)
        if "fp32" in runs:
            runs["fp32"].moloch(n, False)
    if "learned" in runs:
        from .lac import Corrector

        corr = Corrector(runs["learned"], nhid=512, samples_per_phase=args.samples,
                         capacity=max(131072, 3 * args.lead * args.samples), chunk_rows=65536, check_every=50,
                         check_tol=1.0, seed=11, engine=1)
        corr.collect(args.lead)
        it = args.train_iters
        for n_it, lr in ((it // 2, 1e-3), (it // 4, 3e-4), (it // 4, 1e-4)):
            corr.train(n_it, 1024, lr)
        corr.mode("on")
    print(json.dumps({"config": args.config, "grid": [case.jx, case.iy, case.kz], "lead": args.lead,
                      "variants": wanted, "setup_s": time.perf_counter() - t0}), flush=True)
    for n in range(args.lead + 1, args.lead + args.steps + 1):
        ref.moloch(n, False)
        for m in runs.values():
            m.moloch(n, False)
        if (n - args.lead) % args.every == 0 or n == args.lead + args.steps:
            ref.sync()
            base = {f: dm.interior(ref.get_field(f)).copy() for f in PROGNOSTIC}
            line = {"step": n, "elapsed_s": time.perf_counter() - t0}
            for v, m in runs.items():
                m.sync()
                line[v] = _errors(base, m)
            if corr is not None:
                st = corr.stats()
                line["corrector"] = {k: st[k] for k in ("learned_phases", "plain_phases", "checks", "fallback",
                                                         "worst_ratio") if k in st}
            print(json.dumps(line), flush=True)
            del base
            if args.max_seconds > 0 and time.perf_counter() - t0 > args.max_seconds and n < args.lead + args.steps:
                print(json.dumps({"stopped_at_step": n, "reason": "max-seconds", "elapsed_s": time.perf_counter() - t0}),
                      flush=True)
                break
    ref.close()
    for m in runs.values():
        m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
