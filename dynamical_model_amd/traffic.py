"""Measured HBM traffic per kernel kind, from two rocprofv3 counter runs.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d DIR_F -o run -- python3 bench.py ARGS
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d DIR_W -o run -- python3 bench.py ARGS
    python3 -m dynamical_model_amd.traffic DIR_F DIR_W JX IY KZ "ARGS" > profiles/traffic_JXxIYxKZ_fp32.json

FETCH_SIZE and WRITE_SIZE are the L2's reads from and writes to memory in KiB
per dispatch.  Both are calibrated on the run's own device-to-device copies of
one FP32 field allocation (the corrector's phase-start snapshots of u and v,
pitch x (iy + 4) x kz x 4 bytes each): the factor is the copied bytes over
the counter's reading.  Only dispatches of the FP32 run are used (those after
its device case build); the kinds are counted over its last step (from its
last prologue on), so that with two or more steps they are the steady state
of the learned loop (the first learned step checks its corrector, which
takes the phase-start copies the calibration uses and reads them back in
place of the increments the substep kernels write otherwise), grouped into
the kernel kinds that bench.py times;
one ``lac_mlp`` launch is the layer kernels of one inference batch (layer 1,
then layers 2 and 3 in one kernel or in two).
bench.py reads the result as ``roofline.traffic``.
"""
from __future__ import annotations

import csv
import glob
import json
import os
import sys
from collections import defaultdict

JOFF = 31

# kernel-name fragment -> kind (first match wins)
KINDS = [
    ("k_prologue<", "prologue"), ("k_halo<", "halo"), ("k_sound_a<", "sound_a"),
    ("k_sound_bs<", "sound_b"), ("k_sound_b<", "sound_b"), ("k_sound_c<", "sound_c"),
    ("k_adv_in<", "adv_in"), ("k_wafv<", "waf_v"), ("k_wafh<", "waf_h"), ("k_adv_out<", "adv_out"),
    ("k_epilogue<", "epilogue"), ("k_lac_features<", "lac_gather"), ("k_lac_apply<", "lac_apply"),
    ("k_mlp_fused", "mlp_fused"), ("k_mlp_tail", "mlp_rest"), ("k_mlp_dense", "mlp_rest"),
    ("__amd_rocclr_copyBuffer", "copy"),
]
FP32_START = "k_case_build<float>"
STEP_START = "k_prologue<float>"


def kind_of(name: str) -> str | None:
    for frag, kind in KINDS:
        if frag in name:
            return kind
    return None


def read_counter(d: str, counter: str) -> list[tuple[int, str, int, float]]:
    """(dispatch id, kernel name, grid size, value) of every dispatch, in dispatch order."""
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no *counter_collection.csv under {d}")
    rows = {}
    for fn in files:
        with open(fn, newline="") as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                did = int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
                val = float(r["Counter_Value"])
                if did in rows:
                    rows[did][3] += val
                else:
                    rows[did] = [did, r["Kernel_Name"], int(float(r.get("Grid_Size") or 0)), val]
    return [tuple(v) for _, v in sorted(rows.items())]


def per_kind(rows, field_bytes: float):
    """Mean KiB per dispatch by kind over the FP32 run, and the copy calibration."""
    start = next((n for n, (_, name, _, _) in enumerate(rows) if FP32_START in name), None)
    if start is None:
        raise SystemExit(f"no {FP32_START} dispatch: not a bench run with an FP32 variant")
    last = max((n for n, (_, name, _, _) in enumerate(rows) if n >= start and STEP_START in name), default=start)
    acc = defaultdict(list)
    copies = defaultdict(list)
    for n, (_, name, grid, val) in enumerate(rows[start:], start):
        k = kind_of(name)
        if k == "copy":
            copies[grid].append(val)
        elif k and n >= last:
            acc[k].append(val)
    # the snapshot copies (one launch size for all) move u, v (one FP32 field
    # each), w (kz + 1 levels) and pi (FP64): u and v are the smallest
    grid, vals = max(copies.items(), key=lambda kv: len(kv[1]))
    lo = min(vals)
    vals = [v for v in vals if v <= 1.01 * lo]
    factor = field_bytes / (sum(vals) / len(vals) * 1024.0)
    means = {k: sum(v) / len(v) * 1024.0 * factor for k, v in acc.items()}
    counts = {k: len(v) for k, v in acc.items()}
    return means, counts, factor, len(vals)


def main(argv):
    if len(argv) != 7:
        print(__doc__, file=sys.stderr)
        return 2
    dir_f, dir_w, jx, iy, kz, args = argv[1], argv[2], int(argv[3]), int(argv[4]), int(argv[5]), argv[6]
    pitch = (JOFF + jx + 3 + 31) // 32 * 32
    field_bytes = float(pitch * (iy + 4) * kz * 4)
    fm, fc, ff, nf = per_kind(read_counter(dir_f, "FETCH_SIZE"), field_bytes)
    wm, wc, wf, nw = per_kind(read_counter(dir_w, "WRITE_SIZE"), field_bytes)
    kinds = {}
    for k in fm:
        if k in ("mlp_fused", "mlp_rest") or k not in wm:
            continue
        kinds[k] = {"fetch_bytes": fm[k], "write_bytes": wm[k], "bytes": fm[k] + wm[k], "dispatches": fc[k]}
    if "mlp_fused" in fm and "mlp_rest" in fm:
        per = fc["mlp_rest"] / fc["mlp_fused"]  # launches after layer 1 per batch
        fb = fm["mlp_fused"] + per * fm["mlp_rest"]
        wb = wm["mlp_fused"] + per * wm["mlp_rest"]
        kinds["lac_mlp"] = {"fetch_bytes": fb, "write_bytes": wb, "bytes": fb + wb, "dispatches": fc["mlp_fused"]}
    toks = args.split()
    config = toks[toks.index("--config") + 1] if "--config" in toks else "4"
    out = {
        "workload": f"configuration {config}, {jx}x{iy}x{kz}, FP32 state (FP64 Exner function, theta and "
                    "temperature): kernels of the FP32 and FP32+learned steps",
        "commands": [f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d prof -o run -- python3 bench.py {args}",
                     f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d prof -o run -- python3 bench.py {args}",
                     f"python3 -m dynamical_model_amd.traffic <fetch dir> <write dir> {jx} {iy} {kz} '{args}'"],
        "units": "bytes per launch as bench.py times launches (lac_mlp: the layer kernels of one batch)",
        "calibration": {"copy_bytes": field_bytes, "fetch_factor": ff, "write_factor": wf,
                        "copies": min(nf, nw)},
        "kinds": kinds,
    }
    json.dump(out, sys.stdout, indent=1)
    print()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
