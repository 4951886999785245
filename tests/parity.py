"""Shared helpers of the parity tests: run the oracle and the GPU on the same
input and compare fields over the interior."""
from __future__ import annotations

import numpy as np

import dynamical_model_amd as dm
import oracle_lib

PROGNOSTIC = ["u", "v", "w", "pai", "tetav", "qv"]
DIAGNOSTIC = ["t", "p", "rho", "tvirt", "qsat", "ps"]


def small_config2(jx: int, iy: int, kz: int, **kw):
    """Configuration-2 physics (20 km top, jet, RH 60 %, random hills) with the
    terrain scaled to a small test grid so the hills stay well below the top."""
    over = dict(jx=jx, iy=iy, kz=kz, nhills=4, hill_hmin=100.0, hill_hmax=600.0,
                hill_hwmin=2.0, hill_hwmax=max(2.5, min(jx, iy) / 6.0))
    over.update(kw)
    return dm.case_config(2, **over)


def interior(name: str, a: np.ndarray) -> np.ndarray:
    """Interior points a field is defined on after a step."""
    b = dm.interior(a)
    if name == "w":  # levels 1..kz+1 are all state
        return b
    return b


def compare(a: np.ndarray, b: np.ndarray) -> tuple[float, float]:
    """(relative L2, max-abs) of a against reference b."""
    d = a - b
    nb = np.linalg.norm(b)
    rel = np.linalg.norm(d) / nb if nb > 0 else np.linalg.norm(d)
    return float(rel), float(np.max(np.abs(d))) if d.size else 0.0


def run_oracle(hs, params, nsteps: int) -> oracle_lib.Oracle:
    o = oracle_lib.oracle_from_host(hs, params)
    for n in range(1, nsteps + 1):
        o.moloch(n, False)
    return o


def field_errors(gpu: dict, o: oracle_lib.Oracle, names) -> dict:
    out = {}
    for n in names:
        out[n] = compare(interior(n, gpu[n]), interior(n, o.field(n)))
    return out
