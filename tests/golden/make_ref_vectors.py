"""Write tests/golden/ref_vectors.npz from the reference modules.

Runs oracle/_ref/ref_driver (built by `make -C oracle ref` from the
reference's mod_constants, mod_zita and mod_atmosphere sources; only where
the reference tree is mounted) for a few grids and stores what it prints:
the physical constants, the zita levels and gzita, the geometry (zeta, fmz,
fmzf, hx, hy) and the initial state of fill_atmosphere (t, qv, pai, ps).

    python tests/golden/make_ref_vectors.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
OUT = os.path.join(HERE, "ref_vectors.npz")

CONST_NAMES = ["egrav", "rgas", "rwat", "cpd", "cvd", "rdrcv", "cpovr", "rovcp", "p00", "ep1", "ep2",
               "wlhv", "tzero", "lrate", "eomeg2", "degrad", "mathpi", "regrav", "earthrad", "raddeg"]

# (name, jx, iy, kz, ztop, zh, a0, hill height m, hill half-width cells)
CASES = [
    ("hill", 12, 10, 8, 36000.0, 12000.0, 0.0, 500.0, 3.0),
    ("stretched", 16, 12, 10, 20000.0, 20000.0 / 3.0, 0.0, 1500.0, 4.0),
    ("flat", 9, 7, 6, 36000.0, 12000.0, 0.0, 0.0, 0.0),
    ("a0", 10, 8, 6, 36000.0, 12000.0, 0.3, 800.0, 2.0),
]


def read_driver_output(path: str) -> dict:
    raw = open(path, "rb").read()
    jx, iy, kz = (int(x) for x in np.frombuffer(raw[:12], dtype=np.int32))
    d = np.frombuffer(raw[12:], dtype=np.float64)
    pos = 0

    def take(n, shape=None):
        nonlocal pos
        a = d[pos:pos + n].copy()
        pos += n
        return a.reshape(shape) if shape else a

    out = {"dims": np.array([jx, iy, kz])}
    out["constants"] = take(len(CONST_NAMES))
    out["zita"] = take(kz + 1)
    out["zitah"] = take(kz)
    out["gzita_f"] = take(kz + 1)
    out["gzita_h"] = take(kz)
    out["ht"] = take(jx * iy, (iy, jx))
    out["hx"] = take((jx + 1) * iy, (iy, jx + 1))
    out["hy"] = take(jx * (iy + 1), (iy + 1, jx))
    out["zeta"] = take(jx * iy * kz, (kz, iy, jx))
    out["fmz"] = take(jx * iy * kz, (kz, iy, jx))
    out["fmzf"] = take(jx * iy * (kz + 1), (kz + 1, iy, jx))
    out["t"] = take(jx * iy * kz, (kz, iy, jx))
    out["qv"] = take(jx * iy * kz, (kz, iy, jx))
    out["pai"] = take(jx * iy * kz, (kz, iy, jx))
    out["ps"] = take(jx * iy, (iy, jx))
    assert pos == d.size
    return out


def main() -> int:
    if not os.path.exists(DRIVER):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    vec = {"constant_names": np.array(CONST_NAMES)}
    with tempfile.TemporaryDirectory() as tmp:
        for name, jx, iy, kz, ztop, zh, a0, hh, hw in CASES:
            path = os.path.join(tmp, name + ".bin")
            subprocess.run([DRIVER, str(jx), str(iy), str(kz), repr(ztop), repr(zh), repr(a0),
                            repr(hh), repr(hw), path], check=True)
            for k, v in read_driver_output(path).items():
                vec[f"{name}/{k}"] = v
            vec[f"{name}/args"] = np.array([jx, iy, kz, ztop, zh, a0, hh, hw])
    np.savez_compressed(OUT, **vec)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
