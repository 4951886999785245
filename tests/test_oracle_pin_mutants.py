"""The physics pins catch plausible mistakes in the oracle.

Each case plants one mistake (a wrong constant, a dropped or extra factor,
a wrong level array, a transposed coefficient, a flipped sign) in a scratch
copy of oracle/moloch_oracle.c, builds it with gcc into a temporary
directory and runs tests/test_oracle_physics_pins.py against it in a child
process: the pins must fail.  The same run against the unmodified source
must pass.  Nothing is written into the repository.
"""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "moloch_oracle.c")
PINS = os.path.join(HERE, "test_oracle_physics_pins.py")

# (name, text in the source, replacement): every text occurs in the source
MUTANTS = [
    ("eq10 half-level decay", "(zuh + zvh) * S->gzitak[k - 1];", "(zuh + zvh) * S->gzitakh[k - 1];"),
    ("eq10 quarter dropped", "F3(s, j, i, k) = -0.25 * (zuh + zvh)", "F3(s, j, i, k) = -0.5 * (zuh + zvh)"),
    ("surface w half dropped", "F3(w, j, i, nk + 1) = 0.5 * (zuh + zvh);", "F3(w, j, i, nk + 1) = (zuh + zvh);"),
    ("divergence metric", "double zrfmzup = 2.0 / (f0 + F3(fmz, j + 1, i, k));",
     "double zrfmzup = 1.0 / f0;"),
    ("vertical divergence fmzf for fmz", "F3(S->fmz, j, i, k) * dtrdz * (F3(S->s, j, i, k)",
     "F3(S->fmzf, j, i, k) * dtrdz * (F3(S->s, j, i, k)"),
    ("exner rovcp for rdrcv", "(1.0 - RDRCV * F3(zdiv2, j, i, k))", "(1.0 - ROVCP * F3(zdiv2, j, i, k))"),
    ("pgf cvd for cpd", "double zrom1u = 0.5 * CPD *", "double zrom1u = 0.5 * CVD *"),
    ("pgf sign", "                         zcx * zrom1u * (F3(pai, j, i, k) - F3(pai, j - 1, i, k));",
     "                         -zcx * zrom1u * (F3(pai, j, i, k) - F3(pai, j - 1, i, k));"),
    ("terrain pgf interface decay", "zfz * F2(S->hx, j, i) * S->gzitakh[k - 1]",
     "zfz * F2(S->hx, j, i) * S->gzitak[k - 1]"),
    ("gw counter ns-1", "F3(fmzf, j, i, k) * F3(fmzf, j, i, k) * (double)ns * dtrdz",
     "F3(fmzf, j, i, k) * F3(fmzf, j, i, k) * (double)(ns - 1) * dtrdz"),
    ("gw fmzf once", "F3(fmzf, j, i, k) * F3(fmzf, j, i, k) * (double)ns * dtrdz",
     "F3(fmzf, j, i, k) * (double)ns * dtrdz"),
    ("gw sign", "zrom1w = zrom1w - CPD * F3(w, j, i, k)", "zrom1w = zrom1w + CPD * F3(w, j, i, k)"),
    ("zdth counter ns", "(double)(ns - 1) * dts * WLHV * WLHV", "(double)ns * dts * WLHV * WLHV"),
    ("zdth rgas for rwat", "F3(pai, j, i, k - 1) * RWAT * F3(S->t", "F3(pai, j, i, k - 1) * RGAS * F3(S->t"),
    ("zdth pi dropped", "(CPD * F3(pai, j, i, k - 1) * RWAT", "(CPD * RWAT"),
    ("zdth threshold", "F3(qv, j, i, k) > 0.96 * F3(S->qsat, j, i, k)", "F3(qv, j, i, k) > 0.95 * F3(S->qsat, j, i, k)"),
    ("fulleq qv ratio swapped", "(1.0 + 0.86 * qv + 3.2 * qc) / (1.0 + 0.96 * qv + 4.8 * qc)",
     "(1.0 + 0.96 * qv + 3.2 * qc) / (1.0 + 0.86 * qv + 4.8 * qc)"),
    ("fulleq liquid/ice swapped", "4.2 * qwltot + 2.1 * qwitot", "2.1 * qwltot + 4.2 * qwitot"),
]


def _build(tmp, name, text, repl):
    src = open(SRC).read()
    assert src.count(text) >= 1, name
    path = os.path.join(tmp, "mutant.c")
    with open(path, "w") as f:
        f.write(src.replace(text, repl))
    so = os.path.join(tmp, "libmutant.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c99", "-ffp-contract=off", "-fno-fast-math", "-shared",
                    "-I", os.path.join(ROOT, "oracle"), "-o", so, path, "-lm"], check=True)
    return so


def _pins(so):
    env = dict(os.environ, MLC_ORACLE_MUTANT=so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", PINS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-2000:]


def test_pins_pass_on_the_unmodified_oracle(tmp_path):
    so = _build(str(tmp_path), "none", "#include", "#include")
    rc, out = _pins(so)
    assert rc == 0, out


@pytest.mark.parametrize("name,text,repl", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_catch_planted_mistake(tmp_path, name, text, repl):
    so = _build(str(tmp_path), name, text, repl)
    rc, out = _pins(so)
    assert rc != 0, f"the pins did not notice: {name}\n{out}"
