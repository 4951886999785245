"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/moloch_oracle.c is built with gcc together with a small standalone
driver (tests/oracle_san_driver.c) that allocates every array with exactly
the size the Python wrapper gives it; the driver runs init_moloch and a few
steps on a dumped initial state.  Any out-of-bounds access, use after free or
undefined behaviour aborts the program.  Its results must also equal the
production oracle's bit for bit (the same source and floating-point flags).
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import dynamical_model_amd as dm
from oracle_lib import _PTRS_1D, _PTRS_2D, _PTRS_3D, Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAMES = _PTRS_1D + _PTRS_2D + _PTRS_3D
FLAGS = ["-O2", "-std=c99", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas", "-g",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = tmp_path_factory.mktemp("san") / "oracle_san"
    subprocess.run(["gcc", *FLAGS, "-o", str(exe), os.path.join(ROOT, "oracle", "moloch_oracle.c"),
                    os.path.join(ROOT, "tests", "oracle_san_driver.c"), "-lm"], check=True)
    return str(exe)


def _dump(path, o: Oracle, nsteps: int) -> None:
    S = o.S
    with open(path, "wb") as f:
        f.write(struct.pack("4i", S.nj, S.ni, S.nk, nsteps))
        f.write(struct.pack("7d", S.dt, S.dx, S.dzita, S.ztop, S.zh, S.a0, S.anu2))
        f.write(struct.pack("12i", S.nadv, S.nsound, S.nqx, S.iqfrst, S.rotllr, S.divfilter, S.halo_mpi,
                            S.do_fulleq, S.do_divdamp, S.do_vadvtwice, S.do_filterpai, S.do_filtertheta))
        for n in NAMES:
            f.write(np.ascontiguousarray(o.a[n], np.float64).tobytes())


def _cloudy(hs, c):
    kz, ny, nx = hs.fields["qc"].shape
    jj, ii = np.meshgrid((np.arange(nx) - 2) % c.jx, (np.arange(ny) - 2) % c.iy)
    bump = 0.25 * (1 + np.cos(2 * np.pi * jj / c.jx)) * (1 + np.cos(2 * np.pi * ii / c.iy))
    prof = np.exp(-((np.arange(kz) - 0.6 * kz) / (0.2 * kz)) ** 2)[:, None, None]
    for n, amp in (("qc", 3e-4), ("qr", 1e-4), ("qi", 5e-5), ("qs", 5e-5)):
        hs.fields[n][:] = amp * prof * bump[None]


CASES = {
    "hill": (dict(config=0, jx=24, iy=20, kz=10), {}, False),
    "filters_mpi_not_rotllr": (dict(config=0, jx=24, iy=20, kz=10),
                               dict(divfilter=1, filterpai=1, filtertheta=1, halo_mpi=1, rotllr=0), False),
    "minimum_grid": (dict(config=0, jx=3, iy=4, kz=3, hill_hw=1.0, noise_levels=2), {}, False),
    "saturated_cloudy_bubble": (dict(config=1, jx=32, iy=28, kz=16, hill_hw=6.0, bubble_rcells=5.0, rh=0.99), {},
                                True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_clean_under_sanitizers(driver, tmp_path, name):
    over, pover, cloudy = CASES[name]
    over = dict(over)
    c = dm.case_config(over.pop("config"), **over)
    hs = dm.build_case(c)
    if cloudy:
        _cloudy(hs, c)
    params = dm.case_params(c, **pover)
    nsteps = 2
    o = Oracle(hs.jx, hs.iy, hs.kz, params, hs.levels)
    o.load_fields(hs.fields)
    _dump(tmp_path / "in.bin", o, nsteps)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([driver, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], capture_output=True,
                         text=True, env=env, timeout=600)
    assert run.returncode == 0, run.stderr[-4000:]
    assert "runtime error" not in run.stderr, run.stderr[-4000:]
    # the production oracle on the same input
    o.init_moloch()
    for n in range(1, nsteps + 1):
        o.moloch(n, n == nsteps)
    out = np.fromfile(tmp_path / "out.bin", np.float64)
    pos = 0
    for n in NAMES:
        size = o.a[n].size
        np.testing.assert_array_equal(out[pos:pos + size], o.a[n].ravel(), err_msg=n)
        pos += size
    np.testing.assert_array_equal(out[pos:pos + 6], np.array(o.S.diag))
    assert out[pos + 6] == o.S.cfl_violation
