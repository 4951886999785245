"""The device case builder against the host builder (mlc_case_build, whose
reference kind is pinned bit for bit to the reference's fill_atmosphere):
same fields to the last bits of the device exp/pow, the theta noise drawn
from the same splitmix64 sequence, and the same steps afterwards."""
from __future__ import annotations

import time

import numpy as np
import pytest

import dynamical_model_amd as dm
from dynamical_model_amd import _lib
from parity import compare, small_config2

pytestmark = pytest.mark.gpu

CHECK = ["u", "v", "w", "pai", "t", "tetav", "qv", "qc", "ps", "fmz", "fmzf", "zeta", "hx", "hy"]


def _load(case, device: bool, **pover):
    m = dm.Moloch()
    m.allocate_moloch(case.jx, case.iy, case.kz, dm.case_params(case, **pover))
    (m.load_case_device if device else m.load_case)(case)
    return m


CASES = {
    "jet-hills-noise": lambda: small_config2(70, 45, 20),
    "bubble-noise": lambda: dm.case_config(1, jx=64, iy=48, kz=16),
    "reference": lambda: dm.case_config(0, jx=40, iy=36, kz=16, kind=_lib.MLC_CASE_REFERENCE, hill_h=500.0,
                                        hill_hw=10.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_builder_matches_host(name):
    case = CASES[name]()
    print(f"case {name} {case.jx}x{case.iy}x{case.kz}", flush=True)
    host, dev = _load(case, False), _load(case, True)
    for f in CHECK:
        a, b = dev.get_field(f), host.get_field(f)
        rel, mx = compare(dm.interior(a), dm.interior(b))
        # the whole halo-2 ring too (filled periodically by both)
        ring = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        print(f"  {f:5s} rel {rel:.2e} ring {ring:.2e}")
        assert rel <= 1e-13 and ring <= 1e-13, f
    for n in (1, 2):
        host.moloch(n)
        dev.moloch(n)
    for f in ("u", "w", "pai", "tetav"):
        rel, _ = compare(dm.interior(dev.get_field(f)), dm.interior(host.get_field(f)))
        assert rel <= 1e-11, f
    host.close()
    dev.close()


def test_device_builder_is_fast():
    case = dm.case_config(4, jx=1024, iy=1024, kz=60)
    m = dm.Moloch()
    m.allocate_moloch(case.jx, case.iy, case.kz, dm.case_params(case, precision=dm.MLC_FP32))
    t0 = time.perf_counter()
    m.load_case_device(case)
    m.sync()
    td = time.perf_counter() - t0
    t0 = time.perf_counter()
    m.load_case(case)
    m.sync()
    th = time.perf_counter() - t0
    print(f"1024x1024x60: device build + init {td:.2f} s, host build + upload + init {th:.2f} s")
    assert td < th
    m.close()


@pytest.mark.parametrize("precision", [dm.MLC_FP64, dm.MLC_FP32], ids=["fp64", "fp32"])
def test_checkpoint_resume_is_bitwise(tmp_path, precision):
    """A handle restored from the state dump after step 3 and a handle that
    never stopped agree bit for bit after step 6."""
    c = small_config2(48, 40, 16)
    params = dm.case_params(c, precision=precision)
    a = dm.Moloch()
    a.allocate_moloch(c.jx, c.iy, c.kz, params)
    a.load_case_device(c)
    print(f"checkpoint 48x40x16 precision {precision}", flush=True)
    for n in (1, 2, 3):
        a.moloch(n)
    path = str(tmp_path / "state.npz")
    a.save_state(path, 3)
    for n in (4, 5, 6):
        a.moloch(n)
    b = dm.Moloch()
    b.allocate_moloch(c.jx, c.iy, c.kz, params)
    b.load_case_device(c)  # geometry, then the state of step 3
    assert b.restore_state(path) == 3
    for n in (4, 5, 6):
        b.moloch(n)
    try:
        for f in dm.Moloch.STATE_FIELDS:
            assert np.array_equal(dm.interior(a.get_field(f)), dm.interior(b.get_field(f))), f
    finally:
        a.close()
        b.close()
