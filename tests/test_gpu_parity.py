"""GPU (HIP) vs CPU oracle parity of the full moloch step.

The HIP library is called through its C ABI (ctypes).  Inputs are built once
on the host and given byte-identical to both.  FP64 results must agree with
the oracle to round-off (the only differences are the device pow/exp, which
are not correctly rounded); FP32 results are held to the error budget of the
mixed-precision scheme.
"""
from __future__ import annotations

import numpy as np
import pytest

import dynamical_model_amd as dm
from dynamical_model_amd import _lib
from parity import DIAGNOSTIC, PROGNOSTIC, compare, field_errors, run_oracle, small_config2

pytestmark = pytest.mark.gpu

ALL = PROGNOSTIC + DIAGNOSTIC

# FP64 tolerances on relative L2 / max-abs, per number of steps
FP64_REL = {1: 1e-13, 3: 1e-12, 10: 1e-11}


def _report(tag, errs):
    print(f"\n[{tag}]")
    for n, (rel, mx) in errs.items():
        print(f"  {n:6s} relL2 {rel:.3e}  maxabs {mx:.3e}")


def _run(case, nsteps, **pover):
    hs = dm.build_case(case)
    params = dm.case_params(case, **pover)
    print(f"case {case.jx}x{case.iy}x{case.kz} steps={nsteps} params={pover}", flush=True)
    gpu = dm.run_case(hs, params, nsteps)
    ora = run_oracle(hs, params, nsteps)
    return gpu, ora


def _check_fp64(gpu, ora, nsteps, tag, names=ALL):
    errs = field_errors(gpu, ora, names)
    _report(tag, errs)
    tol = FP64_REL[nsteps]
    for n, (rel, mx) in errs.items():
        assert np.isfinite(gpu[n]).all(), n
        assert rel <= tol, f"{tag}: {n} relL2 {rel:.3e} > {tol:.1e}"


@pytest.mark.parametrize("nsteps", [1, 3])
def test_fp64_small_hill(nsteps):
    c = dm.case_config(0, jx=40, iy=36, kz=16)
    gpu, ora = _run(c, nsteps)
    _check_fp64(gpu, ora, nsteps, f"fp64 hill {nsteps}")


def test_fp64_config0_full_grid():
    c = dm.case_config(0)
    gpu, ora = _run(c, 3)
    _check_fp64(gpu, ora, 3, "fp64 config0")


def test_fp64_jet_hills_ragged():
    # configuration-2 physics on a ragged grid (not a multiple of the block)
    c = small_config2(70, 45, 20)
    gpu, ora = _run(c, 3)
    _check_fp64(gpu, ora, 3, "fp64 jet ragged")


def test_fp64_reference_initial_state():
    c = dm.case_config(0, jx=32, iy=28, kz=12, kind=_lib.MLC_CASE_REFERENCE, hill_h=0.0, hill_hw=0.0)
    gpu, ora = _run(c, 3)
    _check_fp64(gpu, ora, 3, "fp64 reference state")


def test_fp64_mpi_halo_order():
    c = dm.case_config(0, jx=30, iy=26, kz=12)
    gpu, ora = _run(c, 3, halo_mpi=1)
    _check_fp64(gpu, ora, 3, "fp64 mpi halo order")


def test_fp64_filters_on():
    c = dm.case_config(0, jx=30, iy=26, kz=12)
    gpu, ora = _run(c, 3, divfilter=1, filterpai=1, filtertheta=1)
    _check_fp64(gpu, ora, 3, "fp64 filters")


def test_fp64_not_rotllr_with_map_factors():
    c = dm.case_config(0, jx=30, iy=26, kz=12)
    hs = dm.build_case(c)
    jj, ii = np.meshgrid(np.arange(-1, c.jx + 3), np.arange(-1, c.iy + 3))
    for name, ph in (("msfu", 0.3), ("msfv", 1.1), ("msfx", 2.0)):
        hs.fields[name][...] = 1.0 + 0.05 * np.sin(2 * np.pi * jj / c.jx + ph) * np.cos(2 * np.pi * ii / c.iy)
        # halos are exact periodic copies, as `param` makes them (:159-162)
        core = hs.fields[name][2:-2, 2:-2]
        hs.fields[name][...] = core[(np.arange(-1, c.iy + 3) - 1) % c.iy][:, (np.arange(-1, c.jx + 3) - 1) % c.jx]
    params = dm.case_params(c, rotllr=0)
    gpu = dm.run_case(hs, params, 3)
    ora = run_oracle(hs, params, 3)
    _check_fp64(gpu, ora, 3, "fp64 not rotllr")


def test_fp64_substep_counts():
    c = dm.case_config(0, jx=24, iy=20, kz=10)
    gpu, ora = _run(c, 3, nsound=2, nadv=1)
    _check_fp64(gpu, ora, 3, "fp64 nsound=2 nadv=1")


def test_fp64_minimum_grid():
    c = dm.case_config(0, jx=3, iy=4, kz=3, hill_hw=1.0, noise_levels=2)
    gpu, ora = _run(c, 3)
    _check_fp64(gpu, ora, 3, "fp64 3x4x3")


def test_fp64_ten_steps():
    c = dm.case_config(1, jx=48, iy=40, kz=20, hill_hw=8.0, bubble_rcells=6.0)
    gpu, ora = _run(c, 10)
    _check_fp64(gpu, ora, 10, "fp64 bubble 10 steps")


def _add_condensates(hs, c):
    """A periodic cloud of all four condensate species in the host state
    (halos included), so that the fulleq moist coefficients (:697-706), the
    epilogue's condensate loading and the advection of qc, qi, qr, qs act on
    nonzero values."""
    kz, ny, nx = hs.fields["qc"].shape
    jj, ii = np.meshgrid((np.arange(nx) - 2) % c.jx, (np.arange(ny) - 2) % c.iy)
    bump = 0.25 * (1 + np.cos(2 * np.pi * jj / c.jx)) * (1 + np.cos(2 * np.pi * ii / c.iy))
    prof = np.exp(-((np.arange(kz) - 0.6 * kz) / (0.2 * kz)) ** 2)[:, None, None]
    for n, amp in (("qc", 3e-4), ("qr", 1e-4), ("qi", 5e-5), ("qs", 5e-5)):
        hs.fields[n][:] = amp * prof * bump[None]


def test_fp64_cold_tropopause():
    """A 185 K tropopause: the upper third of the column is colder than
    -75 C, where the saturation pressures of the prologue and epilogue are
    clamped (:119-181); the other cases stay above -57 C."""
    c = dm.case_config(0, jx=32, iy=28, kz=24, ttrop=185.0)
    gpu, ora = _run(c, 3)
    assert np.count_nonzero(dm.interior(ora.field("t")) < 198.15) > 0
    _check_fp64(gpu, ora, 3, "fp64 cold tropopause")


def test_fp64_saturated_cloudy_bubble():
    """Near-saturated air (RH 99 %) around a warm bubble, with clouds: the
    column solve's moist term (qv > 0.96 qsat where w > 0.1, :645-657) and
    the fulleq coefficients act, which the other cases (RH 50-60 %, no
    condensates) never reach.  The regime is checked on the oracle's state."""
    c = dm.case_config(1, jx=48, iy=40, kz=20, hill_hw=8.0, bubble_rcells=6.0, rh=0.99)
    hs = dm.build_case(c)
    _add_condensates(hs, c)
    params = dm.case_params(c)
    print(f"saturated cloudy bubble {c.jx}x{c.iy}x{c.kz}, 3 steps", flush=True)
    gpu = dm.run_case(hs, params, 3)
    ora = run_oracle(hs, params, 3)
    qv, qsat, w = dm.interior(ora.field("qv")), dm.interior(ora.field("qsat")), dm.interior(ora.field("w"))
    assert np.count_nonzero((qv > 0.96 * qsat) & (w[: qv.shape[0]] > 0.1)) > 0
    assert np.abs(dm.interior(ora.field("qc"))).max() > 1e-4
    _check_fp64(gpu, ora, 3, "fp64 saturated cloudy bubble", ALL + ["qc", "qi", "qr", "qs"])
    # the same case in FP32: default kernels bit-identical to the reference-order ones
    a = dm.run_case(hs, dm.case_params(c, precision=dm.MLC_FP32, unfused=0), 3)
    b = dm.run_case(hs, dm.case_params(c, precision=dm.MLC_FP32, unfused=1), 3)
    for n in PROGNOSTIC + ["qc", "qi", "qr", "qs"]:
        np.testing.assert_array_equal(dm.interior(a[n]), dm.interior(b[n]), err_msg=n)


def test_diagnostics_match_oracle():
    c = dm.case_config(0, jx=40, iy=36, kz=16)
    hs = dm.build_case(c)
    params = dm.case_params(c)
    m = dm.Moloch()
    m.allocate_moloch(c.jx, c.iy, c.kz, params)
    m.load_state(hs)
    m.init_moloch()
    d = m.moloch(1, True)
    m.close()
    o = run_oracle(hs, params, 0)
    o.moloch(1, True)
    got = [d.ps_max, d.ps_min, d.u_max, d.u_min, d.v_max, d.v_min]
    np.testing.assert_allclose(got, list(o.S.diag), rtol=1e-13, atol=1e-13)
    assert d.cfl_violation == 0 == o.S.cfl_violation


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_cfl_violation_detected(bad):
    """A NaN or an Inf in the Exner function at one surface point: after one
    step ps is NaN in a patch of about a sixth of this grid, its extrema over
    the rest are finite, and the lprint check of both the GPU and the oracle
    flags the step (:402-411) with the same extrema; as in the nompi tree,
    the run goes on (one more step)."""
    c = dm.case_config(0, jx=160, iy=120, kz=12)
    hs = dm.build_case(c)
    hs.fields["pai"][c.kz - 1, 2 + 60, 2 + 80] = bad
    params = dm.case_params(c)
    m = dm.Moloch()
    m.allocate_moloch(c.jx, c.iy, c.kz, params)
    m.load_state(hs)
    m.init_moloch()
    d1 = m.moloch(1, True)
    d2 = m.moloch(2, True)
    m.close()
    o = run_oracle(hs, params, 0)
    o.moloch(1, True)
    print(f"{bad}: gpu {d1.cfl_violation} {d2.cfl_violation} oracle {o.S.cfl_violation}", flush=True)
    assert d1.cfl_violation == 1 == o.S.cfl_violation
    assert d2.cfl_violation == 1
    got = [d1.ps_max, d1.ps_min, d1.u_max, d1.u_min, d1.v_max, d1.v_min]
    assert np.isfinite(got[:2]).all()
    np.testing.assert_allclose(got, list(o.S.diag), rtol=1e-13, atol=1e-13)


def test_fp32_error_budget():
    c = dm.case_config(0, jx=40, iy=36, kz=16)
    hs = dm.build_case(c)
    gpu = dm.run_case(hs, dm.case_params(c, precision=dm.MLC_FP32), 3)
    ora = run_oracle(hs, dm.case_params(c), 3)
    errs = field_errors(gpu, ora, PROGNOSTIC)
    _report("fp32 vs fp64 oracle, 3 steps", errs)
    # FP32 storage of u, v, w and the water species with theta, the
    # temperature and the Exner function in FP64: about five times the
    # measured errors (u 2.8e-7, v 3.9e-5, w 4.8e-5, pi 1.6e-9, theta 9.4e-9,
    # qv 2.8e-7), so that a regression shows
    budget = {"u": 1.5e-6, "v": 2e-4, "w": 2.5e-4, "pai": 1e-8, "tetav": 5e-8, "qv": 1.5e-6}
    for n, (rel, _) in errs.items():
        assert rel < budget[n], n


def test_fp32_core_accuracy_over_300_steps():
    """The FP32 core against the FP64 core (itself bit-identical to the
    oracle) over 300 steps of configuration-2 physics (jet, hills, RH 60 %)
    on 192 x 160 x 60: every prognostic field within the 1e-3 relative L2
    target.  Rounding theta or the temperature to FP32 (their storage before)
    put w at 2e-3 after 100 steps of this flow."""
    c = small_config2(192, 160, 60, nhills=6, hill_hmin=200.0, hill_hmax=1500.0, hill_hwmin=6.0,
                      hill_hwmax=24.0)
    runs = {}
    for prec in (dm.MLC_FP64, dm.MLC_FP32):
        m = dm.Moloch()
        m.allocate_moloch(c.jx, c.iy, c.kz, dm.case_params(c, precision=prec))
        m.load_case_device(c)
        runs[prec] = m
    print(f"{c.jx}x{c.iy}x{c.kz} configuration-2 physics: FP32 vs FP64, 300 steps", flush=True)
    for n in range(1, 301):
        for m in runs.values():
            m.moloch(n, False)
    errs = {}
    for f in PROGNOSTIC:
        a = dm.interior(runs[dm.MLC_FP32].get_field(f))
        b = dm.interior(runs[dm.MLC_FP64].get_field(f))
        errs[f] = compare(a, b)
    for m in runs.values():
        m.close()
    _report("fp32 vs fp64 core, 300 steps", errs)
    for n, (rel, _) in errs.items():
        assert np.isfinite(rel) and rel < 1e-3, n


FUSION_OPTS = [{}, {"divfilter": 1}, {"rotllr": 0}, {"halo_mpi": 1}]


@pytest.mark.parametrize("prec", [dm.MLC_FP64, dm.MLC_FP32])
@pytest.mark.parametrize("grid", [(70, 45, 20), (29, 13, 7), (3, 4, 3), (130, 9, 5)])
@pytest.mark.parametrize("opts", range(len(FUSION_OPTS)))
def test_fused_equals_unfused(prec, grid, opts):
    """The two-kernel wafone (vertical passes, then both horizontal sweeps in
    one strip kernel) performs the reference's arithmetic in the reference's
    order: bit-identical to the three-sweep kernels, on grids smaller than
    one strip, with partial strips, and on every map-factor / filter path."""
    jx, iy, kz = grid
    over = FUSION_OPTS[opts]
    if over and grid != (29, 13, 7):
        pytest.skip("option variants run on one grid")
    c = small_config2(jx, iy, kz) if jx > 10 else dm.case_config(0, jx=jx, iy=iy, kz=kz, hill_hw=1.0,
                                                                 noise_levels=2)
    hs = dm.build_case(c)
    print(f"fused vs unfused {grid} prec={prec} {over}", flush=True)
    a = dm.run_case(hs, dm.case_params(c, precision=prec, unfused=1, **over), 2)
    b = dm.run_case(hs, dm.case_params(c, precision=prec, unfused=0, **over), 2)
    for n in PROGNOSTIC:
        np.testing.assert_array_equal(dm.interior(a[n]), dm.interior(b[n]), err_msg=n)


@pytest.mark.parametrize("prec", [dm.MLC_FP64, dm.MLC_FP32])
@pytest.mark.parametrize("kz", [3, 9, 17, 60, 65, 66])
def test_segmented_solve_equals_one_lane(prec, kz):
    """The segmented column solve (a column's levels split over eight lanes,
    elimination coefficients held in registers and handed across segments
    through LDS, divdamp's wind update applied by the momentum kernel) is
    bit-identical to the one-lane sweep in the reference's order: level counts
    that leave segments empty (3), uneven (9, 17), one short of full (60, the
    benchmark's), exactly full (65 = 8 x 8 + 1) and past what the segments
    hold (66, where the library falls back to the sweep)."""
    c = small_config2(45, 7, kz)
    hs = dm.build_case(c)
    print(f"segmented vs one-lane kz={kz} prec={prec}", flush=True)
    a = dm.run_case(hs, dm.case_params(c, precision=prec, unfused=1), 2)
    b = dm.run_case(hs, dm.case_params(c, precision=prec, unfused=0), 2)
    for n in PROGNOSTIC:
        np.testing.assert_array_equal(dm.interior(a[n]), dm.interior(b[n]), err_msg=n)
    if prec == dm.MLC_FP64 and kz == 60:
        _check_fp64(b, run_oracle(hs, dm.case_params(c), 2), 3, "segmented kz=60 vs oracle", PROGNOSTIC)


def test_case_load_matches_host_build():
    c = small_config2(33, 29, 9)
    hs = dm.build_case(c)
    params = dm.case_params(c)
    m = dm.Moloch()
    m.allocate_moloch(c.jx, c.iy, c.kz, params)
    m.load_case(c)
    for name in ("u", "pai", "t", "qv", "fmz", "fmzf", "hx", "hy"):
        got = m.get_field(name)
        np.testing.assert_array_equal(dm.interior(got), dm.interior(hs.fields[name]), err_msg=name)
    m.close()


@pytest.mark.parametrize("precision", [dm.MLC_FP64, dm.MLC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("halo_mpi", [0, 1], ids=["nompi", "mpi"])
def test_poisoned_halos_change_nothing(precision, halo_mpi):
    """Debug mode of the halo fills: NaN in the two-point halo ring of every
    state and scratch field before each step.  Every halo cell a kernel reads
    is refilled in the step before it is read (the nompi tree's stale Exner
    seam included: its halo holds the substep's pre-update values, written
    in the same substep), so the interior comes out finite and bit-identical
    to the run without poison."""
    c = small_config2(70, 45, 20)
    hs = dm.build_case(c)
    params = dm.case_params(c, precision=precision, halo_mpi=halo_mpi)
    print(f"poisoned halos 70x45x20 precision {precision} halo_mpi {halo_mpi}", flush=True)
    out = []
    for poison in (False, True):
        m = dm.Moloch()
        m.allocate_moloch(c.jx, c.iy, c.kz, params)
        m.load_state(hs)
        m.init_moloch()
        m.poison_halos(poison)
        for n in range(1, 4):
            m.moloch(n)
        m.sync()
        out.append({f: dm.interior(m.get_field(f)).copy() for f in PROGNOSTIC + ["t", "ps"]})
        m.close()
    for f in out[0]:
        assert np.isfinite(out[1][f]).all(), f
        assert np.array_equal(out[0][f], out[1][f]), f
