"""Pins of the CPU oracle (oracle/moloch_oracle.c).

The reference has no tests, golden vectors or fixtures, and its timestep
module cannot be built here (nompi/mod_moloch.F90 uses mod_mppparam, which
needs netCDF-Fortran).  The oracle is pinned by:
  * golden vectors from the reference modules that do build (constants,
    gzita, zita levels, geometry and fill_atmosphere), tests/golden/;
  * closed forms and invariants the mathematics fixes, chosen so that a
    dropped term, a wrong sign or index, or a transposed operand fails one.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

import dynamical_model_amd as dm
from dynamical_model_amd import _lib
import oracle_lib
from parity import small_config2

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
I = dm.interior


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


def _case_from_gold(gold, name):
    jx, iy, kz, ztop, zh, a0, hh, hw = gold[f"{name}/args"]
    c = dm.case_config(0, jx=int(jx), iy=int(iy), kz=int(kz), kind=_lib.MLC_CASE_REFERENCE,
                       ztop=ztop, zh=zh, a0=a0, hill_h=hh, hill_hw=hw)
    return c


# ----------------------------------------------------------- golden vectors
def test_constants_match_reference_modules(gold):
    names = [str(n) for n in gold["constant_names"]]
    c = oracle_lib.constants()
    for n, v in zip(names, gold["hill/constants"]):
        assert c[n] == v, n


@pytest.mark.parametrize("name", ["hill", "stretched", "flat", "a0"])
def test_gzita_matches_reference(gold, name):
    L = oracle_lib.lib()
    _, _, kz, ztop, _, a0, _, _ = gold[f"{name}/args"]
    for z, g in zip(gold[f"{name}/zita"], gold[f"{name}/gzita_f"]):
        assert L.ora_gzita(z, ztop, a0) == g
    for z, g in zip(gold[f"{name}/zitah"], gold[f"{name}/gzita_h"]):
        assert L.ora_gzita(z, ztop, a0) == g


@pytest.mark.parametrize("name", ["hill", "stretched", "flat", "a0"])
def test_case_builder_reproduces_fill_atmosphere(gold, name):
    """The package's reference-state builder equals param + fill_atmosphere."""
    c = _case_from_gold(gold, name)
    hs = dm.build_case(c)
    f = hs.fields
    np.testing.assert_array_equal(hs.levels["zita"], gold[f"{name}/zita"])
    np.testing.assert_array_equal(hs.levels["zitah"], gold[f"{name}/zitah"])
    np.testing.assert_array_equal(f["hx"][2:-2, 2:-1], gold[f"{name}/hx"])
    np.testing.assert_array_equal(f["hy"][2:-1, 2:-2], gold[f"{name}/hy"])
    for n in ("zeta", "fmz", "fmzf", "t", "qv", "pai", "ps"):
        np.testing.assert_array_equal(I(f[n]), gold[f"{name}/{n}"], err_msg=n)


# ------------------------------------------------------------- saturation
def _goff_gratch_water(T):
    Ts = 373.16
    lg = (-7.90298 * (Ts / T - 1) + 5.02808 * np.log10(Ts / T)
          - 1.3816e-7 * (10 ** (11.344 * (1 - T / Ts)) - 1)
          + 8.1328e-3 * (10 ** (-3.49149 * (Ts / T - 1)) - 1) + np.log10(1013.246))
    return 100.0 * 10 ** lg


def _goff_gratch_ice(T):
    T0 = 273.16
    lg = (-9.09718 * (T0 / T - 1) - 3.56654 * np.log10(T0 / T)
          + 0.876793 * (1 - T / T0) + np.log10(6.1071))
    return 100.0 * 10 ** lg


def test_pfesat_against_goff_gratch():
    L = oracle_lib.lib()
    for tc in np.linspace(0.0, 45.0, 46):
        es = L.ora_pfesat(273.15 + tc, 1.0e7)
        assert abs(es / _goff_gratch_water(273.15 + tc) - 1) < 3e-3, tc
    for tc in np.linspace(-60.0, -1.0, 60):
        es = L.ora_pfesat(273.15 + tc, 1.0e7)
        assert abs(es / _goff_gratch_ice(273.15 + tc) - 1) < 5e-3, tc


def test_pfesat_clamps_and_cap():
    L = oracle_lib.lib()
    hi = L.ora_pfesat(273.15 + 100.0, 1e9)
    assert L.ora_pfesat(450.0, 1e9) == hi
    lo = L.ora_pfesat(273.15 - 75.0, 1e9)
    assert L.ora_pfesat(100.0, 1e9) == lo
    # es capped at 0.15 p
    assert L.ora_pfesat(303.15, 1000.0) == pytest.approx(150.0, rel=0, abs=1e-12)
    # continuity at 0 C between the ice and water polynomials (both fit ~611 Pa)
    assert abs(L.ora_pfesat(273.15, 1e7) / L.ora_pfesat(273.15 - 1e-9, 1e7) - 1) < 2e-4


def test_pfwsat_is_a_mixing_ratio():
    """w = ep2*es/(p-es) inverts to the vapour pressure e = w*p/(ep2+w)."""
    L = oracle_lib.lib()
    ep2 = oracle_lib.constants()["ep2"]
    assert abs(ep2 - 18.015 / 28.965) < 1e-4
    for T, p in [(300.0, 1e5), (250.0, 5e4), (220.0, 2e4)]:
        w = L.ora_pfwsat(T, p)
        e = w * p / (ep2 + w)
        assert e == pytest.approx(L.ora_pfesat(T, p), rel=1e-12)


# ------------------------------------------------------------- tridiagonal
def test_thomas_column_solves_assembled_system():
    rng = np.random.default_rng(7)
    L = oracle_lib.lib()
    for nk in (3, 8, 40):
        zw = rng.normal(size=nk + 1)
        zu = rng.uniform(0.0, 50.0, size=nk + 1)
        zd = rng.uniform(0.0, 50.0, size=nk + 1)
        w1, wk1 = 0.0, rng.normal()
        out = np.zeros(nk + 1)
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        L.ora_thomas_column(nk, P(zw), P(zu), P(zd), w1, wk1, P(out))
        # (1+zu+zd) w_k - zd w_{k+1} - zu w_{k-1} = zwexpl_k,  k = 2..nk
        n = nk - 1
        A = np.zeros((n, n))
        b = np.zeros(n)
        for r, k in enumerate(range(2, nk + 1)):
            A[r, r] = 1 + zu[k - 1] + zd[k - 1]
            b[r] = zw[k - 1]
            if k > 2:
                A[r, r - 1] = -zu[k - 1]
            else:
                b[r] += zu[k - 1] * w1
            if k < nk:
                A[r, r + 1] = -zd[k - 1]
            else:
                b[r] += zd[k - 1] * wk1
        ref = np.linalg.solve(A, b)
        np.testing.assert_allclose(out[1:nk], ref, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ helpers
def _oracle(case, **pover):
    hs = dm.build_case(case)
    p = dm.case_params(case, **pover)
    return oracle_lib.oracle_from_host(hs, p), hs, p


def _fill_periodic(a):
    """Refill every halo cell of a canonical array from the interior."""
    core = a[..., 2:-2, 2:-2]
    ni, nj = core.shape[-2:]
    jj = (np.arange(-1, nj + 3) - 1) % nj
    ii = (np.arange(-1, ni + 3) - 1) % ni
    a[...] = core[..., ii[:, None], jj[None, :]]
    return a


# ------------------------------------------------------------ init_moloch
def test_xknu_and_coriolis_closed_forms():
    c = dm.case_config(0, jx=8, iy=6, kz=9)
    o, hs, p = _oracle(c)
    xk = o.a["xknu"][:9]
    assert xk[0] == pytest.approx(1.0, abs=1e-15)
    assert xk[-1] == pytest.approx(0.0625, abs=1e-15)
    assert np.all(np.diff(xk) < 0)
    # mid level: 1/16 + 15/16 sin(pi/4)
    assert xk[4] == pytest.approx(0.0625 + 0.9375 * np.sin(np.pi / 4), rel=1e-14)
    om = 7.2921159e-5
    np.testing.assert_allclose(I(o.a["coru"]), 2 * om * np.sin(np.deg2rad(I(hs.fields["ulat"]))), rtol=1e-13)


# ------------------------------------------------------------------ divdamp
def test_divdamp_fourier_symbol_and_gradient():
    c = dm.case_config(0, jx=16, iy=12, kz=4, hill_h=0.0)
    o, hs, p = _oracle(c)
    nj, ni = 16, 12
    m, n = 3, 2
    j = np.arange(1, nj + 1)
    i = np.arange(1, ni + 1)
    Z = np.cos(2 * np.pi * m * j / nj)[None, :] * np.cos(2 * np.pi * n * i / ni)[:, None]
    z = o.a["zdiv2"]
    z[...] = 0
    for k in range(4):
        z[k, 2:-2, 2:-2] = Z * (k + 1)
    o.a["u"][...] = 0
    o.a["v"][...] = 0
    dts = 2.0
    o.call("ora_divdamp", dts)
    lam = 0.125 * (2 * np.cos(2 * np.pi * m / nj) + 2 * np.cos(2 * np.pi * n / ni)) - 0.5
    coef = 0.0625 * 0.015625 / dts
    xk = o.a["xknu"]
    ddamp = 0.125 * 3000.0 / dts
    for k in range(4):
        np.testing.assert_allclose(I(z[k]), (k + 1) * Z * (1 + coef * lam), rtol=0, atol=1e-13)
        # u(j) += xknu*ddamp*(Z(j)-Z(j-1)): backward difference at the west face
        a = 2 * np.pi * m / nj
        dZ = (-2 * np.sin(a * (j - 0.5)) * np.sin(a / 2))[None, :] * np.cos(2 * np.pi * n * i / ni)[:, None]
        np.testing.assert_allclose(I(o.a["u"][k]), xk[k] * ddamp * (k + 1) * dZ, rtol=0, atol=1e-11)
        b = 2 * np.pi * n / ni
        dZi = np.cos(2 * np.pi * m * j / nj)[None, :] * (-2 * np.sin(b * (i - 0.5)) * np.sin(b / 2))[:, None]
        np.testing.assert_allclose(I(o.a["v"][k]), xk[k] * ddamp * (k + 1) * dZi, rtol=0, atol=1e-11)


# ----------------------------------------------------------- interpolators
def test_stag_cross_interpolators_exact_on_cubics():
    c = dm.case_config(0, jx=20, iy=18, kz=12, hill_h=0.0)
    o, hs, p = _oracle(c)
    P = lambda x: 0.3 + 0.2 * x - 0.05 * x ** 2 + 0.004 * x ** 3
    j = np.arange(-1, 23)
    i = np.arange(-1, 21)
    # u at west faces x = j - 1/2, v at south faces y = i - 1/2
    o.a["u"][...] = P(j - 0.5)[None, None, :]
    o.a["v"][...] = P(i - 0.5)[None, :, None]
    o.call("ora_uvstagtox")
    ux = o.a["ux"][0, 2:-2, 2:-2]
    vx = o.a["vx"][0, 2:-2, 2:-2]
    jin = np.arange(1, 21)
    iin = np.arange(1, 19)
    # points whose stencil stays off the (periodically refilled) halo
    np.testing.assert_allclose(ux[:, 1:-2], np.broadcast_to(P(jin[1:-2])[None, :], ux[:, 1:-2].shape), atol=1e-12)
    np.testing.assert_allclose(vx[1:-2, :], np.broadcast_to(P(iin[1:-2])[:, None], vx[1:-2, :].shape), atol=1e-12)
    # back: cross -> face
    o.a["ux"][...] = P(j)[None, None, :].astype(float)
    o.a["vx"][...] = P(i)[None, :, None].astype(float)
    o.call("ora_xtouvstag")
    u = o.a["u"][0, 2:-2, 2:-2]
    v = o.a["v"][0, 2:-2, 2:-2]
    np.testing.assert_allclose(u[:, 2:-1], np.broadcast_to(P(jin[2:-1] - 0.5)[None, :], u[:, 2:-1].shape), atol=1e-12)
    np.testing.assert_allclose(v[2:-1, :], np.broadcast_to(P(iin[2:-1] - 0.5)[:, None], v[2:-1, :].shape), atol=1e-12)
    # vertical: w at interfaces k (k=1 top), wx at mid-levels k+1/2
    kz = 12
    kf = np.arange(1, kz + 2)
    o.a["w"][...] = P(kf.astype(float))[:, None, None]
    o.call("ora_wstagtox")
    wx = o.a["wx"][:, 5, 5]
    np.testing.assert_allclose(wx[1:kz - 1], P(np.arange(2, kz) + 0.5), atol=1e-12)
    L1 = lambda x: 1.5 - 0.25 * x
    o.a["w"][...] = L1(kf.astype(float))[:, None, None]
    o.call("ora_wstagtox")
    assert o.a["wx"][0, 5, 5] == pytest.approx(L1(1.5)) and o.a["wx"][kz - 1, 5, 5] == pytest.approx(L1(kz + 0.5))
    o.a["wx"][...] = P(np.arange(1, kz + 1) + 0.5)[:, None, None]
    o.call("ora_xtowstag")
    np.testing.assert_allclose(o.a["w"][2:kz - 1, 5, 5], P(np.arange(3, kz).astype(float)), atol=1e-12)


def test_uvstagtox_periodic_closure():
    """Brute force of the 2-wide periodic halos: ux on a random periodic u."""
    c = dm.case_config(0, jx=11, iy=9, kz=3, hill_h=0.0)
    o, hs, p = _oracle(c)
    rng = np.random.default_rng(3)
    U = rng.normal(size=(3, 9, 11))
    V = rng.normal(size=(3, 9, 11))
    o.a["u"][...] = np.nan
    o.a["v"][...] = np.nan
    o.a["u"][:, 2:-2, 2:-2] = U
    o.a["v"][:, 2:-2, 2:-2] = V
    o.call("ora_uvstagtox")
    r = lambda a, s, ax: np.roll(a, -s, axis=ax)
    ux = 0.5625 * (r(U, 1, 2) + U) - 0.0625 * (r(U, 2, 2) + r(U, -1, 2))
    vx = 0.5625 * (r(V, 1, 1) + V) - 0.0625 * (r(V, 2, 1) + r(V, -1, 1))
    np.testing.assert_allclose(I(o.a["ux"]), ux, rtol=0, atol=1e-14)
    np.testing.assert_allclose(I(o.a["vx"]), vx, rtol=0, atol=1e-14)


# ------------------------------------------------------------------ wafone
def _waf_setup(nj=24, ni=20, nk=10):
    c = dm.case_config(0, jx=nj, iy=ni, kz=nk, hill_h=0.0)
    o, hs, p = _oracle(c)
    for n in ("fmz", "fmzf", "mu", "mv", "mx", "rmu", "rmv", "mx2"):
        o.a[n][...] = 1.0
    for n in ("u", "v", "s"):
        o.a[n][...] = 0.0
    return o


def test_wafone_preserves_constants_and_zero():
    c = small_config2(22, 18, 10)
    o, hs, p = _oracle(c)
    rng = np.random.default_rng(11)
    o.a["u"][...] = rng.normal(10, 3, size=o.a["u"].shape)
    o.a["v"][...] = rng.normal(0, 3, size=o.a["v"].shape)
    o.a["s"][...] = rng.normal(0, 0.05, size=o.a["s"].shape)
    o.a["s"][0] = 0.0  # s(1) = s(kz+1) = 0 whenever wafone runs (:843-849)
    o.a["s"][-1] = 0.0
    _fill_periodic(o.a["u"])
    _fill_periodic(o.a["v"])
    q = np.full(o.a["t"].shape, 3.7)
    o.wafone(q, 10.0)
    np.testing.assert_allclose(I(q), 3.7, rtol=1e-13)
    z = np.zeros(o.a["t"].shape)
    o.wafone(z, 10.0)
    assert np.all(I(z) == 0.0)


def _toro_waf_1d(q, c, periodic=True):
    """Textbook WAF (Toro, Riemann Solvers, sec. 13.7) with the superbee
    limiter for q_t + a q_x = 0 on a uniform grid, Courant number c.
    Interface m+1/2 lies between cells m and m+1."""
    n = q.size
    qe = np.concatenate([q[-2:], q, q[:2]]) if periodic else np.concatenate([q[:1]] * 2 + [q] + [q[-1:]] * 2)
    F = np.zeros(n + 1)  # F[m] at interface between cells m-1 and m (0-based cells)
    for m in range(n + 1):
        L, R = m - 1 + 2, m + 2
        dloc = qe[R] - qe[L]
        if c > 0:
            dup = qe[L] - qe[L - 1]
        else:
            dup = qe[R + 1] - qe[R]
        r = dup / dloc if dloc != 0 else (np.inf if dup > 0 else (-np.inf if dup < 0 else 0.0))
        B = max(0.0, min(2.0, max(r, min(2 * r, 1.0))))
        phi = np.sign(c) * (1 - (1 - abs(c)) * B)
        F[m] = 0.5 * (1 + phi) * qe[L] + 0.5 * (1 - phi) * qe[R]
    if not periodic:
        F[0] = F[n] = 0.0
    return F


def test_wafone_zonal_is_textbook_waf():
    o = _waf_setup()
    rng = np.random.default_rng(5)
    dta = 10.0
    for cour in (0.37, -0.61, 0.93):
        q = np.zeros(o.a["t"].shape)
        Q = rng.normal(size=(10, 20, 24))
        q[:, 2:-2, 2:-2] = Q
        o.a["u"][...] = cour * 3000.0 / dta
        o.wafone(q, dta)
        for k in (0, 7):
            for i in (0, 13):
                F = _toro_waf_1d(Q[k, i], cour)
                ref = Q[k, i] + cour * (F[:-1] - F[1:])
                np.testing.assert_allclose(q[k, i + 2, 2:-2], ref, rtol=0, atol=1e-12)


def test_wafone_meridional_is_textbook_waf():
    o = _waf_setup()
    rng = np.random.default_rng(6)
    dta = 10.0
    for cour in (0.44, -0.29):
        q = np.zeros(o.a["t"].shape)
        Q = rng.normal(size=(10, 20, 24))
        q[:, 2:-2, 2:-2] = Q
        o.a["v"][...] = cour * 3000.0 / dta
        o.wafone(q, dta)
        for k in (2,):
            for j in (0, 9, 23):
                F = _toro_waf_1d(Q[k, :, j], cour)
                ref = Q[k, :, j] + cour * (F[:-1] - F[1:])
                np.testing.assert_allclose(q[k, 2:-2, j + 2], ref, rtol=0, atol=1e-12)


def test_wafone_vertical_is_textbook_waf_twice():
    """Two half-step vertical passes; level 1 is the top, s>0 carries k+1 -> k."""
    o = _waf_setup()
    rng = np.random.default_rng(8)
    dta, dzita = 10.0, o.S.dzita
    for cour in (0.8, -0.55):  # Courant number of each half pass
        q = np.zeros(o.a["t"].shape)
        Q = rng.normal(size=(10, 20, 24))
        q[:, 2:-2, 2:-2] = Q
        o.a["s"][...] = cour / (0.5 * dta / dzita)
        o.wafone(q, dta)
        col = Q[::-1, 4, 7].copy()  # bottom-up order: flow from k+1 to k is "rightward"
        for _ in range(2):
            F = _toro_waf_1d(col, cour, periodic=False)
            col = col + cour * (F[:-1] - F[1:])
        np.testing.assert_allclose(q[:, 6, 9], col[::-1], rtol=0, atol=1e-12)


def test_wafone_courant_one_is_an_exact_shift():
    o = _waf_setup()
    rng = np.random.default_rng(9)
    q = np.zeros(o.a["t"].shape)
    Q = rng.normal(size=(10, 20, 24))
    q[:, 2:-2, 2:-2] = Q
    o.a["u"][...] = 3000.0 / 10.0
    o.wafone(q, 10.0)
    np.testing.assert_allclose(I(q), np.roll(Q, 1, axis=2), rtol=0, atol=1e-13)


# -------------------------------------------------------- prologue/epilogue
def test_prologue_potential_and_virtual_temperature():
    c = dm.case_config(0, jx=10, iy=8, kz=6, hill_h=0.0)
    o, hs, p = _oracle(c)
    o.call("ora_prologue")
    T, pai, qv = I(o.a["t"]), I(o.a["pai"]), I(o.field("qv"))
    pres = 1e5 * pai ** (1004.70 / 287.06)
    theta_v = T * (1 + 0.6078 * qv) * (1e5 / pres) ** (287.06 / 1004.70)
    np.testing.assert_allclose(I(o.a["tetav"]), theta_v, rtol=2e-6)


def test_epilogue_ps_inverts_reference_fill(gold):
    """fill_atmosphere derives pai(kz) from ps with the same lapse-rate
    extrapolation the epilogue uses to diagnose ps: the round trip is exact."""
    for name in ("hill", "stretched", "a0"):
        c = _case_from_gold(gold, name)
        o, hs, p = _oracle(c)
        o.call("ora_prologue")
        o.call("ora_epilogue")
        np.testing.assert_allclose(I(o.a["ps"]), gold[f"{name}/ps"], rtol=1e-13)
        np.testing.assert_allclose(I(o.a["t"]), gold[f"{name}/t"], rtol=1e-13)


def test_epilogue_ps_hydrostatic_closed_form():
    """For a constant lapse rate equal to lrate the extrapolation is the
    hypsometric equation with the layer-mean temperature."""
    c = dm.case_config(0, jx=6, iy=5, kz=8, hill_h=0.0, rh=0.0, noise_amp=0.0)
    o, hs, p = _oracle(c)
    lr = 0.00649
    zeta = I(o.a["zeta"])
    tvK = 285.0
    o.field("qv")[...] = 0.0
    o.a["t"][...] = 0.0
    o.a["t"][:, 2:-2, 2:-2] = tvK + lr * (zeta[-1] - zeta)
    o.call("ora_prologue")
    o.call("ora_epilogue")
    pK = I(o.a["p"])[-1]
    zK = zeta[-1]
    Ts = tvK + lr * zK
    g, R = 9.80665, oracle_lib.constants()["rgas"]
    exact = pK * (Ts / tvK) ** (g / (R * lr))
    # the mean-temperature form differs from the exact integral at second
    # order: ln(exact/approx) = g*zK/(R*Tm) * (a/Tm)**2/3, a = lr*zK/2
    tm = tvK + 0.5 * lr * zK
    second = g * zK / (R * tm) * (0.5 * lr * zK / tm) ** 2 / 3
    np.testing.assert_allclose(np.log(exact / I(o.a["ps"])), second, rtol=0.02)
    np.testing.assert_allclose(I(o.a["ps"]), exact, rtol=1e-6)


# ---------------------------------------------------------- whole step
def test_rest_state_is_steady():
    c = dm.case_config(0, jx=16, iy=12, kz=10, hill_h=0.0, u0=0.0, noise_amp=0.0, zero_coriolis=1)
    o, hs, p = _oracle(c)
    pai0 = I(o.a["pai"]).copy()
    t0 = I(o.a["t"]).copy()
    for n in range(1, 4):
        o.moloch(n)
    assert np.abs(I(o.a["u"])).max() < 1e-10
    assert np.abs(I(o.a["v"])).max() < 1e-10
    assert np.abs(I(o.a["w"])).max() < 1e-10
    np.testing.assert_allclose(I(o.a["pai"]), pai0, rtol=1e-13)
    np.testing.assert_allclose(I(o.a["t"]), t0, rtol=1e-12)


def test_uniform_wind_flat_is_steady():
    c = dm.case_config(0, jx=14, iy=12, kz=10, hill_h=0.0, u0=7.0, v0=-3.0, noise_amp=0.0, zero_coriolis=1)
    o, hs, p = _oracle(c)
    pai0 = I(o.a["pai"]).copy()
    for n in range(1, 3):
        o.moloch(n)
    np.testing.assert_allclose(I(o.a["u"]), 7.0, rtol=1e-12)
    np.testing.assert_allclose(I(o.a["v"]), -3.0, rtol=1e-12)
    assert np.abs(I(o.a["w"])).max() < 1e-10
    np.testing.assert_allclose(I(o.a["pai"]), pai0, rtol=1e-13)


def _shift_state(hs, dj, di):
    out = hs.copy()
    for n, a in out.fields.items():
        core = a[..., 2:-2, 2:-2]
        a[..., 2:-2, 2:-2] = np.roll(np.roll(core, dj, axis=-1), di, axis=-2)
        _fill_periodic(a)
    return out


@pytest.mark.parametrize("over", [{}, {"halo_mpi": 1}, {"divfilter": 1}, {"rotllr": 0}])
def test_threaded_oracle_matches_serial(over):
    """The all-core CPU baseline (the oracle source built with loop-level
    OpenMP threads) computes the serial oracle's bits: every threaded loop
    nest writes only its own points."""
    c = small_config2(61, 37, 17)
    hs = dm.build_case(c)
    p = dm.case_params(c, **over)
    a = oracle_lib.oracle_from_host(hs, p)
    b = oracle_lib.oracle_from_host(hs, p, threaded=True)
    for n in range(1, 3):
        a.moloch(n, True)
        b.moloch(n, True)
    for f in ("u", "v", "w", "pai", "tetav", "qv", "t", "ps"):
        np.testing.assert_array_equal(a.field(f), b.field(f), err_msg=f)


def test_step_is_translation_equivariant_with_fresh_halos():
    """With the MPI tree's halo order and f = 0 every halo that is read is
    fresh, so the periodic step commutes with shifts bit for bit: any wrong
    halo index breaks it.  (With f != 0 the Coriolis terms read the ud/vd
    halos, which both trees copy at substep start, before divdamp changes the
    interior (:514-529, :730-735), so the seam is visible there.)"""
    c = small_config2(18, 16, 9, zero_coriolis=1)
    hs = dm.build_case(c)
    p = dm.case_params(c, halo_mpi=1)
    dj, di = 5, 3
    hs2 = _shift_state(hs, dj, di)
    a = oracle_lib.oracle_from_host(hs, p)
    b = oracle_lib.oracle_from_host(hs2, p)
    for n in range(1, 3):
        a.moloch(n)
        b.moloch(n)
    for f in ("u", "v", "w", "pai", "tetav", "qv", "t", "ps"):
        A = I(a.field(f))
        np.testing.assert_array_equal(np.roll(np.roll(A, dj, axis=-1), di, axis=-2), I(b.field(f)), err_msg=f)


def test_nompi_seam_staleness_is_local():
    """nompi fills the pai halo before the Exner update (:716-727): after one
    substep only u at j=1 and v at i=1 can differ from the MPI order."""
    c = small_config2(18, 16, 9)
    hs = dm.build_case(c)
    res = {}
    for order in (0, 1):
        p = dm.case_params(c, halo_mpi=order, nsound=1, nadv=1)
        o = oracle_lib.oracle_from_host(hs, p)
        o.call("ora_reset_tendencies")
        o.call("ora_prologue")
        o.call("ora_sound", 2.0)
        res[order] = o
    for f in ("w", "pai", "tetav"):
        np.testing.assert_array_equal(I(res[0].field(f)), I(res[1].field(f)), err_msg=f)
    du = I(res[0].field("u")) != I(res[1].field("u"))
    dv = I(res[0].field("v")) != I(res[1].field("v"))
    assert du[:, :, 0].any() and dv[:, 0, :].any()
    assert not du[:, :, 1:].any() and not dv[:, 1:, :].any()


def _mirror_j(hs):
    out = hs.copy()
    nj = hs.jx
    face = np.concatenate([[0], nj - np.arange(1, nj)])  # face j -> NJ+2-j (0-based)
    cell = np.arange(nj)[::-1]
    for n, a in out.fields.items():
        core = a[..., 2:-2, 2:-2].copy()
        if n in ("u", "hx", "msfu"):
            core = core[..., face]
        else:
            core = core[..., cell]
        if n in ("u", "hx", "ux"):
            core = -core
        a[..., 2:-2, 2:-2] = core
        _fill_periodic(a)
    return out, face, cell


def test_step_mirror_symmetry_in_x():
    """x -> -x with u -> -u maps solutions to solutions when f = 0 and every
    halo read is fresh (MPI halo order; the nompi stale Exner halo at the seam
    is not mirror symmetric)."""
    c = small_config2(20, 14, 9, zero_coriolis=1)
    hs = dm.build_case(c)
    p = dm.case_params(c, halo_mpi=1)
    hm, face, cell = _mirror_j(hs)
    a = oracle_lib.oracle_from_host(hs, p)
    b = oracle_lib.oracle_from_host(hm, p)
    for n in range(1, 3):
        a.moloch(n)
        b.moloch(n)
    for f in ("u", "v", "w", "pai", "tetav", "qv"):
        A = I(a.field(f))
        A = A[..., face] if f == "u" else A[..., cell]
        if f == "u":
            A = -A
        B = I(b.field(f))
        scale = np.abs(B).max() + 1e-300
        assert np.abs(A - B).max() / scale < 1e-11, f
