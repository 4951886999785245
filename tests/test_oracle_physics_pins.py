"""Pins of the acoustic substep of the CPU oracle (oracle/moloch_oracle.c,
ora_substep and its phases, restating nompi/mod_moloch.F90:512-838) by what
the physics and the mathematics fix, independently of the oracle's own
formulas:

  * Equation 10 and the surface w (:533-570): the kinematic boundary
    condition and the terrain-following metric term, in closed form from
    the decay function of mod_zita (pinned by golden vectors);
  * the divergence (:575-630): a uniform wind over terrain has no physical
    divergence, so the horizontal part (:575-597) and the vertical part
    made of Equation 10's s (:627-630) cancel up to the truncation error;
  * the Exner update and the pressure-gradient force (:726-787): on a flat
    isothermal rest state a Lamb wave oscillates at the textbook frequency
    of the forward-backward scheme, sin(w dts/2) = (c dts/dx) sin(k dx/2)
    with c^2 = (cpd/cvd) Rd T, and keeps w = 0 and a vertically uniform
    Exner perturbation; over terrain the hydrostatic rest state keeps u = 0
    up to a truncation error of second order in the level spacing;
  * the gravity-wave term of the w equation (:645-650): the interface theta
    anticipated over the elapsed substeps, linear in the substep counter,
    against the physical heights of the levels;
  * the moist term (:651-657): the latent heating of a saturated parcel
    lifted by w over the elapsed substeps, Clausius-Clapeyron against a
    finite difference of the saturation mixing ratio in T, and its two
    switches;
  * the fulleq corrections (:695-706): the seven coefficients follow, to the
    reference's two significant figures, from the heat capacities of moist
    air with condensate.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import dynamical_model_amd as dm
import oracle_lib

I = dm.interior
DTS = 2.0  # dt / (nadv * nsound) of every configuration: 30 s / 15


def _fill_periodic(a):
    core = a[..., 2:-2, 2:-2]
    ni, nj = core.shape[-2:]
    jj = (np.arange(-1, nj + 3) - 1) % nj
    ii = (np.arange(-1, ni + 3) - 1) % ni
    a[...] = core[..., ii[:, None], jj[None, :]]
    return a


def _isothermal(jx, iy, kz, T0=250.0, **kw):
    """A dry isothermal atmosphere at rest (f = 0), from the case builder's
    discrete hydrostatic recursion, with the prologue's theta in place."""
    over = dict(jx=jx, iy=iy, kz=kz, hill_h=0.0, u0=0.0, v0=0.0, noise_amp=0.0, zero_coriolis=1, rh=0.0,
                tsfc=T0, ttrop=T0)
    over.update(kw)
    c = dm.case_config(0, **over)
    hs = dm.build_case(c)
    p = dm.case_params(c, divdamp=0)
    o = oracle_lib.oracle_from_host(hs, p)
    o.call("ora_reset_tendencies")
    o.call("ora_prologue")
    _fill_periodic(o.a["tetav"])
    return o, p


# ------------------------------------------------- Equation 10, surface w
def test_surface_w_and_eq10_closed_form():
    """w(kz+1) = u.grad(h) at the cell centre (kinematic condition), and
    s(k) = -(u.grad h) gzita(zita_k) at the interfaces: the slope of level k
    of z = zeta + h gzita(zeta) at fixed zeta.  hx, hy are the face slopes of
    `param` (golden vectors), so the centre slope is their face average."""
    jx, iy, kz = 20, 16, 12
    c = dm.case_config(0, jx=jx, iy=iy, kz=kz, hill_h=900.0, hill_hw=3.0)
    o = oracle_lib.oracle_from_host(dm.build_case(c), dm.case_params(c))
    o.call("ora_reset_tendencies")
    U, V = 7.0, -4.0
    o.a["u"][...] = U
    o.a["v"][...] = V
    o.call("ora_sound_eq10")
    hx, hy = o.a["hx"], o.a["hy"]
    slope = U * 0.5 * (hx[2:-2, 2:-2] + hx[2:-2, 3:-1]) + V * 0.5 * (hy[2:-2, 2:-2] + hy[3:-1, 2:-2])
    assert np.abs(slope).max() > 0.5  # the terrain is felt
    np.testing.assert_allclose(I(o.a["w"])[kz], slope, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(I(o.a["s"])[kz], -slope, rtol=1e-14, atol=1e-15)
    L = oracle_lib.lib()
    zita = o.a["zita"][: kz + 1]
    for k in range(2, kz + 1):
        g = L.ora_gzita(zita[k - 1], c.ztop, c.a0)
        np.testing.assert_allclose(I(o.a["s"])[k - 1], -slope * g, rtol=1e-13, atol=1e-14)
        # the half level's decay factor would be distinguishable here
        gh = L.ora_gzita(o.a["zitah"][k - 1], c.ztop, c.a0)
        assert abs(gh - g) > 1e-3 * abs(g)


# ----------------------------------------------------------- divergence
@pytest.mark.parametrize("axis", ["x", "y"])
def test_uniform_wind_over_terrain_has_no_divergence(axis):
    """u = U (or v = V), w = 0 over a hill is divergence-free in physical
    space.  In the terrain-following frame the horizontal divergence (the
    level thickness 1/fmz changes along the flow) and the vertical
    divergence of Equation 10's s cancel; what is left is the truncation
    error, below 1 % of either part and falling as the levels are refined.
    A wrong metric factor in Equation 10 (gzitakh for gzitak, the 1/4) or in
    the divergences (fmz, dtrdz) leaves an O(1) residual."""
    worst = []
    for kz in (30, 60):
        n, m = 32, 6
        jx, iy = (n, m) if axis == "x" else (m, n)
        c = dm.case_config(0, jx=jx, iy=iy, kz=kz, hill_h=800.0, hill_hw=4.0)
        o = oracle_lib.oracle_from_host(dm.build_case(c), dm.case_params(c))
        o.call("ora_reset_tendencies")
        o.a["u"][...] = 10.0 if axis == "x" else 0.0
        o.a["v"][...] = 10.0 if axis == "y" else 0.0
        o.call("ora_sound_eq10")
        o.call("ora_sound_hdiv", DTS)
        hdiv = I(o.a["zdiv2"]).copy()
        o.call("ora_sound_vdiv", DTS)
        total = I(o.a["zdiv2"])
        scale = np.abs(hdiv).max(axis=(1, 2))
        assert scale.min() > 0
        ratio = np.abs(total).max(axis=(1, 2)) / scale
        assert ratio.max() < 0.01, (kz, ratio.max())
        worst.append(ratio.max())
    assert worst[1] < 0.7 * worst[0]


# ---------------------------------------------------------- Lamb wave
@pytest.mark.parametrize("axis", ["x", "y"])
def test_lamb_wave_frequency(axis):
    """A wind perturbation u' = eps cos(k (j - 1/2) dx) / pi(z) on a flat,
    dry, isothermal rest state (no divergence damping) is a discrete Lamb
    wave: w stays 0, the Exner perturbation stays uniform in the vertical,
    and the amplitude a_n = u' pi obeys a_{n+1} + a_{n-1} = 2 cos(w dts) a_n
    with 1 - cos(w dts) = 2 (c dts / dx)^2 sin^2(k dx / 2), c^2 = (cpd/cvd)
    Rd T (forward-backward time stepping on the staggered grid).  This fixes
    the divergence's dts/dx (:575-597), the Exner update's Rd/cvd
    (:726-728) and the pressure gradient's cpd theta and dts/dx (:737-787);
    a sign error makes the mode grow."""
    n, m, kz, T0 = 16, 6, 10, 250.0
    jx, iy = (n, m) if axis == "x" else (m, n)
    o, p = _isothermal(jx, iy, kz, T0)
    pbar = o.a["pai"][:, 2, 2].copy()
    kap = 2 * np.pi * 2 / n
    idx = np.arange(-1, n + 3)
    prof = np.cos(kap * (idx - 0.5))
    eps = 1e-4
    fld = "u" if axis == "x" else "v"
    if axis == "x":
        o.a["u"][...] = eps * prof[None, None, :] / pbar[:, None, None]
    else:
        o.a["v"][...] = eps * prof[None, :, None] / pbar[:, None, None]
    pai0 = o.a["pai"].copy()
    mode = np.cos(kap * (np.arange(1, n + 1) - 0.5))
    mode = mode[None, None, :] if axis == "x" else mode[None, :, None]

    def amplitude():
        return 2.0 * np.mean(I(o.a[fld]) * pbar[:, None, None] * mode)

    A = [amplitude()]
    for _ in range(8):
        o.call("ora_substep", 1, DTS)
        A.append(amplitude())
    A = np.array(A)
    cosw = np.sum((A[2:] + A[:-2]) * A[1:-1]) / (2 * np.sum(A[1:-1] ** 2))
    K = oracle_lib.constants()
    c2 = K["cpd"] / K["cvd"] * K["rgas"] * T0
    expect = 2 * c2 * (DTS / p.dx) ** 2 * np.sin(kap / 2) ** 2
    assert (1 - cosw) == pytest.approx(expect, rel=1e-6)
    # the recurrence holds at every substep, not only on average
    assert np.abs(A[2:] + A[:-2] - 2 * cosw * A[1:-1]).max() < 1e-8 * np.abs(A).max()
    dp = I(o.a["pai"] - pai0)
    assert np.abs(dp).max() > 1e-8
    assert np.abs(dp - dp.mean(axis=0)).max() < 1e-6 * np.abs(dp).max()
    assert np.abs(I(o.a["w"])).max() < 1e-6 * eps


# ----------------------------------------------- terrain pressure gradient
def test_terrain_pressure_gradient_balances_rest_state():
    """An isothermal atmosphere at rest over a hill: along a terrain-following
    level the Exner gradient term cpd theta dpi/dx and the terrain term
    g hx gzitakh (:755-757, :781-783) are each about g dts slope, but their
    sum, the physical horizontal gradient, vanishes: u after one substep is
    below 0.2 % of either and falls as (level spacing)^2."""
    res = []
    for kz in (20, 40):
        o, p = _isothermal(32, 6, kz, 250.0, hill_h=800.0, hill_hw=4.0)
        o.a["u"][...] = 0.0
        o.a["v"][...] = 0.0
        o.call("ora_substep", 1, DTS)
        term = 9.80665 * DTS * np.abs(I(o.a["hx"])).max() * np.abs(o.a["gzitakh"][:kz]).max()
        assert term > 0.5
        r = np.abs(I(o.a["u"])).max() / term
        assert r < 2e-3, (kz, r)
        res.append(r)
    assert res[1] < 0.3 * res[0]


# ------------------------------------------------------- the w equation
def _column_state(kz=12, rh=0.5, **kw):
    c = dm.case_config(0, jx=6, iy=5, kz=kz, hill_h=0.0, rh=rh, noise_amp=0.0, **kw)
    o = oracle_lib.oracle_from_host(dm.build_case(c), dm.case_params(c))
    o.call("ora_reset_tendencies")
    o.call("ora_prologue")
    return o, c


def _coefs(o, ns, j=3, i=2):
    kz = o.kz
    arrs = [np.zeros(kz + 1) for _ in range(4)]
    P = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
    o.L.ora_w_coefficients(C.byref(o.S), j, i, ns, DTS, *P)
    return arrs  # zwexpl, zu, zd, zrom1w at index k-1


def test_gravity_wave_term_anticipates_interface_theta():
    """zrom1w / (cpd fmzf) is the interface theta displaced by w over the ns
    substeps elapsed: theta_face - w ns dts dtheta/dz, with dtheta/dz from
    the physical heights zeta of the two levels (not the coordinate's fmzf
    and dzita): the term's sign, its ns factor and its scale."""
    o, c = _column_state(kz=30, rh=0.0)
    kz = c.kz
    j, i = 3, 2
    x = lambda a, k: a[k - 1, i + 1, j + 1]  # noqa: E731
    fz = o.a["fmzf"][1:kz, i + 1, j + 1]
    assert fz.max() > 1.3 and fz.min() < 0.8  # a stretched coordinate: fmzf^2 differs from fmzf
    o.field("qv")[...] = 0.0
    o.a["w"][...] = 0.0
    _, _, _, r0 = _coefs(o, 1, j, i)
    o.a["w"][1:kz, :, :] = 0.7  # interfaces 2..kz
    K = oracle_lib.constants()
    terms = {}
    for ns in (1, 2, 5):
        _, _, _, r = _coefs(o, ns, j, i)
        terms[ns] = r.copy()
        for k in range(2, kz + 1):
            fmzf = x(o.a["fmzf"], k)
            th_u, th_l = x(o.a["tetav"], k - 1), x(o.a["tetav"], k)
            dthdz = (th_u - th_l) / (x(o.a["zeta"], k - 1) - x(o.a["zeta"], k))
            # without motion: the interface mean of theta
            assert r0[k - 1] / (K["cpd"] * fmzf) == pytest.approx(0.5 * (th_u + th_l), rel=1e-14)
            gw = r[k - 1] - r0[k - 1]
            assert gw < 0  # stable layer, upward motion: cooler air arrives
            assert gw / (K["cpd"] * fmzf) == pytest.approx(-0.7 * ns * DTS * dthdz, rel=2e-3)
    np.testing.assert_allclose(terms[5] - r0, 5 * (terms[1] - r0), rtol=1e-10)
    np.testing.assert_allclose(terms[2] - r0, 2 * (terms[1] - r0), rtol=1e-10)


def test_moist_term_is_latent_heating_of_a_lifted_parcel():
    """Where qv > 0.96 qsat and w > 0.1 the interface gets zdth = cpd dtheta:
    a saturated parcel lifted by dz = w (ns-1) dts cools by g dz / cpd,
    condenses dqs/dT per kelvin and warms by L/(cpd pi) per unit
    condensate, dtheta = (L / (cpd pi)) (dqs/dT) g dz / cpd.  The reference
    takes dqs/dT as (qs / es) des/dT with Clausius-Clapeyron for des/dT,
    L es / (Rv T^2); here des/dT is a centred difference of the saturation
    pressure (pinned against Goff-Gratch), which agrees with it within 3 %.
    Checked over warm, unsaturated-cap states with Exner values 0.6 .. 1,
    so L, Rv, T^2, pi, cpd, g and the ns - 1 factor all show."""
    o, c = _column_state(rh=0.0)
    kz = c.kz
    j, i = 3, 2
    K = oracle_lib.constants()
    L = oracle_lib.lib()
    g, Lv, cpd = K["egrav"], K["wlhv"], K["cpd"]
    w0, ns = 0.8, 4
    for T, pai in ((283.0, 1.0), (293.0, 0.8), (288.0, 0.6), (278.0, 0.9)):
        p = 1e5 * pai ** K["cpovr"]
        qs = L.ora_pfwsat(T, p)
        es = L.ora_pfesat(T, p)
        assert es < 0.15 * p
        for name, val in (("t", T), ("pai", pai), ("qsat", qs)):
            o.a[name][:, i + 1, j + 1] = val
        o.a["w"][...] = 0.0
        o.a["w"][1:kz, i + 1, j + 1] = w0
        o.field("qv")[:, i + 1, j + 1] = 0.5 * qs  # dry branch
        _, _, _, dry = _coefs(o, ns, j, i)
        o.field("qv")[:, i + 1, j + 1] = 0.99 * qs  # saturated branch
        _, _, _, wet = _coefs(o, ns, j, i)
        h = 0.01
        dqsdT = qs / es * (L.ora_pfesat(T + h, p) - L.ora_pfesat(T - h, p)) / (2 * h)
        dz = w0 * (ns - 1) * DTS
        dtheta = Lv / (cpd * pai) * dqsdT * g * dz / cpd
        for k in range(2, kz + 1):
            zdth = (wet[k - 1] - dry[k - 1]) / o.a["fmzf"][k - 1, i + 1, j + 1]
            assert zdth / cpd == pytest.approx(dtheta, rel=0.03), (T, pai, k)


def test_moist_term_switches():
    """The moist term acts only where qv > 0.96 qsat and w > 0.1 (:651-652)."""
    o, c = _column_state(rh=0.0)
    kz = c.kz
    j, i = 3, 2
    qs = o.a["qsat"][:, i + 1, j + 1].copy()
    res = {}
    for tag, frac, w in (("on", 0.9602, 0.1002), ("dry", 0.9598, 0.1002), ("slow", 0.9602, 0.0998),
                         ("base", 0.5, 0.1002), ("base_slow", 0.5, 0.0998)):
        o.a["w"][...] = 0.0
        o.a["w"][1:kz, i + 1, j + 1] = w
        o.field("qv")[:, i + 1, j + 1] = frac * qs
        res[tag] = _coefs(o, 3, j, i)[3][1:kz].copy()
    assert np.all(res["on"] > res["base"])
    np.testing.assert_array_equal(res["dry"], res["base"])
    np.testing.assert_array_equal(res["slow"], res["base_slow"])


# ------------------------------------------------------------ fulleq
def test_fulleq_coefficients_from_moist_heat_capacities():
    """zdiv2 *= (1 + a qv + b qc)/(1 + c qv + d qc) turns the dry adiabatic
    Exner tendency into the moist one: the factor is (cp_m/cpd)/(cv_m/cvd)
    per unit moist mass, so a = cpv/cpd - 1, b = cw/cpd - 1,
    c = cvv/cvd - 1, d = cw/cvd - 1 (cvv = cpv - Rv).  tetav *= 1 + (Rd/cvd)
    zdiv2 (e qv + f ql + g qi) carries the moist adiabat's exponent R_m/cp_m
    against the dry Rd/cpd, per unit dry air: e = cpv/cpd - Rv/Rd,
    f = cw/cpd, g = ci/cpd.  With cw = cpw, ci = cpi (mod_constants.F90:147,
    149) and the textbook cpv = 1870 J/(kg K) every coefficient rounds to the
    reference's two significant figures; the coefficients are read back
    from the routine by probing it, so a transposed qv/qc or liquid/ice
    coefficient fails."""
    K = oracle_lib.constants()
    cpd, cvd, Rd, Rv = K["cpd"], K["cvd"], K["rgas"], K["rwat"]
    cpv, cw, ci = 1870.0, 4186.95, 2117.27
    cvv = cpv - Rv
    phys = {"a": cpv / cpd - 1, "b": cw / cpd - 1, "c": cvv / cvd - 1, "d": cw / cvd - 1,
            "e": cpv / cpd - Rv / Rd, "f": cw / cpd, "g": ci / cpd}
    L = oracle_lib.lib()

    def point(qv, qc, ql, qi, z=1e-3, th=300.0):
        zz, tt = C.c_double(z), C.c_double(th)
        L.ora_fulleq_point(qv, qc, ql, qi, C.byref(zz), C.byref(tt))
        return zz.value, tt.value

    def ratio_coefs(which):
        # f(q) = (1 + x q)/(1 + y q) at two q: (x - y f) q = f - 1
        rows, rhs = [], []
        for q in (0.01, 0.03):
            args = (q, 0.0, 0.0, 0.0) if which == "qv" else (0.0, q, 0.0, 0.0)
            f = point(*args)[0] / 1e-3
            rows.append([q, -q * f])
            rhs.append(f - 1)
        return np.linalg.solve(np.array(rows), np.array(rhs))

    a, c = ratio_coefs("qv")
    b, d = ratio_coefs("qc")
    got = {"a": a, "b": b, "c": c, "d": d}
    for name, args in (("e", (0.01, 0.0, 0.0, 0.0)), ("f", (0.0, 0.0, 0.01, 0.0)), ("g", (0.0, 0.0, 0.0, 0.01))):
        z, th = point(*args)
        got[name] = (th / 300.0 - 1) / (K["rdrcv"] * z) / 0.01
    for name, v in phys.items():
        two_figures = float(f"{v:.2g}")
        assert got[name] == pytest.approx(two_figures, rel=1e-6), (name, got[name], v)
