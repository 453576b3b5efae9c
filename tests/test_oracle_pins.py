"""Pins of the oracle: known answers, closed forms and conservation laws.

Each check is independent of the restatement it pins (a dropped term, a wrong sign or a
transposed operand in the oracle fails one of them).  Citations are to the reference's own
tests where it has one.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from oracle import canopy, interception, params as OP, radiation, snow, soil, solvers, stomata
from oracle import thermo, fire
from oracle._num import ipow

C = OP.physical_constants()


# ---- thermodynamics (tests/test_thermo.py) -------------------------------------------
def test_esat_known_answer_at_freezing():
    # liquid base constant a0 = 6.1078 hPa at 273.16 K exactly (thermo.py:88)
    assert thermo.esat(np.array([273.16]))[0] == pytest.approx(610.78, rel=1e-14)
    # ice branch at 273.15 uses b0 = 6.1091780 and ti = -0.01
    e = thermo.esat(np.array([273.15]))[0]
    ti = -0.01
    poly = ti * (5.0346990e-1 + ti * (1.8860134e-2 + ti * (4.1762237e-4 + ti * (5.8247203e-6 + ti * (4.8388032e-8 + ti * 1.8388269e-10)))))
    assert e == pytest.approx(100.0 * (6.1091780 + poly), rel=1e-14)


def test_esat_monotonic_and_desat_is_its_derivative():
    t = np.linspace(230.0, 320.0, 901)
    e = thermo.esat(t)
    assert np.all(np.diff(e) > 0)
    h = 1e-4
    for tt in (250.0, 265.0, 280.0, 300.0, 315.0):
        num = (thermo.esat(np.array([tt + h])) - thermo.esat(np.array([tt - h]))) / (2 * h)
        assert thermo.desat(np.array([tt]))[0] == pytest.approx(num[0], rel=1e-2)


def test_qsat_closed_form_and_dqsat():
    e = np.array([1000.0, 3000.0])
    p = np.array([101325.0, 90000.0])
    np.testing.assert_allclose(thermo.qsat(e, p), 0.622 * e / (p - 0.378 * e), rtol=1e-15)
    t = np.array([283.0])
    q = lambda tt: thermo.qsat(thermo.esat(tt), np.array([101325.0]))
    num = (q(t + 1e-4) - q(t - 1e-4)) / 2e-4
    assert thermo.dqsat(t, q(t))[0] == pytest.approx(num[0], rel=1e-2)


def test_latent_heats_at_melt():
    # hvapf(tmelt, tmelt) = hvap (tests/test_f77_verification.py:85-90)
    t = np.array([C.tmelt])
    assert thermo.hvapf(t, t, C)[0] == C.hvap
    assert thermo.hsubf(t, t, C)[0] == C.hsub


def test_integer_powers_follow_square_and_multiply():
    x = np.array([0.3, 1.7, 0.0])
    np.testing.assert_allclose(ipow(x, 13), x ** 13, rtol=1e-14)
    assert ipow(np.array([2.0]), 0)[0] == 1.0
    np.testing.assert_array_equal(ipow(x, 3), x * (x * x))


# ---- solvers (tests/test_solvers.py) -------------------------------------------------
def test_tridia_known_3x3():
    a = np.array([[0.0, -1.0, -1.0]])
    b = np.array([[2.0, 2.0, 2.0]])
    c = np.array([[-1.0, -1.0, 0.0]])
    r = np.array([[1.0, 0.0, 1.0]])
    np.testing.assert_allclose(solvers.tridia(a, b, c, r), [[1.0, 1.0, 1.0]], rtol=1e-15)


def test_tridia_against_banded_solver():
    rng = np.random.default_rng(0)
    n, k = 20, 9
    a, c = 0.3 * rng.standard_normal((n, k)), 0.3 * rng.standard_normal((n, k))
    b = 3.0 + rng.random((n, k))
    r = rng.standard_normal((n, k))
    x = solvers.tridia(a, b, c, r)
    for i in range(n):
        ab = np.zeros((3, k))
        ab[0, 1:] = c[i, :-1]
        ab[1] = b[i]
        ab[2, :-1] = a[i, 1:]
        np.testing.assert_allclose(x[i], solve_banded((1, 1), ab, r[i]), rtol=1e-12)


def test_nopivot_gauss_against_lapack_and_closed_form():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((30, 9, 9)) + 12.0 * np.eye(9)[None]
    b = rng.standard_normal((30, 9))
    np.testing.assert_allclose(solvers.gauss_nopivot(A, b), np.linalg.solve(A, b[..., None])[..., 0],
                               rtol=1e-10, atol=1e-12)
    A2 = np.array([[[4.0, 1.0], [2.0, 3.0]]])
    x = solvers.gauss_nopivot(A2, np.array([[1.0, 2.0]]))
    np.testing.assert_allclose(x, [[(3 - 2) / 10.0, (8 - 2) / 10.0]], rtol=1e-15)


# ---- stomata ---------------------------------------------------------------------------
def test_colimit_is_the_smaller_quadratic_root():
    rng = np.random.default_rng(2)
    j1, j2 = rng.random(100) * 1e-5, rng.random(100) * 1e-5
    for th in (0.95, 0.8):
        b, c = j1 + j2, j1 * j2
        root = (b - np.sqrt(b * b - 4 * th * c)) / (2 * th)
        np.testing.assert_allclose(stomata.colimit(j1, j2, th), root, rtol=1e-6)


def test_drystress_root_weighting_and_bounds():
    froot = OP.root_fractions()
    np.testing.assert_allclose(froot.sum(0), [1.0, 1.0], rtol=1e-14)
    n = 5
    wsoi = np.full((n, 6), 0.9)
    tl, tu, sl, su = stomata.drystress(wsoi, np.zeros((n, 6)), np.full((n, 6), 0.2),
                                       np.full((n, 6), 0.6), froot, -5.0)
    np.testing.assert_allclose(tl, 1.0, rtol=1e-14)     # wetter than field capacity: unstressed
    dry = stomata.drystress(np.full((n, 6), 0.1), np.zeros((n, 6)), np.full((n, 6), 0.2),
                            np.full((n, 6), 0.6), froot, -5.0)
    np.testing.assert_allclose(dry[0], 0.0, atol=1e-15)


def test_total_conductance_is_a_series_resistance():
    su, g = np.array([0.05]), np.array([0.2])
    assert stomata.total_conductance(su, g)[0] == pytest.approx(1 / (1 / su[0] + 1 / (0.029 * g[0])), rel=1e-13)


# ---- radiation (tests/test_radiation.py) -----------------------------------------------
def _rad_inputs(n=50, seed=3, coszen=None):
    rng = np.random.default_rng(seed)
    cz = rng.uniform(0.05, 1.0, n) if coszen is None else np.full(n, coszen)
    lai = rng.uniform(0.0, 6.0, (n, 2))
    sai = rng.uniform(0.05, 1.0, (n, 2))
    fw = interception.fwetcal(*(rng.uniform(250, 300, n) for _ in range(3)),
                              *(rng.uniform(0, 0.1, n) for _ in range(6)), OP.interception_params(), C)
    temps = [rng.uniform(250, 310, n) for _ in range(5)]
    return dict(coszen=cz, solad=rng.uniform(0, 400, (n, 2)), solai=rng.uniform(0, 200, (n, 2)),
                fira=rng.uniform(200, 450, n), lai=lai, sai=sai, fu=rng.uniform(0, 1, n),
                fl=rng.uniform(0, 1, n), tu=temps[0], ts=temps[1], tl=temps[2], tg=temps[3],
                ti=np.minimum(temps[4], 273.16), fi=rng.uniform(0, 1, n),
                tsno1=rng.uniform(250, 273.16, n), wsoi1=rng.uniform(0, 1, n),
                wisoi1=rng.uniform(0, 0.5, n), albsav=rng.uniform(0.05, 0.15, n),
                albsan=rng.uniform(0.1, 0.3, n), fw=fw)


def test_solar_energy_closes():
    d = _rad_inputs()
    r = radiation.compute_radiation(**d, vr=OP.veg_rad_params(), c=C)
    incoming = d["solad"].sum(1) + d["solai"].sum(1)
    reflected = (r.asurd * d["solad"]).sum(1) + (r.asuri * d["solai"]).sum(1)
    # solu/sols are per unit area under the upper canopy, soll per unit area of lower
    # canopy, solg per unit snow-free ground and soli per unit snow
    absorbed = (d["fu"] * (r.solu + r.sols) + r.soll * d["fl"] * (1 - d["fi"])
                + (1 - d["fi"]) * r.solg + d["fi"] * r.soli)
    np.testing.assert_allclose(absorbed + reflected, incoming, rtol=1e-10)


def test_infrared_energy_closes():
    d = _rad_inputs()
    f = radiation.irrad(d["fira"], d["lai"], d["sai"], d["fu"], d["fl"], d["fi"], d["tu"], d["ts"],
                        d["tl"], d["tg"], d["ti"], C.stef)
    firu, firs, firl, firg, firi, firb = f
    fu, fl, fi = d["fu"], d["fl"], d["fi"]
    absorbed = fu * (firu + firs) + (1 - fi) * (fl * firl + firg) + fi * firi
    np.testing.assert_allclose(absorbed, d["fira"] - firb, rtol=1e-10, atol=1e-9)


def test_night_has_no_solar():
    d = _rad_inputs(coszen=0.0)
    r = radiation.compute_radiation(**d, vr=OP.veg_rad_params(), c=C)
    for k in ("solu", "sols", "soll", "solg", "soli"):
        assert np.all(getattr(r, k) == 0.0)


def test_bare_surface_reflects_with_ground_albedo():
    d = _rad_inputs()
    d["lai"][:] = 0.0
    d["sai"][:] = 0.0
    d["fu"][:] = 0.0
    d["fl"][:] = 0.0
    r = radiation.compute_radiation(**d, vr=OP.veg_rad_params(), c=C)
    for ib in range(2):
        sod, soi, snd, sni = radiation.solsur(d["coszen"], d["wsoi1"], d["wisoi1"], d["tsno1"],
                                              d["albsav"], d["albsan"], C.tmelt, d["coszen"] > 0, ib)
        np.testing.assert_allclose(r.asurd[:, ib], (1 - d["fi"]) * sod + d["fi"] * snd, rtol=1e-12)


# ---- interception ---------------------------------------------------------------------
def test_mix_is_the_mass_weighted_mean():
    x, t = np.array([1.0, 0.0]), np.array([280.0, 0.0])
    xm, tm = interception.mix(x, t, 3 * x, t + 4.0, 0 * x, t, C.epsilon)
    np.testing.assert_allclose(xm, [4.0, 0.0])
    np.testing.assert_allclose(tm, [283.0, 0.0])


def test_cascade_conserves_water():
    rng = np.random.default_rng(4)
    n = 64
    ip = OP.interception_params()
    lai, sai = rng.uniform(0, 5, (n, 2)), rng.uniform(0, 1, (n, 2))
    fu, fl = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    w0 = [rng.uniform(0, 0.15, n) for _ in range(3)] + [rng.uniform(0, 1.0, n) for _ in range(3)]
    t = [rng.uniform(260, 290, n) for _ in range(5)]
    raina, snowa = rng.uniform(0, 2e-3, n), rng.uniform(0, 1e-3, n)
    r = interception.cascade(*t, *w0, lai, sai, fu, fl, rng.uniform(260, 300, n),
                             rng.uniform(0.001, 0.01, n), np.full(n, 1e5), raina, snowa, ip, 3600.0, C)
    xu, xs, xl = 2 * lai[:, 1], 2 * sai[:, 1], 2 * (lai[:, 0] + sai[:, 0])
    d_up = fu * (xu * ((r.wliqu - w0[0]) + (r.wsnou - w0[3])) + xs * ((r.wliqs - w0[1]) + (r.wsnos - w0[4])))
    d_lo = fl * xl * ((r.wliql - w0[2]) + (r.wsnol - w0[5]))
    np.testing.assert_allclose((raina + snowa) * 3600.0, d_up + d_lo + (r.raing + r.snowg) * 3600.0,
                               rtol=1e-9, atol=1e-12)


def test_wet_bulb_solves_its_balance_when_converged():
    ta, q, p = np.array([300.0, 290.0]), np.array([0.01, 0.005]), np.array([1e5, 1e5])
    tw = interception.twet3(ta, q, p, C)
    resid = (ta - C.tmelt) - (thermo.hvapf(tw, ta, C) / C.cair) * (thermo.qsat(thermo.esat(tw), p) - q) \
        - (tw - C.tmelt)
    assert np.all(np.abs(resid) < 0.02 / 0.2 * 1.5)
    assert np.all(tw < ta)


def test_noveg_sets_surface_temperature():
    n = 3
    z = np.zeros((n, 2))
    out = interception.noveg(*(np.full(n, 290.0) for _ in range(3)), *(np.ones(n) for _ in range(6)),
                             z, z, np.zeros(n), np.zeros(n), np.array([0.0, 0.5, 1.0]),
                             np.full(n, 280.0), np.full(n, 270.0))
    np.testing.assert_allclose(out[0], [280.0, 275.0, 270.0])
    assert np.all(out[3] == 0.0) and np.all(out[8] == 0.0)


# ---- canopy system ----------------------------------------------------------------------
def test_impexp_weights_bounded():
    t = np.array([270.0, 273.16, 280.0])
    w = canopy.impexp(t, np.full(3, 1e4), np.full(3, 0.1), np.full(3, 0.1), C, np.ones(3), False)
    assert np.all((w >= 0) & (w <= 1))
    w2 = canopy.impexp2(np.array([275.0, 260.0]), np.array([270.0, 270.0]), C, np.ones(2), False)
    np.testing.assert_allclose(w2, [(C.tmelt - 270.0) / 5.0, 1.0])


# ---- snow (tests/test_snow.py) ----------------------------------------------------------
def test_vadapt_conserves_thickness_and_heat():
    rng = np.random.default_rng(5)
    h = rng.uniform(0.0, 0.5, (40, 3))
    t = rng.uniform(250, 273.16, (40, 3))
    hn, tn = snow.vadapt(h, t, 0.05, h.sum(1) > 1e-7, 1e-7)
    np.testing.assert_allclose(hn.sum(1), h.sum(1), rtol=1e-12)
    np.testing.assert_allclose((hn * tn).sum(1), (h * t).sum(1), rtol=1e-12)
    thick = h.sum(1) >= 0.15
    np.testing.assert_allclose(hn[thick, 0], 0.05)


def test_snowheat_isothermal_quiescent_and_bottom_flux():
    n = 4
    t = np.full((n, 3), 265.0)
    tn, tl, sflo = snow.snowheat(t, np.full(n, 265.0), np.zeros(n), np.full(n, 1.0), 8436.0,
                                 np.full(n, 0.5), np.full((n, 3), 0.1), 250.0, 0.2, 0.1,
                                 np.full(n, 1.0), np.full(n, 265.0), 3600.0, C.cice)
    np.testing.assert_allclose(tn, 265.0, rtol=1e-14)
    np.testing.assert_allclose(sflo, 0.0, atol=1e-9)
    # heated from above: the bottom flux equals the implicit conduction into the soil
    xl, hs1, k1, ts1 = 1.0, 0.1, 1.0, 260.0
    tn, tl, sflo = snow.snowheat(t, np.full(n, 265.0), np.full(n, 50.0), np.full(n, xl), 8436.0,
                                 np.full(n, 0.5), np.full((n, 3), 0.1), 250.0, 0.2, hs1,
                                 np.full(n, k1), np.full(n, ts1), 3600.0, C.cice)
    con = 1.0 / (xl * (0.5 / 2.0) + 0.5 * hs1 / k1)
    np.testing.assert_allclose(sflo[:, -1], con * (tl - ts1), rtol=1e-10)


def test_snow_melt_energy_and_cap():
    h = np.array([[0.1, 0.1, 0.1]])
    t = np.array([[275.0, 273.0, 273.16 + 0.5]])
    hn, tn, zm, zh = snow.snow_melt(h, t, np.array([1.0]), 250.0, C.cice, C.hfus, C.tmelt)
    xh = 250.0 * h * C.cice * np.maximum(t - C.tmelt, 0.0)
    assert zm[0] == pytest.approx(xh.sum() / C.hfus, rel=1e-12)
    assert np.all(tn <= C.tmelt) and zh[0] == pytest.approx(0.0, abs=1e-9)


def test_snow_coverage_conserves_volume():
    rng = np.random.default_rng(6)
    h = rng.uniform(0.01, 0.4, (30, 3))
    fi = rng.uniform(0.05, 1.0, 30)
    fn, hn = snow.snow_coverage(h, fi, 0.15, np.full(30, 0.5), 1.0, 1e-7)
    ok = fn < 1.0
    np.testing.assert_allclose((fn * hn.sum(1))[ok], (fi * h.sum(1))[ok], rtol=1e-12)


# ---- soil (tests/test_soil.py, tests/test_bedrock.py, tests/test_f77_verification.py) ---
def _soil_inputs(n=8, seed=7):
    rng = np.random.default_rng(seed)
    poros = rng.uniform(0.35, 0.5, (n, 6))
    return dict(poros=poros, wsoi=rng.uniform(0.1, 0.9, (n, 6)), wisoi=rng.uniform(0, 0.3, (n, 6)),
                tsoi=rng.uniform(265, 295, (n, 6)), csoi=np.full((n, 6), 880.5),
                rhosoi=np.full((n, 6), 2636.5), hsoi=OP.HSOI_LAYERS)


def test_setsoi_fractions_sum_to_one_and_conductivity_formula():
    rng = np.random.default_rng(8)
    n = 40
    d = _soil_inputs(n)
    wpud = np.where(rng.random(n) < 0.5, rng.uniform(0, 5, n), 0.0)
    wipud = np.where(rng.random(n) < 0.3, rng.uniform(0, 2, n), 0.0)
    sand, clay = np.full((n, 6), 40.0), np.full((n, 6), 20.0)
    r = soil.setsoi(sand, clay, d["poros"], d["wsoi"], d["wisoi"], d["tsoi"], wpud, wipud, 10.0, 0.5,
                    np.full(n, 270.0), np.full(n, 290.0), C)
    np.testing.assert_allclose(r.qglif.sum(1), 1.0, rtol=1e-12)
    zc = 0.4 * 0.300 + 0.4 * 0.265 + 0.2 * 0.250
    exp = zc * 56.0 ** (d["poros"] * d["wsoi"] * (1 - d["wisoi"])) * 224.0 ** (d["poros"] * d["wisoi"])
    np.testing.assert_allclose(r.consoi, exp, rtol=1e-12)


def test_wadjust_clamps_and_reports_mass():
    w = np.array([[1e-9, 0.5, 1.2, 0.3, 0.2, 0.1]])
    wi = np.array([[0.0, 0.0, 0.0, 1.0, 0.2, 1e-9]])
    p = np.full((1, 6), 0.4)
    wn, win, g = soil.wadjust(w, wi, p, OP.HSOI_LAYERS, 1000.0, 3600.0, 1e-7)
    np.testing.assert_array_equal(wn[0], [0.0, 0.5, 1.0, 0.0, 0.2, 0.1])
    np.testing.assert_array_equal(win[0, 5], 0.0)
    t0 = (OP.HSOI_LAYERS * 0.4 * 1000 * ((1 - wi) * w + wi)).sum()
    t1 = (OP.HSOI_LAYERS * 0.4 * 1000 * ((1 - win) * wn + win)).sum()
    assert g[0] * 3600.0 == pytest.approx(t1 - t0, rel=1e-12)


def test_soilheat_with_bedrock_conserves_energy_in_float64():
    # bedrock.py: d(stored heat) = fhtop dt with no water flow (tests/test_bedrock.py:198-229)
    d = _soil_inputs(6)
    n = 6
    fh = np.linspace(-80, 120, n)
    bp = OP.bedrock_params()
    deep = np.full((n, 3), 279.0)
    z = np.zeros((n, 6))
    ts, hflo, dn = soil.soilheat_with_bedrock(d["tsoi"], deep, d["tsoi"], deep, d["wsoi"], d["wsoi"],
                                             d["wisoi"], d["poros"], np.full((n, 6), 1.2), d["csoi"],
                                             d["rhosoi"], OP.HSOI_LAYERS, z, z, fh, np.zeros((n, 7)),
                                             C.ch2o, C.cice, C.rhow, 3600.0, bp)
    cv = ((1 - d["poros"]) * d["csoi"] * d["rhosoi"] + d["poros"] * (1 - d["wisoi"]) * C.ch2o * C.rhow * d["wsoi"]
          + d["poros"] * d["wisoi"] * C.cice * C.rhow)
    dE = (cv * OP.HSOI_LAYERS * (ts - d["tsoi"])).sum(1) + (2.0e6 * np.array(bp.hsoi) * (dn - deep)).sum(1)
    np.testing.assert_allclose(dE, fh * 3600.0, rtol=1e-10, atol=1e-3)


def test_isothermal_column_without_flux_stays_put():
    d = _soil_inputs(3)
    t = np.full((3, 6), 280.0)
    z = np.zeros((3, 6))
    tn, hflo = soil.soilheat(t, t, d["wsoi"], d["wsoi"], d["wisoi"], d["poros"], np.ones((3, 6)),
                             d["csoi"], d["rhosoi"], OP.HSOI_LAYERS, z, z, np.zeros(3),
                             np.zeros((3, 7)), C.ch2o, C.cice, C.rhow, 3600.0)
    np.testing.assert_allclose(tn, 280.0, rtol=1e-14)


def _enthalpy(t, w, wi, d, c1pud):
    zp = np.maximum(d["poros"], C.epsilon)
    rw = c1pud / (C.rhow * d["hsoi"]) + (1 - zp) * d["csoi"] * d["rhosoi"] / C.rhow
    ch = rw + zp * (1 - wi) * w * C.ch2o + zp * wi * C.cice
    return ch * (t - C.tmelt) - C.hfus * zp * wi


def test_freeze_thaw_conserves_enthalpy():
    d = _soil_inputs(10)
    t = np.linspace(268, 278, 10)[:, None] * np.ones((1, 6))
    z = np.zeros((10, 6))
    tn, wn, win, _ = soil.freeze_thaw(t, d["wsoi"], d["wisoi"], d["poros"], d["csoi"], d["rhosoi"],
                                      d["hsoi"], z, C)
    # liquid mass (pore fraction) must balance ice change
    liq0 = d["poros"] * (1 - d["wisoi"]) * d["wsoi"]
    liq1 = d["poros"] * (1 - win) * wn
    np.testing.assert_allclose(liq0 + d["poros"] * d["wisoi"], liq1 + d["poros"] * win, rtol=1e-12)
    np.testing.assert_allclose(_enthalpy(tn, wn, win, d, z), _enthalpy(t, d["wsoi"], d["wisoi"], d, z),
                               rtol=1e-9, atol=1e-6)


def test_puddle_freeze_thaw_conserves_mass_and_never_goes_negative():
    t0 = np.array([270.0, 276.0, 273.16])
    wp, wip = np.array([2.0, 0.0, 1.0]), np.array([0.0, 3.0, 1.0])
    t1, wp1, wip1 = soil.puddle_freeze_thaw(t0, wp, wip, np.full(3, 0.4), np.full(3, 0.3), np.zeros(3),
                                           np.full(3, 880.5), np.full(3, 2636.5), 0.1, C)
    np.testing.assert_allclose(wp1 + wip1, wp + wip, rtol=1e-14)
    assert np.all(wp1 >= 0) and np.all(wip1 >= 0)


def test_soilh2o_closes_the_column_mass_balance():
    rng = np.random.default_rng(9)
    n = 20
    d = _soil_inputs(n)
    wisoi = np.zeros((n, 6))
    poros = d["poros"]
    por = poros * np.maximum(1 - wisoi, 1e-7)
    fwtop = -rng.uniform(0, 1e-5, n)
    fwpud = rng.uniform(0, 2e-4, n)
    r = soil.soilh2o(d["wsoi"], wisoi, poros, por, np.full((n, 6), 2e-6), np.full((n, 6), 0.2),
                     np.full((n, 6), 5.0), OP.HSOI_LAYERS, fwtop, fwpud, np.full(n, 1.0), 0.04, 3600.0,
                     1000.0, 1e-7)
    m0 = (1000 * poros * OP.HSOI_LAYERS * d["wsoi"]).sum(1)
    m1 = (1000 * poros * OP.HSOI_LAYERS * r.wsoi).sum(1)
    # storage change = top inflow - bottom drainage - squeeze returned to the puddle
    np.testing.assert_allclose(m1 - m0, (fwtop + fwpud - r.wflo[:, -1] - r.fsqueez) * 3600.0,
                               rtol=1e-8, atol=1e-9)


# ---- fire ---------------------------------------------------------------------------------
def test_fire_running_mean_starts_at_the_sample():
    w, wi, sw = np.array([0.3]), np.array([0.1]), np.array([0.2])
    fd = fire.fueldry(w, wi, sw)
    assert fire.update_firefac(np.array([0.7]), 1.0, w, wi, sw)[0] == fd[0]
    assert 0.0 <= fd[0] <= 1.0
