"""Pins of the daily-step oracle (oracle/daily.py) by what the mathematics and the
reference's own tests fix (tests/test_biogeochem.py:259-405 and tests/test_f77_verification.py:
693-735 of the reference)."""
import numpy as np
import pytest

from oracle import daily as OD

EPS = 1.0e-7
HSOI = np.array([0.10, 0.15, 0.25, 0.50, 1.00, 2.00, 4.00])


def veg_case(n, seed=1):
    rng = np.random.default_rng(seed)
    v = dict(tempu=rng.uniform(0, 1, n), templ=rng.uniform(0, 1, n),
             agddu=rng.uniform(0, 300, n), agddl=rng.uniform(0, 300, n),
             dropu=rng.uniform(0.1, 1, n), dropls=rng.uniform(0.1, 1, n),
             dropl3=rng.uniform(0.1, 1, n), dropl4=rng.uniform(0.1, 1, n),
             plai=rng.uniform(0, 6, (n, 12)), tc=rng.uniform(-30, 20, n),
             frac=rng.uniform(0, 1, (n, 12)), lai=rng.uniform(0, 6, (n, 2)),
             ztop=np.stack([rng.uniform(0.5, 1.5, n), rng.uniform(10, 30, n)], 1),
             zbot=np.stack([rng.uniform(0.05, 0.3, n), rng.uniform(3, 8, n)], 1),
             fu=rng.uniform(0.2, 0.95, n), fl=rng.uniform(0.2, 0.95, n))
    a10 = dict(a10td=rng.uniform(250, 305, n), a10ancub=rng.uniform(-2e-6, 8e-6, n),
               a10ancls=rng.uniform(-2e-6, 8e-6, n), a10ancl3=rng.uniform(-2e-6, 8e-6, n),
               a10ancl4=rng.uniform(-2e-6, 8e-6, n))
    return v, a10


def carbon_case(n, seed=2):
    rng = np.random.default_rng(seed)
    c = {k: rng.uniform(0.01, 2.0, n) for k in OD.CARBON}
    c["csoipas"] = rng.uniform(1, 20, n)
    c["csoislon"] = rng.uniform(0.5, 5, n)
    c["csoislop"] = rng.uniform(0.5, 5, n)
    c["totcmic"] = rng.uniform(0.01, 0.3, n)
    c["decompl"], c["decomps"] = rng.uniform(0.001, 3, n), rng.uniform(0.001, 3, n)
    sand = rng.uniform(10, 90, (n, 6))
    clay = rng.uniform(5, 50, (n, 6))
    summ = dict(storedn=np.zeros(n), yrleach=np.zeros(n), ynleach=np.zeros(n))
    return c, sand, clay, summ


# ---- phenology --------------------------------------------------------------------------------

def test_fractions_of_each_canopy_sum_to_one_and_lai_is_capped():
    v, a10 = veg_case(500)
    v["plai"][:10] = 50.0                                  # reference: lai capped at 12
    o = OD.pheno(v, a10, np.full(500, 290.0), EPS, OD.PhenoParams(), np.ones(500), np.ones(500))
    up, lo = o["frac"][:, :8].sum(1), o["frac"][:, 8:].sum(1)
    np.testing.assert_allclose(up, 1.0, rtol=1e-14)
    np.testing.assert_allclose(lo, 1.0, rtol=1e-14)
    assert (o["lai"] <= 12.0).all() and (o["lai"][:10] == 12.0).any()
    assert (o["ztop"][:, 0] >= 0.25).all() or (o["ztop"][:, 0] <= v["zbot"][:, 1] - 0.5).all()
    assert (o["zbot"][:, 0] == 0.05).all()


def test_cold_lowers_tempu_warmth_raises_agddu():
    v, a10 = veg_case(200)
    cold = dict(a10, a10td=np.full(200, 260.0))
    o = OD.pheno(v, cold, np.full(200, 260.0), EPS, OD.PhenoParams(), np.ones(200), np.ones(200))
    assert (o["tempu"] <= v["tempu"]).all()
    np.testing.assert_allclose(o["tempu"], np.maximum(0.0, v["tempu"] - 1.0 / 15.0))
    warm = dict(a10, a10td=np.full(200, 285.0))
    o = OD.pheno(v, warm, np.full(200, 300.0), EPS, OD.PhenoParams(), np.ones(200), np.ones(200))
    np.testing.assert_allclose(o["agddu"], v["agddu"] + 300.0 - 273.16)
    assert (o["agddu"] >= v["agddu"]).all()


def test_negative_assimilation_drops_leaves():
    v, a10 = veg_case(200)
    v["dropu"][:] = 1.0
    v["dropls"][:] = 1.0
    dry = dict(a10, a10ancub=np.full(200, -1e-6), a10ancls=np.full(200, -1e-6))
    o = OD.pheno(v, dry, np.full(200, 290.0), EPS, OD.PhenoParams(), np.ones(200), np.ones(200))
    assert (o["dropu"] < v["dropu"]).all() and (o["dropls"] < v["dropls"]).all()
    np.testing.assert_allclose(o["dropu"], 1.0 - 1.0 / 15.0)
    # the floor holds
    v["dropu"][:] = 0.11
    o = OD.pheno(v, dry, np.full(200, 290.0), EPS, OD.PhenoParams(), np.ones(200), np.ones(200))
    assert (o["dropu"] == 0.1).all()


# ---- decomposition factors --------------------------------------------------------------------

def test_climate_factors_at_their_reference_points():
    """Lloyd-Taylor is exactly 1 at 288.16 K; the moisture factor is exp(-(w-60)^2/800) below
    60 % WFPS and the quadratic above, so 0.9776 at 60 % exactly."""
    n = 3
    t = np.full((n, 6), 288.16)
    w = np.tile(np.array([0.60, 0.30, 0.80])[:, None], (1, 6))
    dl, ds = OD.climate_factors(t, w, np.zeros((n, 6)), HSOI, OD.BGCParams())
    np.testing.assert_allclose(dl[0], 0.000371 * 3600 - 0.0748 * 60 + 4.13, rtol=1e-14)
    np.testing.assert_allclose(dl[1], np.exp(-(30.0 - 60.0) ** 2 / 800.0), rtol=1e-14)
    wf = 80.0
    np.testing.assert_allclose(dl[2], 0.000371 * wf * wf - 0.0748 * wf + 4.13, rtol=1e-14)
    np.testing.assert_array_equal(dl, ds)                  # uniform columns: same mean
    # ice takes pore space: wfps = (1 - wisoi) wsoi
    dl2, _ = OD.climate_factors(t, np.full((n, 6), 0.8), np.full((n, 6), 0.5), HSOI, OD.BGCParams())
    np.testing.assert_allclose(dl2, np.exp(-(40.0 - 60.0) ** 2 / 800.0), rtol=1e-14)
    # warm soil, capped at 10; very cold soil uses T0 + 10 K
    hot, _ = OD.climate_factors(np.full((1, 6), 400.0), np.full((1, 6), 0.3), np.zeros((1, 6)),
                                HSOI, OD.BGCParams())
    lt = min(np.exp(344.0 * (1 / (288.16 - 227.13) - 1 / (400.0 - 227.13))), 10.0)
    np.testing.assert_allclose(hot, lt * np.exp(-(30.0 - 60.0) ** 2 / 800.0), rtol=1e-14)


def test_litter_splits_are_verberne_and_sum_to_one():
    b = OD.BGCParams()
    (fll, fls, flm), (frl, frs, frm), (fwl, fws, fwm) = OD.litter_splits(b)
    np.testing.assert_allclose(fll, 0.45 * 1600.0 / (1200.0 + 1600.0), rtol=1e-15)
    np.testing.assert_allclose(fwl, 0.45 * 40000.0 / (1200.0 + 40000.0), rtol=1e-15)
    for a, s, m in ((fll, fls, flm), (frl, frs, frm), (fwl, fws, fwm)):
        assert abs(a + s + m - 1.0) < 1e-15
    # the split reproduces the litter's C:N: lignin at cnr[4], structural at the structural
    # pool's, metabolic at cnr[6]
    cnr = b.cnr
    np.testing.assert_allclose(fll / cnr[4] + fls / cnr[5] + flm / cnr[6], 1.0 / b.cnleaf, rtol=1e-13)
    np.testing.assert_allclose(fwl / cnr[4] + fws / cnr[7] + fwm / cnr[6], 1.0 / b.cnwood, rtol=1e-13)


# ---- soil biogeochemistry ---------------------------------------------------------------------

def test_soilbgc_carbon_balance_defines_tco2mic():
    n = 400
    c, sand, clay, summ = carbon_case(n)
    b = OD.BGCParams()
    out, tco2mic, tnmin, s = OD.soilbgc(c, sand, clay, HSOI, b, 365, summ)
    pools = lambda x: sum(x[k] for k in OD.CARBON[:12]) + x["totcmic"]
    cin = (c["falll"] + c["fallr"] + c["fallw"]) / 365.0
    cleach = s["yrleach"] - summ["yrleach"]
    resp = tco2mic * 86400.0 * 0.012
    np.testing.assert_allclose(pools(c) + cin - pools(out) - cleach, resp, rtol=1e-9, atol=1e-14)
    for k in OD.CARBON[:12] + ["totcmic"]:
        assert (out[k] >= 0).all(), k                      # reference: pools stay non-negative
    assert (tnmin >= 0).all() and np.isfinite(tco2mic).all()
    np.testing.assert_allclose(out["totlit"], sum(out[k] for k in OD.CARBON[:9]), rtol=1e-14)
    np.testing.assert_allclose(s["storedn"], tnmin)


def test_soilbgc_reference_conservation_case_and_thirty_days():
    """The reference's F77 verification case: litter 0.1 / 0.05 / 0.08 / 0.04, sand 50 %,
    clay 20 %, everything else zero; 30 days stay finite and non-negative."""
    n = 2
    c = {k: np.zeros(n) for k in OD.CARBON}
    c["clitlm"][:], c["clitls"][:], c["clitrm"][:], c["clitrs"][:] = 0.1, 0.05, 0.08, 0.04
    c["decompl"][:], c["decomps"][:] = 1.0, 1.0
    sand, clay = np.full((n, 6), 50.0), np.full((n, 6), 20.0)
    summ = dict(storedn=np.zeros(n), yrleach=np.zeros(n), ynleach=np.zeros(n))
    for _ in range(30):
        c, tco2mic, tnmin, summ = OD.soilbgc(c, sand, clay, HSOI, OD.BGCParams(), 365, summ)
        assert all((c[k] >= 0).all() for k in OD.CARBON)
    assert np.isfinite(c["csoipas"]).all() and np.isfinite(c["totcmic"]).all()
    assert (c["totcmic"] > 0).all()                       # microbes grew on the litter


def test_nsoilcall_two_is_two_days_of_soilbgc():
    n = 50
    v, a10 = veg_case(n)
    c, sand, clay, summ = carbon_case(n)
    rng = np.random.default_rng(9)
    day = dict(c, **a10, **summ, ayprcp=np.zeros(n), ayaet=np.zeros(n), ayco2mic=np.zeros(n),
               ayco2root=np.zeros(n), tco2root=rng.uniform(0, 1e-6, n))
    sums = {k: rng.uniform(0, 5, n) for k in ("adrain", "adsnow", "adaet")}
    sums.update({k: rng.uniform(-0.05, 0.5, n) for k in ("a10ancub", "a10ancls", "a10ancl3", "a10ancl4")})
    tsoi = rng.uniform(260, 300, (n, 6))
    wsoi, wisoi = rng.uniform(0.1, 0.9, (n, 6)), rng.uniform(0, 0.3, (n, 6))
    td = rng.uniform(260, 300, n)
    out2, t2, _ = OD.day_end(day, sums, td, tsoi, wsoi, wisoi, sand, clay, HSOI, OD.BGCParams(),
                             nsoilcall=2)
    dl, ds = OD.climate_factors(tsoi, wsoi, wisoi, HSOI, OD.BGCParams())
    cc = dict(c, decompl=dl, decomps=ds)
    s = dict(summ)
    for _ in range(2):
        cc, tt, _, s = OD.soilbgc(cc, sand, clay, HSOI, OD.BGCParams(), 365, s)
    for k in OD.CARBON:
        np.testing.assert_array_equal(out2[k], cc[k], err_msg=k)
    np.testing.assert_array_equal(t2, tt)
    # the 10-day means: exp(-1/10) weight, day sums per second
    z = np.exp(-0.1)
    np.testing.assert_allclose(out2["a10td"], z * a10["a10td"] + (1 - z) * td, rtol=1e-15)
    np.testing.assert_allclose(out2["a10ancub"], z * a10["a10ancub"] + (1 - z) * sums["a10ancub"] / 86400.0,
                               rtol=1e-15)
    np.testing.assert_allclose(out2["ayprcp"], sums["adrain"] + sums["adsnow"], rtol=1e-15)


# ---- seasonal states (state/seasonal.py; tests/test_seasonal_formulations.py) -------------

def test_daylength_is_astronomically_right():
    eq = OD.daylength_hours(np.zeros(365), np.arange(1, 366))
    assert np.all(np.abs(eq - 12.0) < 0.05)
    n45 = OD.daylength_hours(np.deg2rad(45.0), 172)
    s45 = OD.daylength_hours(np.deg2rad(45.0), 355)
    assert 15.2 < n45 < 15.6 and 8.4 < s45 < 8.8
    assert abs(OD.daylength_hours(np.deg2rad(-45.0), 172) - s45) < 0.1
    assert OD.daylength_hours(np.deg2rad(80.0), 172) == 24.0
    assert OD.daylength_hours(np.deg2rad(80.0), 355) == 0.0


def test_acclimation_recursion_matches_preles():
    """S = S + (T - S) / tau, f_S = clip((S - S0) / Smax, 0, 1) (PRELES fS_model)."""
    pp = OD.PhenoParams()
    sea = dict(acclim_x=np.zeros(1), ssen_u=np.zeros(1), ssen_l=np.zeros(1))
    for t in np.r_[np.full(40, -15.0), np.full(60, 20.0)]:
        sea, _, _ = OD.seasonal_update(sea, np.array([t + 273.16]), np.zeros(1), 100, pp, 10.93)
    assert sea["acclim_x"][0] > 19.0 and (sea["acclim_x"][0] + 3.063) / 17.72 >= 1.0
    x = 0.0
    for t in np.full(60, -15.0):
        x = x + (t - x) / 10.93
    assert min(max((x + 3.063) / 17.72, 0.0), 1.0) == 0.0


def test_senescence_sums_accumulate_only_while_days_shorten_below_pstart():
    pp = OD.PhenoParams(sen_photoperiod_u=True, sen_ycrit_u=30.0)
    lat = np.deg2rad(np.array([50.0, 50.0, -50.0]))
    td = np.full(3, 273.16 + 5.0)                 # 15 degree-days below Tb = 20 C per day
    sea = dict(acclim_x=np.zeros(3), ssen_u=np.array([0.0, 20.0, 20.0]), ssen_l=np.zeros(3))
    new, on_u, on_l = OD.seasonal_update(sea, td, lat, 260, pp, 10.93)   # September
    np.testing.assert_allclose(new["ssen_u"][:2], [15.0, 35.0])
    assert new["ssen_u"][2] == 0.0                # southern spring: days lengthen, reset
    assert list(on_u) == [False, True, False] and on_l is None
    spring, _, _ = OD.seasonal_update(sea, td, lat, 100, pp, 10.93)
    assert spring["ssen_u"][0] == 0.0 and spring["ssen_u"][1] == 0.0


def test_senescence_trigger_forces_leaf_fall_and_holds_degree_days():
    n = 3
    v, a10 = veg_case(n, 2)
    v["tempu"], v["agddu"] = np.ones(n), np.full(n, 900.0)
    a10["a10td"] = np.full(n, 288.0)
    td = np.full(n, 288.0)
    pp = OD.PhenoParams()
    s = np.full(n, 1.0)
    a = OD.pheno(v, a10, td, 1e-7, pp, s, s)
    b = OD.pheno(v, a10, td, 1e-7, pp, s, s, sen_on_u=np.zeros(n, bool), sen_on_l=np.zeros(n, bool))
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    c = OD.pheno(v, a10, td, 1e-7, pp, s, s, sen_on_u=np.array([True, False, True]))
    np.testing.assert_allclose(c["tempu"], [1.0 - 1.0 / 15.0, a["tempu"][1], 1.0 - 1.0 / 15.0])
    assert c["agddu"][0] == 0.0 and c["agddu"][2] == 0.0 and c["agddu"][1] > 0
