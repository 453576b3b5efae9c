"""Pins of the oracle's optional StomataParams switches (``oracle/stomata.py``).

Each check restates a property the reference's own tests hold for the switch
(``tests/test_sunshade_canopy.py``, ``test_canopy_capacity.py``, ``test_vmax_mix.py``,
``test_pft_slopes_acclim.py``, ``test_cold_season.py``) or a closed form the mathematics
fixes, chosen so that a dropped term, a swapped exponent or a wrong index fails one of them.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import params as OP
from oracle import radiation as RAD
from oracle import stomata as ST

SP = OP.stomata_params()

# the reference's sunshade cases (tests/test_sunshade_canopy.py:39-40)
CASES = dict(coszen=[0.95, 0.6, 0.3, 0.15, 0.8], lai=[4.7, 1.0, 8.0, 3.0, 0.3],
             sai=[0.85, 0.2, 1.5, 0.5, 0.1], solad=[200.0, 300.0, 50.0, 100.0, 0.0],
             solai=[100.0, 50.0, 150.0, 100.0, 120.0])


def _two_stream(coszen, lai, sai, solad, solai, asur=0.1):
    """Upper-canopy VIS two-stream with IBIS leaf optics: scaling coefficients, unscattered beam
    term, exponents and the absorbed flux per unit ground."""
    z, L, S = (np.asarray(v, dtype=np.float64) for v in (coszen, lai, sai))
    zero = np.zeros_like(z)
    t = np.full_like(z, 293.0)
    sunlit = z > 0.0
    om, bd, bi, av, gd = RAD.twoset(z, L, S, zero, zero, t, 0.10, 0.05, 0.0, 0.0, 0.9, 273.16,
                                    True, sunlit, fwet_stem=zero, rliq_stem=zero, t_stem=t)
    a = np.full_like(z, asur)
    ts = RAD.twostr(z, L, S, a, a, om, bd, bi, av, gd, 1.0e-7, sunlit, True)
    sd, si = np.asarray(solad, dtype=np.float64), np.asarray(solai, dtype=np.float64)
    A = ts.term[0] * sd
    B = ts.term[1] * sd + ts.term[2] * si
    C = ts.term[3] * sd + ts.term[4] * si
    sc = np.stack([A, B, C, A + B + C], 1)
    k, h = ts.term[5], ts.term[6]
    dbeam = (1.0 - ts.omega) * k * sd
    return sc, dbeam, k, h, L, S, sd * ts.abvegd + si * ts.abvegi


def test_e1_is_the_exponential_integral_and_its_small_argument_limit():
    plai = np.array([0.5, 3.0, 8.0, 2.0, 2.0])
    a = np.array([0.7, -0.3, 1e-9, 1.0e-6, 2.0])
    x = np.linspace(0.0, 1.0, 200001)[None, :] * plai[:, None]
    quad = np.trapezoid(np.exp(-a[:, None] * x), x, axis=1)
    np.testing.assert_allclose(ST.e1(a, plai), quad, rtol=1e-9)


def test_two_stream_profile_integrates_to_the_absorbed_flux():
    sc, dbeam, k, h, L, S, absorbed = _two_stream(**CASES)
    sun, sha, *_ = ST.sunshade_canopy(sc, dbeam, k, h, L, S, 0.15, 1.0e-7)
    layer = (sun + sha) * (L + S) / L                      # leaves' share -> whole layer
    np.testing.assert_allclose(layer, absorbed, rtol=2e-3, atol=1e-3)


def test_partition_conserves_light_leaf_area_and_capacity():
    sc, dbeam, k, h, L, S, _ = _two_stream(**CASES)
    kn = 0.15
    sun, sha, lsun, lsha, vsun, vsha = ST.sunshade_canopy(sc, dbeam, k, h, L, S, kn, 1.0e-7)
    for v in (sun, sha, lsun, lsha, vsun, vsha):
        assert (v >= 0).all()
    np.testing.assert_allclose(lsun + lsha, L, rtol=1e-12)
    np.testing.assert_allclose(vsun + vsha, (1.0 - np.exp(-kn * L)) / kn, rtol=1e-12)
    assert (sun / lsun >= sha / lsha * 0.999).all()
    assert (vsun / lsun >= vsha / lsha).all()
    beam = np.asarray(CASES["solad"]) > 0
    assert ((sun / lsun)[beam] > 1.5 * (sha / lsha)[beam]).all()


def test_uniform_light_and_capacity_reduce_to_one_big_leaf():
    L = np.array([1.0, 4.7, 8.0])
    S = np.array([0.1, 0.85, 1.5])
    B = np.array([30.0, 20.0, 10.0])
    z = np.zeros(3)
    sc = np.stack([z, B, z, B], 1)
    k, h = np.full(3, 0.6), np.full(3, 1.0e-7)
    ss = ST.sunshade_canopy(sc, z, k, h, L, S, 0.0, 1.0e-7)
    ci, vmax = np.full(3, 280e-6), np.full(3, 45e-6)
    gam, kc, ko = np.full(3, 40e-6), np.full(3, 300e-6), np.full(3, 0.3)
    rd = 0.015 * vmax
    agc, anc = ST.sunshade_c3(ss, ci, vmax, gam, kc, ko, SP.o2conc, rd, L, SP)
    fl = L / (L + S)
    ag1, an1 = ST.farquhar_c3(B * (L + S) * fl, ci, vmax * L, gam, kc, ko, SP.o2conc, rd * L, SP)
    np.testing.assert_allclose(agc, ag1 / L, rtol=1e-6)
    np.testing.assert_allclose(anc, an1 / L, rtol=1e-6)


def test_capacity_integral_limits():
    L = np.array([0.0, 1e-6, 0.5, 5.0, 50.0])
    f = ST.capacity_integral(L, 0.15)
    assert f[0] == 0.0 and abs(f[1] - 1.0) < 1e-6 and (np.diff(f[1:]) < 0).all()
    np.testing.assert_allclose(f[3:], (1 - np.exp(-0.15 * L[3:])) / (0.15 * L[3:]), rtol=1e-14)


def _c3(par):
    n = len(par)
    return dict(par=np.asarray(par, dtype=np.float64), ci=np.full(n, 250e-6),
                vmax=np.full(n, 37e-6), gamstar=np.full(n, 40e-6), kc=np.full(n, 230e-6),
                ko=np.full(n, 0.28), o2=0.209, rdark=np.full(n, 0.5e-6))


def test_equal_light_and_capacity_profiles_are_the_f77_rate_times_the_profile():
    x = _c3(np.linspace(0.0, 600.0, 25))
    f = np.full(25, 0.245)
    ag0, an0 = ST.farquhar_c3(**x, sp=SP)
    ag1, an1 = ST.farquhar_c3(**x, sp=SP, f_light=f, f_cap=f)
    # exact up to colimit's 1e-15 guard, which does not scale with the rates
    np.testing.assert_allclose(ag1, ag0 * f, rtol=1e-8, atol=1e-20)
    np.testing.assert_allclose(an1, an0 * f, rtol=1e-8, atol=1e-20)
    g0, _ = ST.farquhar_c4(x["par"], x["ci"], x["vmax"], x["rdark"], SP)
    g1, _ = ST.farquhar_c4(x["par"], x["ci"], x["vmax"], x["rdark"], SP, f_light=f, f_cap=f)
    np.testing.assert_allclose(g1, g0 * f, rtol=1e-8, atol=1e-20)


def test_shallow_capacity_profile_lifts_the_plateau_not_the_low_light_rate():
    x = _c3([20.0, 400.0])
    fl, fc = np.full(2, 0.245), np.full(2, 0.63)
    ag0, _ = ST.farquhar_c3(**x, sp=SP)
    ag1, _ = ST.farquhar_c3(**x, sp=SP, f_light=fl, f_cap=fc)
    r = ag1 / (ag0 * fl)
    assert r[0] < 1.10 and r[1] > 1.8


def test_light_rate_factor_scales_the_light_limited_rate_only():
    par, ci, vmax = np.array([50.0, 50.0]), np.full(2, 2.5e-4), np.full(2, 6e-5)
    g, kc, ko, rd = np.full(2, 4e-5), np.full(2, 3e-4), np.full(2, 0.3), np.full(2, 1e-6)
    ag0, _ = ST.farquhar_c3(par, ci, vmax, g, kc, ko, 0.209, rd, SP)
    ag1, _ = ST.farquhar_c3(par, ci, vmax, g, kc, ko, 0.209, rd, SP, je_fac=np.ones(2))
    ag2, _ = ST.farquhar_c3(par, ci, vmax, g, kc, ko, 0.209, rd, SP, je_fac=np.array([0.2, 1.0]))
    np.testing.assert_array_equal(ag0, ag1)
    assert ag2[0] < 0.5 * ag0[0] and ag2[1] == ag0[1]


VP = np.broadcast_to(np.arange(1.0, 13.0) * 1e-6, (1, 12))


def test_vmax_mix_off_is_the_priority_pick_and_every_entry_is_reachable_on():
    rng = np.random.default_rng(0)
    exist = (rng.random((50, 12)) > 0.5).astype(float)
    frac = rng.random((50, 12))
    vp = np.broadcast_to(VP, (50, 12))
    ub, uc, ls, l3, l4 = ST.select_vmax(exist, vp, frac, mix=False)
    np.testing.assert_array_equal(ub, np.where(exist[:, 0] > 0.5, vp[:, 0],
                                               np.where(exist[:, 2] > 0.5, vp[:, 2], vp[:, 4])))
    np.testing.assert_array_equal(uc, np.where(exist[:, 3] > 0.5, vp[:, 3], vp[:, 5]))
    assert (ls == vp[:, 8]).all() and (l3 == vp[:, 11]).all() and (l4 == vp[:, 10]).all()
    for j in range(10):
        f1 = np.zeros((1, 12))
        f1[0, j] = 1.0
        got = ST.select_vmax(np.ones((1, 12)), VP, f1, mix=True)
        leaf = 0 if j in ST.BROADLEAF else 1 if j in ST.CONIFER else 2
        assert abs(got[leaf][0] - VP[0, j]) < 1e-18


def test_vmax_mix_is_the_lai_weighted_mean_and_falls_back_without_leaves():
    frac = np.zeros((2, 12))
    frac[0, 0], frac[0, 1] = 0.25, 0.75
    vp = np.broadcast_to(VP, (2, 12))
    ub, uc, *_ = ST.select_vmax(np.ones((2, 12)), vp, frac, mix=True)
    assert abs(ub[0] - (0.25 * VP[0, 0] + 0.75 * VP[0, 1])) < 1e-18
    assert ub[1] == VP[0, 0] and uc[0] == VP[0, 3]


def test_capacity_shares_keep_the_leaf_area_total():
    rng = np.random.default_rng(3)
    frac = rng.random((20, 12))
    frac[0] = 0.0
    vp = rng.random((20, 12)) * 1e-5
    sh = ST.capacity_shares(frac, vp, ST.BROADLEAF)
    np.testing.assert_allclose(sh.sum(1), frac[:, list(ST.BROADLEAF)].sum(1), rtol=1e-12)
    w = frac[1, list(ST.BROADLEAF)] * vp[1, list(ST.BROADLEAF)]
    np.testing.assert_allclose(sh[1] / sh[1].sum(), w / w.sum(), rtol=1e-12)
    np.testing.assert_array_equal(sh[0], 0.0)


def test_acclimation_factor_per_type_and_scalar_with_frost():
    sp = SimpleNamespace(acclim_on=True, acclim_pfts=(3, 5, 8), acclim_per_pft=False,
                         acclim_s0=-3.063, acclim_smax=17.72, frost_on=False)
    x = np.array([-10.0, 0.0, 5.0, 30.0])
    v = ST.acclim_factor(x, None, sp, 4)
    fs = np.clip((x + 3.063) / 17.72, 0.0, 1.0)
    for j in range(12):
        np.testing.assert_array_equal(v[:, j], fs if j in (3, 5, 8) else 1.0)
    sp.acclim_per_pft = True
    sp.acclim_s0_pft = np.full(12, -3.063)
    sp.acclim_smax_pft = np.full(12, 17.72)
    sp.acclim_s0_pft[5] = -6.6
    v2 = ST.acclim_factor(x, None, sp, 4)
    np.testing.assert_array_equal(v2[:, 3], v[:, 3])
    np.testing.assert_array_equal(v2[:, 5], np.clip((x + 6.6) / 17.72, 0.0, 1.0))
    sp.frost_on, sp.frost_t0, sp.frost_t1 = True, -5.0, 5.0
    td = 273.16 + np.array([-8.0, 0.0, 2.5, 10.0])
    v3 = ST.acclim_factor(x, td, sp, 4)
    ff = np.array([0.0, 0.5, 0.75, 1.0])
    np.testing.assert_allclose(v3[:, 5], v2[:, 5] * ff, rtol=1e-15)
    np.testing.assert_array_equal(v3[:, 0], 1.0)
    sp.acclim_on = False
    assert ST.acclim_factor(x, td, sp, 4) is None


def _stomata_inputs(n=2, lai=3.0, sai=0.5):
    t = np.full(n, 298.0)
    z = np.zeros(n)
    L, S = np.full(n, lai), np.full(n, sai)
    scal = np.tile(np.array([30.0, 40.0, 5.0, 75.0]), (n, 1))
    rad = SimpleNamespace(topparu=np.full(n, 60.0), topparl=np.full(n, 20.0),
                          termu6=np.full(n, 0.6), termu7=np.full(n, 0.9),
                          terml6=np.full(n, 0.6), terml7=np.full(n, 0.9), scalcoefu=scal,
                          scalcoefl=0.3 * scal, dbeamu=np.full(n, 50.0), dbeaml=np.full(n, 15.0))
    exist = np.zeros((n, 12))
    exist[:, [4, 3, 8, 10, 11]] = 1.0
    a10 = (np.full(n, 0.5), np.full(n, 100.0), np.full(n, 0.5), np.full(n, 100.0))
    args = dict(tu=t, tl=t, t12=t, t34=t, q12=np.full(n, 0.012), q34=np.full(n, 0.012),
                su=np.full(n, 0.05), sl=np.full(n, 0.03), psurf=np.full(n, 101325.0), rad=rad,
                ci_u=np.full((n, 2), 280e-6), gs_u=np.full((n, 2), 0.05),
                cs_u=np.full((n, 2), 340e-6), ci_l=np.full((n, 3), 280e-6),
                gs_l=np.full((n, 3), 0.05), cs_l=np.full((n, 3), 340e-6), fwetu=z, fwetl=z,
                lai=np.stack([L, L], 1), sai=np.stack([S, S], 1), exist=exist,
                stresstu=np.ones(n), stresstl=np.ones(n), a10=a10, co2=380e-6, dtime=3600.0,
                eps=1.0e-7)
    return args, exist.copy()


def _sp(**kw):
    sp = OP.stomata_params()
    sp.coefm_pft = np.array([10.0, 10.0, 10.0, 6.0, 10.0, 6.0, 10.0, 6.0, 9.0, 9.0, 4.0, 9.0])
    sp.cap_kn = 0.15
    for k, v in kw.items():
        setattr(sp, k, v)
    return sp


def test_one_type_per_leaf_slopes_are_the_group_values_and_the_c4_entry_moves_only_c4():
    x, frac = _stomata_inputs()
    off = ST.stomata(**x, sp=_sp(), frac=frac)
    on = ST.stomata(**x, sp=_sp(coefm_mix=True), frac=frac)
    for k in ("gs_upper", "gs_lower", "ci_upper", "ci_lower", "totcondub", "totcondl4"):
        np.testing.assert_array_equal(getattr(on, k), getattr(off, k), err_msg=k)
    cp = _sp().coefm_pft.copy()
    cp[10] = 12.0
    c4 = ST.stomata(**x, sp=_sp(coefm_mix=True, coefm_pft=cp), frac=frac)
    assert not np.array_equal(c4.gs_lower[:, 2], off.gs_lower[:, 2])
    np.testing.assert_array_equal(c4.gs_lower[:, :2], off.gs_lower[:, :2])
    np.testing.assert_array_equal(c4.gs_upper, off.gs_upper)


def test_scale_leaf_only_is_the_f77_scale_times_l_over_l_plus_s():
    rng = np.random.default_rng(1)
    n = 40
    toppar = np.full(n, 400.0)
    t6, t7 = rng.uniform(0.4, 1.5, n), rng.uniform(0.3, 1.0, n)
    sc = rng.uniform(0.1, 1.0, (n, 4))
    lai = rng.uniform(0.2, 4.0, n)
    sai = 0.25 * lai
    one = np.ones(n)
    s0, _, _ = ST.canopy_scaling(toppar, t6, t7, sc, lai, sai, one, one, 3600.0, 1e-7)
    s1, _, _ = ST.canopy_scaling(toppar, t6, t7, sc, lai, sai, one, one, 3600.0, 1e-7,
                                 leaf_only=True)
    np.testing.assert_allclose(s1, s0 * lai / (lai + sai), rtol=1e-13)


def test_sunshade_changes_the_canopy_only_when_on():
    sc, dbeam, k, h, L, S, _ = _two_stream(coszen=[0.8, 0.8], lai=[4.7, 0.0], sai=[0.85, 0.85],
                                           solad=[200.0, 200.0], solai=[100.0, 100.0])
    x, frac = _stomata_inputs()
    t = np.full(2, 298.0)
    x["rad"] = SimpleNamespace(topparu=sc[:, 3] * L / (L + S), topparl=0.3 * sc[:, 3] * L / (L + S),
                               termu6=k, termu7=h, terml6=k, terml7=h, scalcoefu=sc,
                               scalcoefl=0.3 * sc, dbeamu=dbeam, dbeaml=0.3 * dbeam)
    x["lai"], x["sai"] = np.stack([L, L], 1), np.stack([S, S], 1)
    x["exist"] = np.zeros((2, 12))
    x["exist"][:, [4, 5, 9]] = 1.0
    x["tu"] = x["tl"] = x["t12"] = x["t34"] = t
    off = ST.stomata(**x, sp=_sp(), frac=frac)
    on = ST.stomata(**x, sp=_sp(sunshade_u=True, sunshade_l=True), frac=frac)
    assert on.ag_upper[0, 0] > 1.3 * off.ag_upper[0, 0]
    assert on.ag_upper[1, 0] == 0.0 and on.an_upper[1, 0] == 0.0
    assert (on.ag_lower[0] > 0.0).all()
    with pytest.raises(ValueError):
        del x["rad"].dbeamu
        ST.stomata(**x, sp=_sp(sunshade_u=True), frac=frac)


def test_cap_scale_with_kn_equal_to_the_light_profile_is_the_f77_canopy():
    """With Kn chosen so that the capacity integral equals the light scaling factor, the
    canopy co-limitation is the F77 leaf rate times the factor (homogeneity of degree one)."""
    x, frac = _stomata_inputs(n=1, lai=2.0, sai=0.0)
    rad = x["rad"]
    rad.scalcoefu = np.array([[0.0, 40.0, 0.0, 40.0]])
    rad.termu6 = rad.termu7 = np.array([1.0])             # extinction k = 1: f = (1 - e^-L)/L
    off = ST.stomata(**x, sp=_sp(), frac=frac)
    on = ST.stomata(**x, sp=_sp(cap_scale_u=True, cap_kn=1.0), frac=frac)
    np.testing.assert_allclose(on.ag_upper, off.ag_upper, rtol=1e-8)
    np.testing.assert_allclose(on.gs_upper, off.gs_upper, rtol=1e-8)
    np.testing.assert_array_equal(on.ag_lower, off.ag_lower)


def test_cropland_factor_scales_only_the_grasses():
    crop = SimpleNamespace(c4_share=-1.0, vmax_scale_c3=0.33, vmax_scale_c4=1.16)
    cf = np.array([0.0, 0.5, 1.0])
    g = ST.crop_factor(None, cf, crop, 3)
    np.testing.assert_array_equal(g[:, :10], 1.0)
    np.testing.assert_allclose(g[:, 10], 1.0 + cf * 0.16, rtol=1e-15)
    np.testing.assert_allclose(g[:, 11], 1.0 - cf * 0.67, rtol=1e-15)
    crop.c4_share = 0.2                              # both grasses take the mix of the scales
    g2 = ST.crop_factor(None, cf, crop, 3)
    mix = 0.8 * 0.33 + 0.2 * 1.16
    np.testing.assert_allclose(g2[:, 10], 1.0 + cf * (mix - 1.0), rtol=1e-15)
    np.testing.assert_array_equal(g2[:, 10], g2[:, 11])
    v = np.full((3, 12), 0.5)
    np.testing.assert_allclose(ST.crop_factor(v, cf, crop, 3), 0.5 * g2, rtol=1e-15)
    assert ST.crop_factor(v, None, crop, 3) is v
