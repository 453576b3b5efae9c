"""Photosynthesis, stomatal conductance and soil-water stress (FP64), restating
``python-port/ibis/physics/stomata.py`` on its F77 path (vmax_mix, coefm_mix, cap_scale,
sunshade, scale_leaf_only and acclimation off).

Pinned by: ``colimit`` against the smaller root of theta x^2 - (j1+j2) x + j1 j2 from the
quadratic formula (textbook; ``stomata.py:198-221``); drystress bounds and root-weighted sum
(``tests/test_stomata.py``); ``_total_conductance`` as a series resistance.  The composed
``stomata`` outputs have only bound/sign checks in the reference
(``tests/test_stomata.py:404-535``): parity unpinned beyond those.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from .thermo import esat, qsat


def colimit(j1, j2, theta):
    """Smaller root of theta x^2 - (j1 + j2) x + j1 j2 = 0 (stomata.py:198-221)."""
    b = j1 + j2
    cc = j1 * j2
    disc = np.maximum(b * b - 4.0 * theta * cc, 0.0)
    q = 0.5 * (b + np.sqrt(disc)) + 1.0e-15
    return np.minimum(q / theta, cc / q)


def farquhar_c3(par, ci, vmax, gamstar, kc, ko, o2, rdark, sp):
    """C3 leaf rates (stomata.py:228-272)."""
    je = par * 4.59e-06 * sp.alpha3 * (ci - gamstar) / (ci + 2.0 * gamstar)
    jc = vmax * (ci - gamstar) / (ci + kc * (1.0 + o2 / ko))
    js = vmax / 2.2
    ag = colimit(colimit(je, jc, sp.theta3), js, sp.beta3)
    return ag, ag - rdark


def farquhar_c4(par, ci, vmax, rdark, sp):
    """C4 leaf rates (stomata.py:279-314)."""
    je = par * 4.59e-06 * sp.alpha4
    jc = vmax
    ji = 18.0e3 * vmax * ci
    ag = colimit(colimit(je, jc, sp.theta4), ji, sp.beta4)
    return ag, ag - rdark


def ball_berry(an, cs0, gs0, ci0, co2, gbco2, rh, coefm, coefb, stress, gsmin, gmin, cimax,
               c4=False):
    """One relaxed step of cs, gs, ci (stomata.py:321-372)."""
    cs = 0.5 * (cs0 + co2 - an / gbco2)
    cs = np.maximum(1.0e-8 if c4 else gmin, cs)
    gs = 0.5 * (gs0 + coefm * an * rh / cs + coefb * stress)
    gs = np.maximum(gs, np.maximum(gsmin, coefb * stress))
    ci = 0.5 * (ci0 + cs - 1.6 * an / gs)
    ci = np.clip(ci, 0.0 if c4 else gmin, cimax)
    return cs, gs, ci


def canopy_scaling(toppar, t6, t7, sc, lai_v, sai_v, a10s, a10d, dtime, eps):
    """PAR-extinction canopy scaling and 10-day means (stomata.py:442-508), F77 denominator."""
    ext = (t6 * sc[:, 0] + t7 * sc[:, 1] - t7 * sc[:, 2]) / np.maximum(sc[:, 3], eps)
    ext = np.clip(ext, 0.1, 10.0)
    pxai = ext * (lai_v + sai_v)
    plai = ext * lai_v
    zw = np.exp(-1.0 / (10.0 * 86400.0 / dtime))
    has = plai > 0.0
    day = toppar > 10.0
    sday = (1.0 - np.exp(-pxai)) / np.maximum(plai, eps)
    upd = has & day
    a10s_n = np.where(upd, zw * a10s + (1.0 - zw) * sday * toppar, a10s)
    a10d_n = np.where(upd, zw * a10d + (1.0 - zw) * toppar, a10d)
    snight = a10s / np.maximum(a10d, eps)
    scale = np.where(has, np.where(day, sday, snight), 0.0)
    return scale, a10s_n, a10d_n


def drystress(wsoi, wisoi, swilt, sfield, froot, stressfac):
    """Soil-water stress (stomata.py:515-562): returns stresstl, stresstu, stressl, stressu."""
    sf = np.asarray(stressfac, dtype=np.float64)
    if sf.ndim == 1:
        sf = sf[:, None]
    znorm = 1.0 - np.exp(sf)
    awc = np.clip((wsoi * (1.0 - wisoi) - swilt) / np.maximum(sfield - swilt, 1e-10), 0.0, 1.0)
    zwilt = np.clip((1.0 - np.exp(sf * awc)) / znorm, 0.0, 1.0)
    stressl = zwilt * froot[None, :, 0]
    stressu = zwilt * froot[None, :, 1]
    # jnp.sum over the layer axis: left-to-right partial sums for six terms
    tl = stressl[:, 0]
    tu = stressu[:, 0]
    for k in range(1, stressl.shape[1]):
        tl = tl + stressl[:, k]
        tu = tu + stressu[:, k]
    return tl, tu, stressl, stressu


def _temp_dep(tleaf, sp):
    """Temperature kinetics (stomata.py:661-689)."""
    rw = 3.47e-03 - 1.0 / tleaf
    tau = sp.tau15 * np.exp(-4500.0 * rw)
    kc = sp.kc15 * np.exp(6000.0 * rw)
    ko = sp.ko15 * np.exp(1500.0 * rw)
    gamstar = sp.o2conc / (2.0 * tau)
    tc = tleaf - 273.16
    base = np.exp(3500.0 * rw)
    cold3 = 1.0 + np.exp(0.40 * (5.0 - tc))
    cold4 = 1.0 + np.exp(0.40 * (10.0 - tc))
    hot = 1.0 + np.exp(0.40 * (tc - 50.0))
    return kc, ko, gamstar, base / (cold3 * hot), base / (cold4 * hot)


def total_conductance(su, gsc):
    """Series leaf + boundary conductance for vapour, F77 operation order (stomata.py:696-725)."""
    dump = 1.0 / 0.029
    return 1.0 / (1.0 / su + dump / gsc)


def vmax_table(sp, n):
    """(n, 12) per-cell Vmax: the (12,) defaults broadcast, or a per-cell table as given."""
    v = np.asarray(sp.vmax_pft, dtype=np.float64)
    return np.broadcast_to(v, (n, 12)) if v.ndim == 1 else v


def stomata(tu, tl, t12, t34, q12, q34, su, sl, psurf, rad, ci_u, gs_u, cs_u, ci_l, gs_l,
            cs_l, fwetu, fwetl, lai, sai, exist, stresstu, stresstl, a10, co2, dtime, sp, eps):
    """Five big leaves (stomata.py:732-1080), F77 switches.  ``a10`` = (spu, dlu, spl, dll)."""
    n = tu.shape[0]
    vp = vmax_table(sp, n)
    gam, cm, cb, gsmin = sp.gamma, sp.coefm, sp.coefb, sp.gsmin

    kc, ko, gstar, tv3, _ = _temp_dep(tu, sp)
    gmin = 1.05 * gstar
    ci_ub = np.clip(ci_u[:, 0], gmin, sp.cimax)
    ci_uc = np.clip(ci_u[:, 1], gmin, sp.cimax)
    gb = np.clip(su * 25.5, 0.1, 10.0)
    rh12 = np.maximum(0.30, q12 / np.maximum(qsat(esat(t12), psurf), eps))
    vub = np.where(exist[:, 0] > 0.5, vp[:, 0], np.where(exist[:, 2] > 0.5, vp[:, 2], vp[:, 4]))
    vuc = np.where(exist[:, 3] > 0.5, vp[:, 3], vp[:, 5])

    vmax = vub * tv3 * stresstu
    rd = gam[0] * vub * tv3
    ag_ub, an_ub = farquhar_c3(rad.topparu, ci_ub, vmax, gstar, kc, ko, sp.o2conc, rd, sp)
    cs_ub, gs_ub, ci_ub = ball_berry(an_ub, cs_u[:, 0], gs_u[:, 0], ci_ub, co2, gb, rh12, cm[0],
                                     cb[0], stresstu, gsmin[0], gmin, sp.cimax)
    vmax = vuc * tv3 * stresstu
    rd = gam[1] * vuc * tv3
    ag_uc, an_uc = farquhar_c3(rad.topparu, ci_uc, vmax, gstar, kc, ko, sp.o2conc, rd, sp)
    cs_uc, gs_uc, ci_uc = ball_berry(an_uc, cs_u[:, 1], gs_u[:, 1], ci_uc, co2, gb, rh12, cm[1],
                                     cb[1], stresstu, gsmin[1], gmin, sp.cimax)
    scale, a10su, a10du = canopy_scaling(rad.topparu, rad.termu6, rad.termu7, rad.scalcoefu,
                                         lai[:, 1], sai[:, 1], a10[0], a10[1], dtime, eps)
    agc_ub, agc_uc = ag_ub * scale, ag_uc * scale
    anc_ub, anc_uc = an_ub * scale, an_uc * scale
    csc = np.maximum(gmin, co2 - anc_ub / gb)
    gsc_ub = np.maximum(np.maximum(gsmin[0], cb[0] * stresstu),
                        cm[0] * anc_ub * rh12 / csc + cb[0] * stresstu)
    csc = np.maximum(gmin, co2 - anc_uc / gb)
    gsc_uc = np.maximum(np.maximum(gsmin[1], cb[1] * stresstu),
                        cm[1] * anc_uc * rh12 / csc + cb[1] * stresstu)
    tc_ub = total_conductance(su, gsc_ub)
    tc_uc = total_conductance(su, gsc_uc)
    dry = 1.0 - fwetu
    agc_ub, agc_uc, anc_ub, anc_uc = agc_ub * dry, agc_uc * dry, anc_ub * dry, anc_uc * dry

    kc, ko, gstar, tv3, tv4 = _temp_dep(tl, sp)
    gmin = 1.05 * gstar
    ci_ls = np.clip(ci_l[:, 0], gmin, sp.cimax)
    ci_l3 = np.clip(ci_l[:, 1], gmin, sp.cimax)
    ci_l4 = np.clip(ci_l[:, 2], 0.0, sp.cimax)
    gb = np.clip(sl * 25.5, 0.1, 10.0)
    rh34 = np.maximum(0.30, q34 / np.maximum(qsat(esat(t34), psurf), eps))
    vls = vp[:, 8]
    vmax = vls * tv3 * stresstl
    rd = gam[2] * vls * tv3
    ag_ls, an_ls = farquhar_c3(rad.topparl, ci_ls, vmax, gstar, kc, ko, sp.o2conc, rd, sp)
    cs_ls, gs_ls, ci_ls = ball_berry(an_ls, cs_l[:, 0], gs_l[:, 0], ci_ls, co2, gb, rh34, cm[2],
                                     cb[2], stresstl, gsmin[2], gmin, sp.cimax)
    vl3 = vp[:, 11]
    vmax = vl3 * tv3 * stresstl
    rd = gam[4] * vl3 * tv3
    ag_l3, an_l3 = farquhar_c3(rad.topparl, ci_l3, vmax, gstar, kc, ko, sp.o2conc, rd, sp)
    cs_l3, gs_l3, ci_l3 = ball_berry(an_l3, cs_l[:, 1], gs_l[:, 1], ci_l3, co2, gb, rh34, cm[4],
                                     cb[4], stresstl, gsmin[4], gmin, sp.cimax)
    vl4 = vp[:, 10]
    vmax = vl4 * tv4 * stresstl
    rd = gam[3] * vl4 * tv4
    ag_l4, an_l4 = farquhar_c4(rad.topparl, ci_l4, vmax, rd, sp)
    cs_l4, gs_l4, ci_l4 = ball_berry(an_l4, cs_l[:, 2], gs_l[:, 2], ci_l4, co2, gb, rh34, cm[3],
                                     cb[3], stresstl, gsmin[3], gmin, sp.cimax, c4=True)
    scale, a10sl, a10dl = canopy_scaling(rad.topparl, rad.terml6, rad.terml7, rad.scalcoefl,
                                         lai[:, 0], sai[:, 0], a10[2], a10[3], dtime, eps)
    agc_ls, agc_l3, agc_l4 = ag_ls * scale, ag_l3 * scale, ag_l4 * scale
    anc_ls, anc_l3, anc_l4 = an_ls * scale, an_l3 * scale, an_l4 * scale
    csc_ls = np.maximum(gmin, co2 - anc_ls / gb)
    csc_l3 = np.maximum(gmin, co2 - anc_l3 / gb)
    csc_l4 = np.maximum(1.0e-8, co2 - anc_l4 / gb)
    gsc_ls = np.maximum(np.maximum(gsmin[2], cb[2] * stresstl),
                        cm[2] * anc_ls * rh34 / csc_ls + cb[2] * stresstl)
    gsc_l3 = np.maximum(np.maximum(gsmin[4], cb[4] * stresstl),
                        cm[4] * anc_l3 * rh34 / csc_l3 + cb[4] * stresstl)
    gsc_l4 = np.maximum(np.maximum(gsmin[3], cb[3] * stresstl),
                        cm[3] * anc_l4 * rh34 / csc_l4 + cb[3] * stresstl)
    tc_ls = total_conductance(sl, gsc_ls)
    tc_l3 = total_conductance(sl, gsc_l3)
    tc_l4 = total_conductance(sl, gsc_l4)
    dry = 1.0 - fwetl
    agc_ls, agc_l3, agc_l4 = agc_ls * dry, agc_l3 * dry, agc_l4 * dry
    anc_ls, anc_l3, anc_l4 = anc_ls * dry, anc_l3 * dry, anc_l4 * dry
    st = np.stack
    return SimpleNamespace(
        ci_upper=st([ci_ub, ci_uc], 1), gs_upper=st([gs_ub, gs_uc], 1),
        cs_upper=st([cs_ub, cs_uc], 1), an_upper=st([anc_ub, anc_uc], 1),
        ag_upper=st([agc_ub, agc_uc], 1),
        ci_lower=st([ci_ls, ci_l3, ci_l4], 1), gs_lower=st([gs_ls, gs_l3, gs_l4], 1),
        cs_lower=st([cs_ls, cs_l3, cs_l4], 1), an_lower=st([anc_ls, anc_l3, anc_l4], 1),
        ag_lower=st([agc_ls, agc_l3, agc_l4], 1),
        totcondub=tc_ub, totconduc=tc_uc, totcondls=tc_ls, totcondl3=tc_l3, totcondl4=tc_l4,
        a10=(a10su, a10du, a10sl, a10dl))
