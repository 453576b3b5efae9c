"""Oracle of the daily step (test infrastructure only: imported by tests, never by the product).

NumPy FP64 restatement of the once-a-day work of the reference's production driver around
its hourly steps (python-port/ibis/timestepping/loops_gswp3.py:382-583):

  day begin   the 10-day means are kept, the daily sums (adrain adsnow adaet a10anc*) reset
              (:403-416), leaf phenology (physics/phenology.py:94-287) from the 10-day means
              and the day's mean temperature;
  day end     10-day means from the day's sums with the daily weight exp(-1/10)
              (:470-477), annual precipitation and AET (:480-481), the decomposition
              climate factors (Lloyd-Taylor temperature x water-filled-pore-space moisture,
              :494-553), soil biogeochemistry nsoilcall times (physics/biogeochem.py:23-278),
              heterotrophic and root respiration added to the annual sums (:577-585).

Configuration restated: the default one (scalar lt_e0, wfps_opt and k_scale; SOM sensitivity
off, so the slow and passive pools take the litter factor; no per-type habits, no stress
accumulator), and the optional seasonal states with photoperiod senescence
(``seasonal_update``; state/seasonal.py, loops_gswp3.py:418-440).

Pins (tests/test_oracle_daily.py): closed forms of the litter splits and of the climate
factors at their reference points; the carbon balance that defines tco2mic; the reference's
own qualitative checks (tests/test_biogeochem.py:259-405, tests/test_f77_verification.py:
693-735): cold lowers tempu, warmth raises agddu, negative assimilation drops leaves, LAI
capped at 12, fractions of each canopy sum to one, pools stay non-negative, nitrogen
mineralisation non-negative, 30 days stay finite; and the brute-force equality of one call
with nsoilcall = 2 against two calls.  The seasonal states by the reference's
tests/test_seasonal_formulations.py: day length 12 h at the equinox and the polar limits,
the acclimation state's first-order relaxation to a constant temperature, senescence sums
accumulating only while days shorten below Pstart and resetting otherwise, the trigger
forcing leaf fall.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

TMELT = 273.16
#: CarbonState fields in dataclass order (python-port/ibis/state/carbon.py)
CARBON = ("clitlm clitls clitll clitrm clitrs clitrl clitwm clitws clitwl csoipas csoislon "
          "csoislop decompl decomps falll fallr fallw cdisturb totcmic totlit totalit totrlit "
          "totcsoi totfall").split()


@dataclass
class PhenoParams:
    """physics/phenology.py:20-85 (the thresholds that are not seasonal-state switches)."""
    tthresh_u_offset: float = 5.0
    gthresh_u_c: float = 0.0
    gdd_thresh_u: float = 100.0
    gdd_ramp_u: float = 50.0
    tthresh_l_c: float = 0.0
    gthresh_l_c: float = -5.0
    gdd_thresh_l: float = 150.0
    gdd_ramp_l: float = 50.0
    drop_thresh: float = 0.0
    drop_floor: float = 0.1
    ddays: float = 15.0
    drop_stress_thresh: float = -1.0
    tthresh_u_abs: float = -100.0
    # photoperiod-gated senescence (physics/phenology.py:79-86; state/seasonal.py:16-24)
    sen_photoperiod_u: bool = False
    sen_pstart_u: float = 14.0
    sen_tb_u: float = 20.0
    sen_ycrit_u: float = 300.0
    sen_photoperiod_l: bool = False
    sen_pstart_l: float = 14.0
    sen_tb_l: float = 20.0
    sen_ycrit_l: float = 300.0


@dataclass
class BGCParams:
    """io/parameters.py:176-250."""
    lig_frac: float = 0.50
    fbsom: float = 0.017
    effac: float = 0.40
    fmax: float = 0.45
    rconst: float = 1200.0
    cnleaf: float = 40.0
    cnroot: float = 60.0
    cnwood: float = 200.0
    cnr: tuple = (8.0, 15.0, 10.0, 15.0, 100.0, 150.0, 6.0, 250.0)
    klm: float = 0.15
    kls: float = 0.01
    kll: float = 0.01
    krm: float = 0.10
    krs: float = 0.005
    krl: float = 0.005
    kwm: float = 0.001
    kws: float = 0.001
    kwl: float = 0.001
    kbn: float = 0.045
    kbp: float = 0.005
    knb: float = 0.001
    kns: float = 1.0e-6
    kpb: float = 0.0001
    kps: float = 1.0e-6
    ksb: float = 8.0e-7
    ylm: float = 0.4
    yls: float = 0.3
    yll: float = 1.0
    yrm: float = 0.4
    yrs: float = 0.3
    yrl: float = 1.0
    ywm: float = 0.4
    yws: float = 0.3
    ywl: float = 1.0
    ybn: float = 1.0
    ybp: float = 1.0
    yps: float = 1.0
    yns: float = 1.0
    ysb: float = 0.20
    ypb: float = 0.20
    ynb: float = 0.25
    lt_e0: float = 344.0
    wfps_opt: float = 60.0
    k_scale: float = 1.0


def daylength_hours(lat_rad, doy):
    """Astronomical day length (h): declination 0.40910518 sin(2 pi (doy - 81) / 365) and the
    sunrise hour angle, clipped for polar day and night (state/seasonal.py:53-58)."""
    decl = 0.40910518 * np.sin(2.0 * np.pi * (doy - 81.0) / 365.0)
    x = np.clip(-np.tan(lat_rad) * np.tan(decl), -1.0, 1.0)
    return 24.0 / np.pi * np.arccos(x)


def seasonal_update(sea, td, lat_rad, jday, pp: PhenoParams, acclim_tau):
    """Advance the seasonal states once a day from the day's mean temperature
    (loops_gswp3.py:418-440): the acclimation state S += (T - S) / tau and the senescence sums,
    which accumulate max(Tb - T, 0) while day length is decreasing and below Pstart and reset
    otherwise.  ``sea``: dict acclim_x, ssen_u, ssen_l.  Returns the new states and the
    senescence triggers (None where the canopy's photoperiod switch is off)."""
    tc = td - TMELT
    dl = daylength_hours(lat_rad, jday)
    dec = dl < daylength_hours(lat_rad, jday - 1)
    new = dict(acclim_x=sea["acclim_x"] + (tc - sea["acclim_x"]) / acclim_tau,
               ssen_u=np.where(dec & (dl < pp.sen_pstart_u),
                               sea["ssen_u"] + np.maximum(pp.sen_tb_u - tc, 0.0), 0.0),
               ssen_l=np.where(dec & (dl < pp.sen_pstart_l),
                               sea["ssen_l"] + np.maximum(pp.sen_tb_l - tc, 0.0), 0.0))
    on_u = new["ssen_u"] >= pp.sen_ycrit_u if pp.sen_photoperiod_u else None
    on_l = new["ssen_l"] >= pp.sen_ycrit_l if pp.sen_photoperiod_l else None
    return new, on_u, on_l


def pheno(v, a10, td, eps, pp: PhenoParams, stresstu, stresstl, sen_on_u=None, sen_on_l=None):
    """Leaf phenology (physics/phenology.py:94-287).  v: dict of vegetation fields (tempu,
    templ, agddu, agddl, dropu, dropls, dropl3, dropl4 (n,), plai (n,12), tc, frac (n,12),
    lai (n,2), ztop, zbot (n,2), fu, fl); a10: 10-day means (a10td, a10ancub, a10ancls,
    a10ancl3, a10ancl4); ``sen_on_u`` / ``sen_on_l``: the photoperiod senescence triggers,
    which force the cold branch and the degree-day reset (:148-150, :170-172).  Returns the
    updated fields."""
    ddfac = 1.0 / pp.ddays
    a10td = a10["a10td"]
    # cold phenology, upper (:140-158) and lower (:160-178) canopy
    tthr_u = np.maximum(np.maximum(0.0 + TMELT, v["tc"] + pp.tthresh_u_offset + TMELT),
                        pp.tthresh_u_abs + TMELT)
    gthr_u = pp.gthresh_u_c + TMELT
    cold_u, greset_u = a10td < tthr_u, a10td < gthr_u
    if sen_on_u is not None:
        cold_u, greset_u = cold_u | sen_on_u, greset_u | sen_on_u
    agddu = np.where(greset_u, 0.0, v["agddu"] + td - gthr_u)
    tempu = np.where(cold_u, np.maximum(0.0, v["tempu"] - ddfac),
                     np.minimum(1.0, np.maximum(0.0, agddu - pp.gdd_thresh_u) / pp.gdd_ramp_u))
    tthr_l = pp.tthresh_l_c + TMELT
    gthr_l = pp.gthresh_l_c + TMELT
    cold_l, greset_l = a10td < tthr_l, a10td < gthr_l
    if sen_on_l is not None:
        cold_l, greset_l = cold_l | sen_on_l, greset_l | sen_on_l
    agddl = np.where(greset_l, 0.0, v["agddl"] + td - gthr_l)
    templ = np.where(cold_l, np.maximum(0.0, v["templ"] - ddfac),
                     np.minimum(1.0, np.maximum(0.0, agddl - pp.gdd_thresh_l) / pp.gdd_ramp_l))

    # drought phenology (:180-203)
    def drop(x, anc, s):
        return np.where((anc < pp.drop_thresh) | (s < pp.drop_stress_thresh),
                        np.maximum(pp.drop_floor, x - ddfac), np.minimum(1.0, x + ddfac))
    dropu = drop(v["dropu"], a10["a10ancub"], stresstu)
    dropls = drop(v["dropls"], a10["a10ancls"], stresstl)
    dropl4 = drop(v["dropl4"], a10["a10ancl4"], stresstl)
    dropl3 = drop(v["dropl3"], a10["a10ancl3"], stresstl)

    # displayed leaf area and fractions (:205-262)
    p = v["plai"]
    avglaiu = (p[:, 0] + p[:, 1] * dropu + p[:, 2] + p[:, 3] + p[:, 4] * tempu + p[:, 5]
               + p[:, 6] * tempu + p[:, 7] * tempu)
    su = np.maximum(avglaiu, eps)
    frac = v["frac"].copy()
    frac[:, 0] = p[:, 0] / su
    frac[:, 1] = p[:, 1] * dropu / su
    frac[:, 2] = p[:, 2] / su
    frac[:, 3] = p[:, 3] / su
    frac[:, 4] = p[:, 4] * tempu / su
    frac[:, 5] = p[:, 5] / su
    frac[:, 6] = p[:, 6] * tempu / su
    frac[:, 7] = p[:, 7] * tempu / su
    m9, m10, m11 = np.minimum(templ, dropls), np.minimum(templ, dropl4), np.minimum(templ, dropl3)
    avglail = p[:, 8] + p[:, 9] * m9 + p[:, 10] * m10 + p[:, 11] * m11
    sl = np.maximum(avglail, eps)
    frac[:, 8] = p[:, 8] / sl
    frac[:, 9] = p[:, 9] * m9 / sl
    frac[:, 10] = p[:, 10] * m10 / sl
    frac[:, 11] = p[:, 11] * m11 / sl
    lai_l = np.minimum(avglail / np.maximum(v["fl"], eps), 12.0)
    lai_u = np.minimum(avglaiu / np.maximum(v["fu"], eps), 12.0)
    lai = np.stack([lai_l, lai_u], 1)
    zbot = v["zbot"].copy()
    zbot[:, 0] = 0.05
    ztop = v["ztop"].copy()
    ztop[:, 0] = np.minimum(np.maximum(0.25, lai_l * 0.25), v["zbot"][:, 1] - 0.5)
    return dict(tempu=tempu, templ=templ, agddu=agddu, agddl=agddl, dropu=dropu,
                dropls=dropls, dropl3=dropl3, dropl4=dropl4, frac=frac, lai=lai, ztop=ztop,
                zbot=zbot)


def ten_day_means(a10, td, sums, day_sec=86400.0):
    """End-of-day 10-day means (loops_gswp3.py:470-477): weight exp(-1/10) per day."""
    z = np.exp(-1.0 / 10.0)
    al = 1.0 - z
    out = {"a10td": z * a10["a10td"] + al * td}
    for k in ("a10ancub", "a10ancls", "a10ancl3", "a10ancl4"):
        out[k] = z * a10[k] + al * (sums[k] / day_sec)
    return out


def climate_factors(tsoi, wsoi, wisoi, hsoi, bgc: BGCParams):
    """decompl (top layer) and decomps (top-4-layer thickness-weighted means), Lloyd-Taylor
    temperature x WFPS moisture, clipped to [0.001, 10] (loops_gswp3.py:494-553)."""
    tconst = bgc.lt_e0
    wshift = bgc.wfps_opt - 60.0
    btemp, tzero = 288.16, 227.13

    def ftemp(t):
        st = np.maximum(t, tzero + 10.0)
        return np.minimum(np.exp(tconst * (1.0 / (btemp - tzero) - 1.0 / (st - tzero))), 10.0)

    def fmoist(w, wi):
        wfps = (1.0 - wi) * w * 100.0 - wshift
        wet = 0.000371 * wfps * wfps - 0.0748 * wfps + 4.13
        dry = np.exp((wfps - 60.0) ** 2 / -800.0)
        return np.where(wfps >= 60.0, wet, dry)

    fl = ftemp(tsoi[:, 0]) * fmoist(wsoi[:, 0], wisoi[:, 0])
    decompl = np.clip(fl * bgc.k_scale, 0.001, 10.0)
    h = hsoi
    h4 = h[0] + h[1] + h[2] + h[3]
    avg = lambda x: (x[:, 0] * h[0] + x[:, 1] * h[1] + x[:, 2] * h[2] + x[:, 3] * h[3]) / h4
    fs = ftemp(avg(tsoi)) * fmoist(avg(wsoi), avg(wisoi))
    decomps = np.clip(fs * bgc.k_scale, 0.001, 10.0)
    return decompl, decomps


def litter_splits(bgc: BGCParams):
    """Lignin / structural / metabolic split of leaf, root and wood litter (Verberne;
    biogeochem.py:62-77)."""
    cnr = bgc.cnr

    def split(cn, cn_struct):
        fl = bgc.fmax * cn ** 2 / (bgc.rconst + cn ** 2)
        fs = (1.0 / cn - fl / cnr[4] - (1.0 - fl) / cnr[6]) / (1.0 / cn_struct - 1.0 / cnr[6])
        return fl, fs, 1.0 - fl - fs
    return split(bgc.cnleaf, cnr[5]), split(bgc.cnroot, cnr[5]), split(bgc.cnwood, cnr[7])


def soilbgc(c, sand, clay, hsoi, bgc: BGCParams, ndaypy, summ):
    """One day of CENTURY-like soil C/N (physics/biogeochem.py:23-278).  c: dict of CARBON
    fields (n,); summ: storedn, yrleach, ynleach.  Returns (carbon, tco2mic, tnmin, summ)."""
    totts = 1.0 / ndaypy
    cnr = bgc.cnr
    (fll, fls, flm), (frl, frs, frm), (fwl, fws, fwm) = litter_splits(bgc)
    rdepth = 1.0 / (hsoi[0] + hsoi[1] + hsoi[2] + hsoi[3])
    texfact = np.sum(0.01 * sand[:, :4] * hsoi[:4], axis=1) * rdepth
    fbpom = 0.50
    totcin = (c["falll"] + c["fallr"] + c["fallw"]) * totts
    totcbegin = (c["clitlm"] + c["clitls"] + c["clitrm"] + c["clitrs"] + c["clitwm"]
                 + c["clitws"] + c["csoislop"] + c["csoislon"] + c["csoipas"] + c["totcmic"]
                 + c["clitll"] + c["clitrl"] + c["clitwl"])
    cmicmx = bgc.fbsom * totcbegin
    cmicp = np.minimum(c["totcmic"], cmicmx)
    cmicn = np.maximum(0.0, c["totcmic"] - cmicmx)
    lm = c["clitlm"] + flm * c["falll"] * totts
    ls = c["clitls"] + fls * c["falll"] * totts
    ll = c["clitll"] + fll * c["falll"] * totts
    rm = c["clitrm"] + frm * c["fallr"] * totts
    rs = c["clitrs"] + frs * c["fallr"] * totts
    rl = c["clitrl"] + frl * c["fallr"] * totts
    wm = c["clitwm"] + fwm * c["fallw"] * totts
    ws = c["clitws"] + fws * c["fallw"] * totts
    wl = c["clitwl"] + fwl * c["fallw"] * totts
    dl, ds = c["decompl"], c["decomps"]
    b = bgc
    o_rs = np.minimum(ds * b.krs * rs, rs)
    o_ws = np.minimum(dl * b.kws * ws, ws)
    o_ls = np.minimum(dl * b.kls * ls, ls)
    o_lm = np.minimum(dl * b.klm * lm, lm)
    o_rm = np.minimum(ds * b.krm * rm, rm)
    o_wm = np.minimum(dl * b.kwm * wm, wm)
    o_nb = np.minimum(ds * b.knb * c["csoislon"], c["csoislon"])
    o_pb = np.minimum(ds * b.kpb * c["csoislop"], c["csoislop"])
    o_sb = np.minimum(ds * b.ksb * c["csoipas"], c["csoipas"])
    o_bp = np.minimum(b.kbp * cmicp, cmicp) * b.effac
    o_bn = np.minimum(b.kbn * cmicn, cmicn) * b.effac
    dbdt = (o_rs * b.yrs + o_ws * b.yws + o_ls * b.yls + o_lm * b.ylm + o_rm * b.yrm
            + o_wm * b.ywm + o_nb * b.ynb + o_pb * b.ypb + o_sb * b.ysb - o_bp - o_bn)
    o_ll = np.minimum(dl * b.kll * ll, ll)
    o_rl = np.minimum(ds * b.krl * rl, rl)
    o_wl = np.minimum(dl * b.kwl * wl, wl)
    o_ns = np.minimum(ds * b.kns * c["csoislon"], c["csoislon"])
    lig = b.lig_frac * (o_ll * b.yll + o_rl * b.yrl + o_wl * b.ywl)
    dcndt = lig + (1.0 - fbpom) * (b.ybn * o_bn + b.ybp * o_bp) - o_nb - o_ns
    o_ps = np.minimum(ds * b.kps * c["csoislop"], c["csoislop"])
    dcpdt = lig + fbpom * (b.ybn * o_bn + b.ybp * o_bp) - o_pb - o_ps
    fleach = 0.30e-03 / 18.0 * (0.01 + 0.04 * texfact)
    dcsdt = (b.yns * o_ns + b.yps * o_ps) - o_sb - fleach * c["csoipas"]
    cleach = fleach * (c["csoipas"] + c["csoislop"] + c["csoislon"])
    ynleach = (summ["ynleach"] + fleach * c["csoipas"] / cnr[1]
               + fleach * c["csoislop"] / cnr[2] + fleach * c["csoislon"] / cnr[3])
    dcndt = dcndt - fleach * c["csoislon"]
    dcpdt = dcpdt - fleach * c["csoislop"]
    yrleach = summ["yrleach"] + cleach

    def nmin_imm(y, cfrom, cto, out):
        net = (1.0 / cto - y / cfrom) * out
        imm = y / cfrom > 1.0 / cto
        return np.where(imm, 0.0, net), np.where(imm, net, 0.0)
    pairs = [nmin_imm(b.yrs, cnr[0], cnr[5], o_rs), nmin_imm(b.yls, cnr[0], cnr[5], o_ls),
             nmin_imm(b.yws, cnr[0], cnr[7], o_ws), nmin_imm(b.ywm, cnr[0], cnr[7], o_wm),
             nmin_imm(b.ylm, cnr[0], cnr[6], o_lm), nmin_imm(b.yrm, cnr[0], cnr[6], o_rm),
             nmin_imm(b.ynb, cnr[0], cnr[3], o_nb), nmin_imm(b.ypb, cnr[0], cnr[2], o_pb),
             nmin_imm(b.ysb, cnr[0], cnr[1], o_sb)]
    totimm = sum(p[1] for p in pairs)
    totmin = sum(p[0] for p in pairs)
    nrel = (o_ps * (1.0 / cnr[2] - 1.0 / cnr[1]) + o_ns * (1.0 / cnr[3] - 1.0 / cnr[1])
            + ((1.0 - fbpom) * o_bn * (1.0 / cnr[0] - 1.0 / cnr[3])
               + (1.0 - fbpom) * o_bp * (1.0 / cnr[0] - 1.0 / cnr[3]))
            + (fbpom * o_bp * (1.0 / cnr[0] - 1.0 / cnr[2])
               + fbpom * o_bn * (1.0 / cnr[0] - 1.0 / cnr[2]))
            + (b.lig_frac * o_ll * (1.0 / cnr[4] - 1.0 / cnr[2])
               + b.lig_frac * o_ll * (1.0 / cnr[4] - 1.0 / cnr[3]))
            + (b.lig_frac * o_rl * (1.0 / cnr[4] - 1.0 / cnr[2])
               + b.lig_frac * o_rl * (1.0 / cnr[4] - 1.0 / cnr[3]))
            + (b.lig_frac * o_wl * (1.0 / cnr[4] - 1.0 / cnr[2])
               + b.lig_frac * o_wl * (1.0 / cnr[4] - 1.0 / cnr[3])))
    netmin = totmin + totimm + nrel
    tnmin = np.maximum(0.0, netmin) / (86400.0 * 0.014)
    out = dict(c)
    out["totcmic"] = np.maximum(0.0, c["totcmic"] + dbdt)
    out["csoislon"] = np.maximum(0.0, c["csoislon"] + dcndt)
    out["csoislop"] = np.maximum(0.0, c["csoislop"] + dcpdt)
    out["csoipas"] = np.maximum(0.0, c["csoipas"] + dcsdt)
    out["clitlm"] = np.maximum(0.0, lm - o_lm)
    out["clitls"] = np.maximum(0.0, ls - o_ls)
    out["clitll"] = np.maximum(0.0, ll - o_ll)
    out["clitrm"] = np.maximum(0.0, rm - o_rm)
    out["clitrs"] = np.maximum(0.0, rs - o_rs)
    out["clitrl"] = np.maximum(0.0, rl - o_rl)
    out["clitwm"] = np.maximum(0.0, wm - o_wm)
    out["clitws"] = np.maximum(0.0, ws - o_ws)
    out["clitwl"] = np.maximum(0.0, wl - o_wl)
    out["totlit"] = (out["clitlm"] + out["clitls"] + out["clitll"] + out["clitrm"]
                     + out["clitrs"] + out["clitrl"] + out["clitwm"] + out["clitws"]
                     + out["clitwl"])
    out["totalit"] = (out["clitlm"] + out["clitls"] + out["clitwm"] + out["clitll"]
                      + out["clitws"] + out["clitwl"])
    out["totrlit"] = out["clitrm"] + out["clitrs"] + out["clitrl"]
    out["totcsoi"] = out["csoipas"] + out["csoislop"] + out["totcmic"] + out["csoislon"]
    out["totfall"] = c["falll"] + c["fallr"] + c["fallw"]
    totcend = out["totlit"] + out["totcsoi"]
    tco2mic = (totcbegin + totcin - totcend - cleach) / (86400.0 * 0.012)
    s = dict(storedn=summ["storedn"] + tnmin, yrleach=yrleach, ynleach=ynleach)
    return out, tco2mic, tnmin, s


def day_end(day, sums, td, tsoi, wsoi, wisoi, sand, clay, hsoi, bgc: BGCParams, ndaypy=365,
            nsoilcall=1):
    """End-of-day work (loops_gswp3.py:470-585).  day: dict with the 10-day means, annual sums
    (ayprcp ayaet ayco2mic ayco2root storedn yrleach ynleach), tco2root and the CARBON
    fields; sums: the day's adrain adsnow adaet a10ancub a10ancls a10ancl3 a10ancl4.
    Returns (day', tco2mic, tnmin)."""
    out = dict(day)
    out.update(ten_day_means(day, td, sums))
    out["ayprcp"] = day["ayprcp"] + sums["adrain"] + sums["adsnow"]
    out["ayaet"] = day["ayaet"] + sums["adaet"]
    dl, ds = climate_factors(tsoi, wsoi, wisoi, hsoi, bgc)
    c = {k: day[k] for k in CARBON}
    c["decompl"], c["decomps"] = dl, ds
    summ = {k: day[k] for k in ("storedn", "yrleach", "ynleach")}
    tco2mic = tnmin = None
    for _ in range(nsoilcall):
        c, tco2mic, tnmin, summ = soilbgc(c, sand, clay, hsoi, bgc, ndaypy, summ)
    out.update(c)
    out.update(summ)
    out["ayco2mic"] = day["ayco2mic"] + tco2mic * (12.0e-3 * 86400.0)
    out["ayco2root"] = day["ayco2root"] + day["tco2root"] * (12.0e-3 * 86400.0)
    return out, tco2mic, tnmin


def stress_accum(acc, td, stresstu, stresstl, tmelt, tmin):
    """The soil-water stress accumulator at day end (loops_gswp3.py:587-599,
    state/stress_accum.py:15-26): growing-season days (td > tmelt + tmin) add the canopies'
    stress factors and one day.  Pinned by its closed form over consecutive days
    (tests/test_daily.py)."""
    grow = (np.asarray(td) > tmelt + tmin).astype(np.float64)
    return dict(aystressu=acc["aystressu"] + stresstu * grow,
                aystressl=acc["aystressl"] + stresstl * grow, nday=acc["nday"] + grow)

