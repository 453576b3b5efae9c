"""The reference's hourly driver around the step (FP64) -- TEST INFRASTRUCTURE.

* ``acc_increments`` -- one step's increment of each per-step accumulator, restating
  ``python-port/ibis/timestepping/loops_gswp3.py:161-302`` (daily rain, snow, AET and big-leaf
  photosynthesis sums; NPP per PFT with Lloyd-Taylor maintenance and growth respiration;
  total NPP and GPP; runoff and drainage; ET components; night and day NEE), in the order of
  ``include/ibis_driver.h``'s ``IbisAcc``.
* ``diurnal`` -- the daily-to-hourly disaggregation of ``python-port/ibis/physics/weather.py:483-601``
  with the wind's uniform draw passed in.
* ``precip_timing`` -- ``compute_precip_timing`` (``weather.py:395-441``) with its two
  uniform draws passed in.

Pinned by (``tests/test_driver.py``): the accumulator identities -- the 12 per-PFT NPP
increments sum to the total, the ET components sum to AET, night + day NEE is the
heterotrophic release minus NPP, zero respiration leaves (1 - rgrowth) x GPP -- and the
diurnal cycle's own closed forms: the day's precipitation total recovered from the hourly
rates, the solar-noon zenith cos(lat - decl), air temperature within [tmin, tmax], humidity
at most 99 % of saturation, wind within [2.5, 10] m/s.  The remaining formula detail
(coefficients of the Friend transmissivity and the Campbell-Norman cycle) has no
reference-held vector: parity unpinned beyond those properties.
"""
from __future__ import annotations

import numpy as np

from .thermo import esat, qsat

#: IbisAcc order (include/ibis_driver.h)
ACC_NAMES = (["adrain", "adsnow", "adaet", "a10ancub", "a10ancls", "a10ancl3", "a10ancl4"]
             + [f"aynpp{j}" for j in range(12)]
             + ["aynpptot", "aygpptot", "aytrunof", "aytdrain", "et_trans", "et_interc",
                "et_soilev", "et_snowsub", "nee_night", "nee_day"])
NACC = len(ACC_NAMES)
_BROAD, _CONIF = (0, 1, 2, 4, 6), (3, 5, 7)


def acc_increments(st, vx, p):
    """(NACC, N) increments of one step.  ``st``: dict of state fields after the step
    (reference shapes); ``vx``: dict with cbiow (N, 12), cbior (N, 12), sapfrac, tco2mic;
    ``p``: namespace with rwood, rroot, rgrowth (12,), froot (6, 2), dtime, ndaypy and, for
    StomataParams.gpp_split_by_capacity (loops_gswp3.py:197-206), ``vmax`` (N, 12) per-slot
    Vmax: the NPP of each upper and shrub big leaf is then shared by leaf area x Vmax."""
    dt = p.dtime
    an_u, an_l = st["canopy.an_upper"], st["canopy.an_lower"]
    ag_u, ag_l = st["canopy.ag_upper"], st["canopy.ag_lower"]
    lai_l, lai_u = st["vegetation.lai"][:, 0], st["vegetation.lai"][:, 1]
    fu, fl, fi = st["vegetation.fu"], st["vegetation.fl"], st["snow.fi"]
    frac = st["vegetation.frac"]
    smask = 1.0 - fi
    froot = np.asarray(p.froot)
    roottemp = np.zeros_like(fi)
    for k in range(6):
        roottemp = roottemp + st["soil.tsoi"][:, k] * 0.5 * (froot[k, 0] + froot[k, 1])
    funca = np.exp(3500.0 * (1.0 / 288.16 - 1.0 / np.maximum(st["canopy.ts"], 200.0)))
    funcb = np.exp(3500.0 * (1.0 / 288.16 - 1.0 / np.maximum(roottemp, 200.0)))
    c_fac = 12.0e-3 * dt
    per_s = 1.0 / (p.ndaypy * 86400.0) * (1000.0 / 12.0)
    rw = np.asarray(p.rwood) * per_s
    rr = np.asarray(p.rroot) * per_s
    rg = np.asarray(p.rgrowth)
    inc = np.zeros((NACC,) + fi.shape)
    npp_tot = np.zeros_like(fi)
    share = frac
    vmax = getattr(p, "vmax", None)
    if vmax is not None:
        from .stomata import capacity_shares
        share = frac.copy()
        for idx in (_BROAD, _CONIF, (8, 9)):
            share[:, list(idx)] = capacity_shares(frac, vmax, idx)
    for j in range(12):
        if j in _BROAD or j in _CONIF:
            an = an_u[:, 1] if j in _CONIF else an_u[:, 0]
            gpp = an * share[:, j] * lai_u * fu
            resp = (rw[j] * vx["cbiow"][:, j] * vx["sapfrac"] * funca
                    + rr[j] * vx["cbior"][:, j] * funcb)
        elif j in (8, 9):
            gpp = an_l[:, 0] * share[:, j] * lai_l * fl * smask
            resp = rw[j] * vx["cbiow"][:, j] * funca + rr[j] * vx["cbior"][:, j] * funcb
        else:
            an = an_l[:, 2] if j == 10 else an_l[:, 1]
            gpp = an * frac[:, j] * lai_l * fl * smask
            resp = rr[j] * vx["cbior"][:, j] * funcb
        npp = (1.0 - rg[j]) * (gpp - resp)
        inc[7 + j] = npp * c_fac
        npp_tot = npp_tot + npp * c_fac
    fb = frac[:, list(_BROAD)].sum(axis=1)
    fc = frac[:, list(_CONIF)].sum(axis=1)
    gpp_u = (ag_u[:, 0] * fb + ag_u[:, 1] * fc) * lai_u * fu
    gpp_l = (ag_l[:, 0] * (frac[:, 8] + frac[:, 9]) + ag_l[:, 2] * frac[:, 10]
             + ag_l[:, 1] * frac[:, 11]) * lai_l * fl * smask
    h = lambda k: st[f"hydrology.{k}"]
    inc[0] = st["atm.raina"] * dt
    inc[1] = st["atm.snowa"] * dt
    inc[2] = (h("ginvap") + h("gsuvap") + h("gtrans")) * dt
    inc[3] = an_u[:, 0] * dt
    inc[4] = an_l[:, 0] * dt
    inc[5] = an_l[:, 1] * dt
    inc[6] = an_l[:, 2] * dt
    inc[19] = npp_tot
    inc[20] = (gpp_u + gpp_l) * c_fac
    inc[21] = h("grunof") * dt
    inc[22] = h("gdrain") * dt
    snowsub = st["canopy.fvapi"] * fi * dt
    inc[23] = h("gtrans") * dt
    inc[24] = h("ginvap") * dt
    inc[25] = h("gsuvap") * dt - snowsub
    inc[26] = snowsub
    nee = vx["tco2mic"] * c_fac - npp_tot
    night = st["atm.coszen"] <= 0.0
    inc[27] = np.where(night, nee, 0.0)
    inc[28] = np.where(night, 0.0, nee)
    return inc


def precip_timing(precip, dtime, u_len, u_start):
    """Event length, start and end (s) of the day's precipitation (weather.py:395-441)."""
    niter = int(86400.0 / dtime)
    plenmin = 1 + int((4.0 * 3600.0 - 1.0) / dtime)
    plenmax = max(int(24.0 * 3600.0 / dtime), plenmin)
    plen = np.clip(plenmin + (u_len * (plenmax - plenmin + 1)).astype(np.int64), plenmin, plenmax)
    plen = np.clip(np.maximum(plen, np.ceil(precip / 15.0).astype(np.int64)), plenmin, plenmax)
    plens = dtime * plen
    max_start = np.maximum(0, niter - plen)
    start = np.minimum((u_start * (max_start + 1)).astype(np.int64), max_start)
    startp = dtime * start
    return plens, startp, startp + plens


def diurnal(time, jday, dw, lat_deg, dtime, rn_wind):
    """Hourly forcing from daily weather ``dw`` (dict of td, tmax, tmin, qd, ud, precip,
    cloud, psurf, plens, startp, endp) at ``time`` s after midnight (weather.py:483-601)."""
    pi = np.pi
    rtime = time / 3600.0
    orbit = 2.0 * pi * jday / 365.2425
    xdecl = (0.006918 - 0.399912 * np.cos(orbit) + 0.070257 * np.sin(orbit)
             - 0.006758 * np.cos(2.0 * orbit) + 0.000907 * np.sin(2.0 * orbit)
             - 0.002697 * np.cos(3.0 * orbit) + 0.001480 * np.sin(3.0 * orbit))
    angle = 2.0 * pi * (rtime - 12.0) / 24.0
    sw = 1370.0 * (1.000110 + 0.034221 * np.cos(orbit) + 0.001280 * np.sin(orbit)
                   + 0.000719 * np.cos(2.0 * orbit) + 0.000077 * np.sin(2.0 * orbit))
    xlat = lat_deg * pi / 180.0
    coszen = np.maximum(0.0, np.sin(xlat) * np.sin(xdecl)
                        + np.cos(xlat) * np.cos(xdecl) * np.cos(angle))
    cloud = dw["cloud"]
    trans = 0.251 + 0.509 * (1.0 - cloud)
    fdiff = np.where(trans > 0.75, 0.166,
                     1.0045 + 0.0435 * trans - 3.5227 * trans * trans
                     + 2.6313 * trans * (trans * trans))
    frac = np.array([0.46, 0.54])
    solad = sw * coszen[:, None] * frac[None, :] * trans[:, None] * (1.0 - fdiff[:, None])
    solai = sw * coszen[:, None] * frac[None, :] * trans[:, None] * fdiff[:, None]
    gamma = 0.44 - 0.46 * np.sin(pi / 12.0 * rtime + 0.9) + 0.11 * np.sin(2.0 * pi / 12.0 * rtime + 0.9)
    ta = dw["tmax"] * gamma + dw["tmin"] * (1.0 - gamma)
    ps = dw["psurf"]
    qmin = np.minimum(dw["qd"], 0.99 * qsat(esat(dw["tmin"]), ps))
    qmax = (dw["qd"] - 0.56 * qmin) / 0.44
    qa = np.minimum(0.99 * qsat(esat(ta), ps), qmax * gamma + qmin * (1.0 - gamma))
    ua = np.clip(1.13989 * dw["ud"] * (-np.log(np.maximum(rn_wind, 1e-10))) ** 0.30, 2.5, 10.0)
    emb = 0.01 * (ps * qa / (0.622 + qa))
    ea = 0.70 + 5.95e-5 * emb * np.exp(1500.0 / ta)
    t4 = (ta * ta) * (ta * ta)
    fira = (1.0 - cloud) * ea * 5.67e-8 * t4 + cloud * 0.950 * 5.67e-8 * t4
    in_event = (time >= dw["startp"]) & (time < dw["endp"])
    warm = (ta - 273.15) > 2.5
    rate = dw["precip"] / np.maximum(dw["plens"], dtime)
    return {"atm.coszen": coszen, "atm.ta": ta, "atm.qa": qa, "atm.ua": ua, "atm.fira": fira,
            "atm.solad": solad, "atm.solai": solai, "atm.psurf": ps,
            "atm.raina": np.where(in_event & warm, rate, 0.0),
            "atm.snowa": np.where(in_event & ~warm, rate, 0.0)}
