"""Canopy interception, wet fractions and evaporation cascade (FP64), restating
``python-port/ibis/physics/interception.py``.

Pinned by: cascade water conservation -- rain + snow in = change of stored water + what
reaches the ground + drip/blow-off (``tests/test_f77_verification.py:747-787``); the wet-bulb
temperature against a brute-force root of its defining equation when the relaxation
converges; ``mix`` as a mass-weighted mean; ``noveg`` on bare cells giving the areal
surface temperature (``interception.py:470-495``).  ``twet3``'s latch result after 20
iterations where it has not converged is parity unpinned.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from .thermo import esat, qsat, hvapf, hsubf


def fwetcal(tu, ts, tl, wliqu, wliqs, wliql, wsnou, wsnos, wsnol, ip, c):
    """Wet fractions and liquid share (interception.py:62-94)."""
    def one(t, wl, ws, wlmax, wsmax):
        xliq = wl / max(wlmax, 0.01)
        xtot = xliq + ws / max(wsmax, 0.01)
        fwet = np.minimum(ip.fmax, xtot)
        rliq = xliq / np.maximum(xtot, c.epsilon)
        rliq = np.where(fwet == 0.0, np.where(t < c.tmelt - 1.0e-9, 0.0, 1.0), rliq)
        return fwet, rliq
    fu_, ru = one(tu, wliqu, wsnou, ip.wliqumax, ip.wsnoumax)
    fs_, rs = one(ts, wliqs, wsnos, ip.wliqsmax, ip.wsnosmax)
    fl_, rl = one(tl, wliql, wsnol, ip.wliqlmax, ip.wsnolmax)
    return SimpleNamespace(fwetu=fu_, fwets=fs_, fwetl=fl_, rliqu=ru, rliqs=rs, rliql=rl)


def mix(x1, t1, x2, t2, x3, t3, eps):
    """Calorimetric mix of three masses (interception.py:101-119)."""
    xm = x1 + x2 + x3
    ok = np.abs(xm) >= eps
    safe = np.where(ok, xm, 1.0)
    tm = np.where(ok, (t1 * x1 + t2 * x2 + t3 * x3) / safe, 0.0)
    return np.where(ok, xm, 0.0), tm


def twet3(tak, q, p, c):
    """Wet-bulb temperature by 20 latched relaxation steps (interception.py:126-156)."""
    qs0 = qsat(esat(tak), p)
    ta = tak - c.tmelt
    twk = ta * q / np.maximum(qs0, 1e-12) + c.tmelt
    done = np.zeros(twk.shape, dtype=bool)
    for _ in range(20):
        twold = twk - c.tmelt
        hv = hvapf(twk, tak, c)
        qs = qsat(esat(twk), p)
        twb = ta - (hv / c.cair) * (qs - q)
        new = twold + 0.2 * (twb - twold) + c.tmelt
        conv = np.abs(new - c.tmelt - twold) < 0.02
        twk = np.where(done, twk, new)
        done = done | conv
    return twk


def steph2o(tveg, wliq, wsno, xai, rain, train, snow, tsnow, tdrip, tblow, wliqmax, wsnomax,
            wliqmin, wsnomin, dtime, c):
    """Interception, drip, blow-off and phase change on one component (interception.py:174-235)."""
    eps = c.epsilon
    with np.errstate(divide="ignore", invalid="ignore"):
        fint = np.where(xai >= eps, (1.0 - np.exp(-0.5 * xai)) / xai, 0.5)
    rw = 1.0 / dtime
    drip = xai * wliq / tdrip
    wliq = wliq * (1.0 - dtime / tdrip)
    wliq = wliq + dtime * rain * fint
    pflux = rain * fint * (tveg - train) * c.ch2o
    rain = rain * (1.0 - xai * fint)
    x = wliq
    wliq = np.minimum(wliq, wliqmax)
    wliq = np.where(wliq < wliqmin, 0.0, wliq)
    drip = drip + xai * (x - wliq) * rw

    blow = xai * wsno / tblow
    wsno = wsno * (1.0 - dtime / tblow)
    wsno = wsno + dtime * snow * fint
    pflux = pflux + snow * fint * (tveg - tsnow) * c.cice
    snow = snow * (1.0 - xai * fint)
    x = wsno
    wsno = np.minimum(wsno, wsnomax)
    wsno = np.where(wsno < wsnomin, 0.0, wsno)
    blow = blow + xai * (x - wsno) * rw

    dw = np.where(tveg < c.tmelt - 1.0e-9, wliq, 0.0)
    cw = c.ch2o - c.cice
    pflux = pflux + dw * (cw * (c.tmelt - tveg) - c.hfus) * rw
    wliq = wliq - dw
    wsno = wsno + dw
    dw = np.where(tveg > c.tmelt + 1.0e-9, wsno, 0.0)
    pflux = pflux + dw * (cw * (tveg - c.tmelt) + c.hfus) * rw
    wsno = wsno - dw
    wliq = wliq + dw

    z = np.zeros_like(rain)
    rain, train = mix(rain, train, drip, tveg, z, z, eps)
    snow, tsnow = mix(snow, tsnow, blow, tveg, z, z, eps)
    return SimpleNamespace(wliq=wliq, wsno=wsno, pflux=pflux, rain=rain, train=train,
                           snow=snow, tsnow=tsnow)


def cascade(tu, ts, tl, t12, t34, wliqu, wliqs, wliql, wsnou, wsnos, wsnol, lai, sai, fu, fl,
            ta, qa, psurf, raina, snowa, ip, dtime, c):
    """Precipitation through the two canopies (interception.py:263-357)."""
    eps, tmelt = c.epsilon, c.tmelt
    z = np.zeros_like(tu)
    wliqmin = 0.001 * (dtime / 3600.0) * (ip.wliqumax / 0.2)
    wsnomin = 0.001 * (dtime / 3600.0) * (ip.wsnoumax / 2.0)
    twb = np.where(ta > tmelt + 1.0e-9, twet3(ta, qa, psurf, c), tmelt)
    trainu = np.maximum(twb, tmelt)
    rainu, trainu = mix(raina, trainu, z, np.maximum(t12, tmelt), z, z, eps)
    tsnowu = np.minimum(ta, tmelt)
    snowu, tsnowu = mix(snowa, tsnowu, z, np.minimum(t12, tmelt), z, z, eps)

    u = steph2o(tu, wliqu, wsnou, 2.0 * lai[:, 1], rainu, trainu, snowu, tsnowu,
                ip.tdripu, ip.tblowu, ip.wliqumax, ip.wsnoumax, wliqmin, wsnomin, dtime, c)
    s = steph2o(ts, wliqs, wsnos, 2.0 * sai[:, 1], u.rain, u.train, u.snow, u.tsnow,
                ip.tdrips, ip.tblows, ip.wliqsmax, ip.wsnosmax, wliqmin, wsnomin, dtime, c)
    rainl, trainl = mix(fu * s.rain, s.train, (1.0 - fu) * rainu, trainu, z,
                        np.maximum(t34, tmelt), eps)
    snowl, tsnowl = mix(fu * s.snow, s.tsnow, (1.0 - fu) * snowu, tsnowu, z,
                        np.minimum(t34, tmelt), eps)
    lo = steph2o(tl, wliql, wsnol, 2.0 * (lai[:, 0] + sai[:, 0]), rainl, trainl, snowl, tsnowl,
                 ip.tdripl, ip.tblowl, ip.wliqlmax, ip.wsnolmax, wliqmin, wsnomin, dtime, c)
    raing, traing = mix(fl * lo.rain, lo.train, (1.0 - fl) * rainl, trainl, z, z, eps)
    snowg, tsnowg = mix(fl * lo.snow, lo.tsnow, (1.0 - fl) * snowl, tsnowl, z, z, eps)
    return SimpleNamespace(wliqu=u.wliq, wliqs=s.wliq, wliql=lo.wliq, wsnou=u.wsno,
                           wsnos=s.wsno, wsnol=lo.wsno, pfluxu=u.pflux, pfluxs=s.pflux,
                           pfluxl=lo.pflux, raing=raing, traing=traing, snowg=snowg,
                           tsnowg=tsnowg, rainl=rainl, trainl=trainl, snowl=snowl,
                           tsnowl=tsnowl)


def steph2o2(tveg, wliq, wsno, fveg, xai, rliq, fvapw, cveg, dtime, ta, fvapa, fsena, c):
    """Remove intercepted evaporation, heat-conserving phase correction (interception.py:364-416)."""
    tmelt, hfus = c.tmelt, c.hfus
    wliq = wliq - dtime * rliq * fvapw
    wsno = wsno - dtime * (1.0 - rliq) * fvapw
    has = fveg * xai > 0.0
    zm = np.where(has, np.maximum(-wliq, 0.0) * fveg * xai / dtime, 0.0)
    fvapa = fvapa + zm
    fsena = fsena - zm * hvapf(tveg, ta, c)
    wliq = np.maximum(wliq, 0.0)
    zm = np.where(has, np.maximum(-wsno, 0.0) * fveg * xai / dtime, 0.0)
    fvapa = fvapa + zm
    fsena = fsena - zm * hsubf(tveg, ta, c)
    wsno = np.maximum(wsno, 0.0)

    rw = 1.0 / hfus
    chav = cveg + c.ch2o * wliq + c.cice * wsno
    below = (tveg < tmelt - 1.0e-9) & (wliq > 0.0)
    dh = chav * (tmelt - tveg)
    dw = np.where(below, np.minimum(wliq, np.maximum(0.0, dh * rw)), 0.0)
    wliq = wliq - dw
    wsno = wsno + dw
    chn = cveg + c.ch2o * wliq + c.cice * wsno
    tveg = np.where(below, tmelt - (dh - hfus * dw) / chn, tveg)

    above = (tveg > tmelt + 1.0e-9) & (wsno > 0.0)
    chav = cveg + c.ch2o * wliq + c.cice * wsno
    dh = chav * (tveg - tmelt)
    dw = np.where(above, np.minimum(wsno, np.maximum(0.0, dh * rw)), 0.0)
    wsno = wsno - dw
    wliq = wliq + dw
    chn = cveg + c.ch2o * wliq + c.cice * wsno
    tveg = np.where(above, tmelt + (dh - hfus * dw) / chn, tveg)
    return tveg, wliq, wsno, fvapa, fsena


def cascad2(tu, ts, tl, wliqu, wliqs, wliql, wsnou, wsnos, wsnol, rliqu, rliqs, rliql,
            fvapuw, fvaps, fvaplw, fu, fl, fi, lai, sai, cp, dtime, ta, fvapa, fsena, c):
    """End-of-hour evaporation cascade (interception.py:423-463)."""
    tu, wliqu, wsnou, fvapa, fsena = steph2o2(tu, wliqu, wsnou, fu, 2.0 * lai[:, 1], rliqu,
                                              fvapuw, cp.chu, dtime, ta, fvapa, fsena, c)
    ts, wliqs, wsnos, fvapa, fsena = steph2o2(ts, wliqs, wsnos, fu, 2.0 * sai[:, 1], rliqs,
                                              fvaps, cp.chs, dtime, ta, fvapa, fsena, c)
    tl, wliql, wsnol, fvapa, fsena = steph2o2(tl, wliql, wsnol, (1.0 - fi) * fl,
                                              2.0 * (lai[:, 0] + sai[:, 0]), rliql, fvaplw,
                                              cp.chl, dtime, ta, fvapa, fsena, c)
    return tu, ts, tl, wliqu, wliqs, wliql, wsnou, wsnos, wsnol, fvapa, fsena


def noveg(tu, ts, tl, wliqu, wliqs, wliql, wsnou, wsnos, wsnol, lai, sai, fu, fl, fi, tg, ti):
    """Vegetation temperatures where a canopy is absent (interception.py:470-495)."""
    tav = (1.0 - fi) * tg + fi * ti
    m = (lai[:, 1] == 0.0) | (fu == 0.0)
    tu = np.where(m, tav, tu)
    wliqu = np.where(m, 0.0, wliqu)
    wsnou = np.where(m, 0.0, wsnou)
    m = (sai[:, 1] == 0.0) | (fu == 0.0)
    ts = np.where(m, tav, ts)
    wliqs = np.where(m, 0.0, wliqs)
    wsnos = np.where(m, 0.0, wsnos)
    x = 2.0 * (lai[:, 0] + sai[:, 0])
    y = fl * (1.0 - fi)
    m = (x == 0.0) | (y == 0.0)
    tl = np.where(m, tav, tl)
    wliql = np.where(m, 0.0, wliql)
    wsnol = np.where(m, 0.0, wsnol)
    return tu, ts, tl, wliqu, wliqs, wliql, wsnou, wsnos, wsnol
