"""Soil water and heat, bedrock, freeze/thaw (FP64), restating
``python-port/ibis/physics/soil_physics.py`` and ``python-port/ibis/physics/bedrock.py``.

Pinned by: ``setsoi`` qglif summing to 1 (``tests/test_soil.py:54-73``) and the conductivity
formula zcondry*56^powliq*224^powice (``tests/test_f77_verification.py:219-242``);
``wadjust`` clamps (``tests/test_soil.py:103-148``); ``soilh2o`` column mass balance
(sum of storage change = fluxes through top and bottom, ``tests/test_water_balance.py:35-47``);
``soilheat`` energy balance -- fhtop - hflo[bottom] = stored heat change, exactly in FP64
across the bedrock seam (``tests/test_bedrock.py:198-229``) -- and an isothermal column
staying put (``tests/test_soil.py:203-222``); ``freeze_thaw`` and ``puddle_freeze_thaw``
conserving enthalpy (``tests/test_f77_verification.py:301-332``).
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from ._num import ipow
from .solvers import tridia
from .thermo import hvapf, hsubf


def setsoi(sand, clay, poros, wsoi, wisoi, tsoi, wpud, wipud, wpudmax, zwpmax, tsno1, ta, c):
    """Conductivity, evaporation sources and surface latent heats (soil_physics.py:37-140)."""
    eps = c.epsilon
    fsand = 0.01 * sand
    fclay = 0.01 * clay
    fsilt = 1.0 - fsand - fclay
    powliq = poros * wsoi * (1.0 - wisoi)
    powice = poros * wisoi
    zcondry = fsand * 0.300 + fsilt * 0.265 + fclay * 0.250
    consoi = zcondry * (56.0 ** powliq) * (224.0 ** powice)
    pud = wpud + wipud
    zwpud = np.clip(zwpmax * pud / max(wpudmax, eps), 0.0, zwpmax)
    w1, wi1 = wsoi[:, 0], wisoi[:, 0]
    zwsoi = (1.0 - wi1) * w1 + wi1
    soilw = zwsoi >= eps
    hasp = zwpud >= eps
    r1 = 1.0 / np.maximum(zwsoi, eps)
    r2 = 1.0 / np.maximum(pud, eps)
    above = tsoi[:, 0] >= c.tmelt - 1.0e-9
    z = np.zeros_like(w1)
    q1 = np.where(soilw, np.where(hasp, (1.0 - zwpud) * (1.0 - wi1) * w1 * r1,
                                  (1.0 - wi1) * w1 * r1), z)
    q2 = np.where(soilw, np.where(hasp, (1.0 - zwpud) * wi1 * r1, wi1 * r1), z)
    q3 = np.where(hasp, zwpud * wpud * r2, np.where(soilw, z, np.where(above, 1.0, 0.0)))
    q4 = np.where(hasp, zwpud * wipud * r2, np.where(soilw, z, np.where(above, 0.0, 1.0)))
    zvap = hvapf(tsoi[:, 0], ta, c)
    zsub = hsubf(tsoi[:, 0], ta, c)
    hvasug = (q1 + q3) * zvap + (q2 + q4) * zsub
    hvasui = hsubf(tsno1, ta, c)
    return SimpleNamespace(consoi=consoi, qglif=np.stack([q1, q2, q3, q4], 1), hvasug=hvasug,
                           hvasui=hvasui)


def wadjust(wsoi, wisoi, poros, hsoi, rhow, dtime, eps):
    """Clamp wsoi/wisoi to [0, 1] and report the mass created (soil_physics.py:147-182)."""
    t0 = hsoi[None, :] * poros * rhow * ((1.0 - wisoi) * wsoi + wisoi)
    w = np.minimum(1.0, np.where(wsoi < eps, 0.0, wsoi))
    wi = np.minimum(1.0, np.where(wisoi < eps, 0.0, wisoi))
    w = np.where(wi >= 1.0 - eps, 0.0, w)
    t1 = hsoi[None, :] * poros * rhow * ((1.0 - wi) * w + wi)
    d = t1 - t0
    acc = d[:, 0]
    for k in range(1, d.shape[1]):
        acc = acc + d[:, k]
    return w, wi, acc / dtime


def soilh2o(wsoi, wisoi, poros, porosflo, hydraul, suction, bex, hsoi, fwtop, fwpud, wpud,
            bperm, dtime, rhow, eps, consistent_weights=False):
    """Implicit soil-water step (soil_physics.py:196-373), transpiration already removed."""
    n, ns = wsoi.shape
    dmin = 1e-9
    e = np.zeros((n, ns + 1))
    f = np.zeros((n, ns + 1))
    g = np.zeros((n, ns + 1))
    for k in range(1, ns):
        hm = 0.5 * (hsoi[k - 1] + hsoi[k])
        weip = 0.5 * hsoi[k] / hm
        weim = 1.0 - weip
        a = weim * hydraul[:, k - 1] + weip * hydraul[:, k]
        b = (weim * hydraul[:, k - 1] * suction[:, k - 1] * bex[:, k - 1]
             + weip * hydraul[:, k] * suction[:, k] * bex[:, k])
        zbex = weim * bex[:, k - 1] + weip * bex[:, k]
        nr = np.floor(zbex + 0.5).astype(np.int64)
        m = 2 * nr + 3
        nn = nr + 2
        if consistent_weights:
            wm = weip * wsoi[:, k - 1] + weim * wsoi[:, k]
        else:
            wm = weim * wsoi[:, k - 1] + weip * wsoi[:, k]
        wa = np.minimum(wm, 1.0)
        bwn1 = b * ipow(wa, nn - 1)
        bwn = bwn1 * wa
        bwn1 = np.where(bwn < dmin, 0.0, bwn1)
        bwn = np.maximum(bwn, dmin)
        dw = (wsoi[:, k] - wsoi[:, k - 1]) / hm
        ek = ((-1.0 + 1.0 * m) * a * ipow(wa, m)
              + ((1.0 - 1.0) * bwn - 1.0 * nn * bwn1 * wa) * dw)
        fk = -1.0 * m * a * ipow(wa, m - 1) + 1.0 * nn * bwn1 * dw
        gk = 1.0 * bwn
        zi = np.maximum(0.0, 1.0 - np.maximum(wisoi[:, k - 1], wisoi[:, k]))
        zi = zi * zi
        e[:, k] = zi * ek
        f[:, k] = zi * fk
        g[:, k] = zi * gk
    wb = np.minimum(wsoi[:, ns - 1], 1.0)
    mb = 2 * np.floor(bex[:, ns - 1] + 0.5).astype(np.int64) + 3
    eb = -hydraul[:, ns - 1] * ipow(wb, mb) * bperm
    zi = np.maximum(0.0, 1.0 - wisoi[:, ns - 1])
    e[:, ns] = (zi * zi) * eb

    d1 = np.zeros((n, ns))
    d2 = np.zeros((n, ns))
    d3 = np.zeros((n, ns))
    r = np.zeros((n, ns))
    for k in range(ns):
        dt = dtime / np.maximum(porosflo[:, k] * hsoi[k], 1e-30)
        hm = 0.5 * hsoi[0] if k == 0 else 0.5 * hsoi[k - 1] + 0.5 * hsoi[k]
        hm = max(hm, 1e-30)
        km1 = max(k - 1, 0)
        d1k = dt * (f[:, k] * 0.5 * hsoi[k] / hm - g[:, k] / hm)
        rk = wsoi[:, k] + dt * (e[:, k + 1] - e[:, k])
        if k == 0:
            rk = rk + dt * (fwtop + fwpud) / rhow
            d1k = np.zeros(n)
        if k < ns - 1:
            hp = max(0.5 * (hsoi[k] + hsoi[k + 1]), 1e-30)
            d2k = 1.0 + dt * (-f[:, k + 1] * 0.5 * hsoi[k + 1] / hp
                              + f[:, k] * 0.5 * hsoi[km1] / hm
                              + g[:, k + 1] / hp + g[:, k] / hm)
            d3k = dt * (-f[:, k + 1] * 0.5 * hsoi[k] / hp - g[:, k + 1] / hp)
        else:
            d2k = 1.0 + dt * (-f[:, k + 1] + f[:, k] * 0.5 * hsoi[km1] / hm + g[:, k] / hm)
            d3k = np.zeros(n)
        d1[:, k], d2[:, k], d3[:, k], r[:, k] = d1k, d2k, d3k, rk
    wn = tridia(d1, d2, d3, r)
    wflo = np.zeros((n, ns + 1))
    wflo[:, ns] = -rhow * e[:, ns]
    fsq = np.zeros(n)
    for k in range(ns - 1, -1, -1):
        zz = rhow * poros[:, k] * np.maximum(eps, 1.0 - wisoi[:, k]) * hsoi[k]
        wk = wn[:, k] + dtime * fsq / zz
        fsq_n = np.maximum(wk - 1.0, 0.0) * zz / dtime
        wk = np.minimum(wk, 1.0)
        wflo[:, k] = wflo[:, k + 1] + (wk - wsoi[:, k]) * zz / dtime
        wn[:, k] = wk
        fsq = fsq_n
    wpud_n = np.maximum(wpud + (fsq - fwpud) * dtime, 0.0)
    wflo[:, 0] = wflo[:, 0] + fsq - fwpud
    return SimpleNamespace(wsoi=wn, wpud=wpud_n, wflo=wflo, fsqueez=fsq)


def soilheat(tsoi, otsoi, owsoi, wsoi, wisoi, poros, consoi, csoi, rhosoi, hsoi, c0pud, c1pud,
             fhtop, wflo, ch2o, cice, rhow, dtime, n_hydro=None):
    """Implicit conduction (soil_physics.py:385-490); n_hydro < K leaves thermal-only layers."""
    n, ns = tsoi.shape
    nh = ns if n_hydro is None else n_hydro
    con = np.zeros((n, ns + 1))
    wh = np.zeros((n, ns + 1))
    wh[:, 0] = wflo[:, 0] * ch2o * tsoi[:, 0]
    for k in range(1, ns):
        con[:, k] = 1.0 / (0.5 * (hsoi[k - 1] / np.maximum(consoi[:, k - 1], 1e-12)
                                  + hsoi[k] / np.maximum(consoi[:, k], 1e-12)))
        if k < nh:
            tav = (hsoi[k] * tsoi[:, k - 1] + hsoi[k - 1] * tsoi[:, k]) / (hsoi[k - 1] + hsoi[k])
            wh[:, k] = wflo[:, k] * ch2o * tav
    drain = wflo[:, nh] * ch2o * tsoi[:, nh - 1]
    for k in range(nh, ns + 1):
        wh[:, k] = drain
    r1 = (1.0 - poros) * csoi * rhosoi
    r2 = poros * (1.0 - wisoi) * ch2o * rhow
    r3 = poros * wisoi * cice * rhow
    c0 = c0pud + (r1 + r2 * owsoi + r3) * hsoi[None, :]
    c1 = c1pud + (r1 + r2 * wsoi + r3) * hsoi[None, :]
    d1 = np.zeros((n, ns))
    d2 = np.zeros((n, ns))
    d3 = np.zeros((n, ns))
    r = np.zeros((n, ns))
    for k in range(ns):
        km1, kp1 = max(k - 1, 0), min(k + 1, ns - 1)
        rw = dtime / c1[:, k]
        d1[:, k] = -rw * 1.0 * con[:, k]
        d2[:, k] = 1.0 + rw * 1.0 * (con[:, k] + con[:, k + 1])
        d3[:, k] = -rw * 1.0 * con[:, k + 1]
        r[:, k] = ((c0[:, k] / c1[:, k]) * tsoi[:, k]
                   + rw * (0.0 * con[:, k] * (tsoi[:, km1] - tsoi[:, k])
                           + 0.0 * con[:, k + 1] * (tsoi[:, kp1] - tsoi[:, k])
                           + wh[:, k] - wh[:, k + 1]))
        if k == 0:
            r[:, k] = r[:, k] + (dtime / c1[:, k]) * fhtop
    tn = tridia(d1, d2, d3, r)
    hflo = np.zeros((n, ns + 1))
    hflo[:, 0] = fhtop
    for k in range(ns):
        hflo[:, k + 1] = hflo[:, k] - (c1[:, k] * tn[:, k] - c0[:, k] * otsoi[:, k]) / dtime
    return tn, hflo


def soilheat_with_bedrock(tsoi, tdeep, otsoi, otdeep, owsoi, wsoi, wisoi, poros, consoi, csoi,
                          rhosoi, hsoi, c0pud, c1pud, fhtop, wflo, ch2o, cice, rhow, dtime, bp):
    """Append thermal-only bedrock layers to the conduction solve (bedrock.py:88-153)."""
    n, ns = tsoi.shape
    nb = len(bp.hsoi)
    z = np.zeros((n, nb))
    cat = lambda a, b: np.concatenate([a, b], axis=1)
    tn, hflo = soilheat(cat(tsoi, tdeep), cat(otsoi, otdeep), cat(owsoi, z), cat(wsoi, z),
                        cat(wisoi, z), cat(poros, z), cat(consoi, np.full((n, nb), bp.consoi)),
                        cat(csoi, np.full((n, nb), bp.csoi)),
                        cat(rhosoi, np.full((n, nb), bp.rhosoi)),
                        np.concatenate([np.asarray(hsoi, dtype=np.float64),
                                        np.asarray(bp.hsoi, dtype=np.float64)]),
                        cat(c0pud, z), cat(c1pud, z), fhtop, cat(wflo, z),
                        ch2o, cice, rhow, dtime, n_hydro=ns)
    return tn[:, :ns], hflo[:, :ns + 1], tn[:, ns:]


def freeze_thaw(tsoi, wsoi, wisoi, poros, csoi, rhosoi, hsoi, c1pud, c):
    """Heat-conserving freeze then melt per layer (soil_physics.py:497-560)."""
    eps, tmelt, hfus = c.epsilon, c.tmelt, c.hfus
    tsoi, wsoi, wisoi = tsoi.copy(), wsoi.copy(), wisoi.copy()
    for k in range(tsoi.shape[1]):
        zp = np.maximum(poros[:, k], eps)
        rw = c1pud[:, k] / (c.rhow * hsoi[k]) + (1.0 - zp) * csoi[:, k] * rhosoi[:, k] / c.rhow
        chav = rw + zp * (1.0 - wisoi[:, k]) * wsoi[:, k] * c.ch2o + zp * wisoi[:, k] * c.cice
        zw = np.minimum(1.0, wsoi[:, k])
        dh = chav * (tmelt - tsoi[:, k])
        dw = np.minimum(zp * (1.0 - wisoi[:, k]) * zw, np.maximum(0.0, dh / hfus))
        wik = wisoi[:, k] + dw / zp
        wk = wsoi[:, k] - (dw / zp) * (1.0 - zw) / np.maximum(eps, 1.0 - wik)
        chav = rw + zp * (1.0 - wik) * wk * c.ch2o + zp * wik * c.cice
        tk = tmelt - (dh - hfus * dw) / chav
        dh = chav * (tk - tmelt)
        dw = np.minimum(zp * wik, np.maximum(0.0, dh / hfus))
        wik = wik - dw / zp
        wk = wk + (dw / zp) * (1.0 - wk) / np.maximum(eps, 1.0 - wik)
        chav = rw + zp * (1.0 - wik) * wk * c.ch2o + zp * wik * c.cice
        tk = tmelt + (dh - hfus * dw) / chav
        tsoi[:, k], wsoi[:, k], wisoi[:, k] = tk, wk, wik
    return tsoi, wsoi, wisoi, poros * np.maximum(eps, 1.0 - wisoi)


def puddle_freeze_thaw(t0, wpud, wipud, poros0, wsoi0, wisoi0, csoi0, rhosoi0, hsoi0, c):
    """Puddle phase change against the top layer (soil_physics.py:567-615)."""
    tmelt, hfus = c.tmelt, c.hfus
    rw = ((1.0 - poros0) * csoi0 * rhosoi0 + poros0 * (1.0 - wisoi0) * wsoi0 * c.ch2o * c.rhow
          + poros0 * wisoi0 * c.cice * c.rhow) * hsoi0
    chav = c.ch2o * wpud + c.cice * wipud + rw
    dh = chav * (tmelt - t0)
    dw = np.minimum(wpud, np.maximum(0.0, dh / hfus))
    wipud = wipud + dw
    wpud = wpud - dw
    chav = c.ch2o * wpud + c.cice * wipud + rw
    t0 = tmelt - (dh - hfus * dw) / chav
    dh = chav * (t0 - tmelt)
    dw = np.minimum(wipud, np.maximum(0.0, dh / hfus))
    wipud = wipud - dw
    wpud = wpud + dw
    chav = c.ch2o * wpud + c.cice * wipud + rw
    t0 = tmelt + (dh - hfus * dw) / chav
    return t0, np.maximum(wpud, 0.0), np.maximum(wipud, 0.0)
