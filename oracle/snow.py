"""Snow layers (FP64), restating ``python-port/ibis/physics/snow_physics.py``.

Pinned by: ``vadapt`` conserving total thickness and the thickness-weighted temperature
integral and setting the top layer to hsnotop (``tests/test_snow.py:33-66``); ``snowheat``
leaving an isothermal pack with zero flux unchanged (``tests/test_snow.py:86-104``) and
closing its heat budget (sflo[0] - sflo[-1] = stored heat change); ``snow_melt`` energy
(melt x hfus + sensible = excess heat) and no melt below tmelt (``tests/test_snow.py:152-195``);
``snow_coverage`` conserving the snow volume fi x sum(hsno) (``tests/test_snow.py:207-216``).
"""
from __future__ import annotations

import numpy as np

from .solvers import tridia


def vadapt(hcur, tcur, htop, has_snow, eps):
    """Re-layer the pack, conserving sum(h*T) (snow_physics.py:27-96)."""
    n, nl = hcur.shape
    zold = np.zeros((n, nl + 1))
    for k in range(nl):
        zold[:, k + 1] = zold[:, k] + hcur[:, k]
    ht = zold[:, nl]
    thick = ht >= nl * htop
    h1 = np.where(thick, (ht - htop) / (nl - 1), np.maximum(ht / nl, eps))
    hnew = np.repeat(h1[:, None], nl, axis=1)
    hnew[:, 0] = np.where(thick, htop, h1)
    znew = np.zeros((n, nl + 1))
    for k in range(nl):
        znew[:, k + 1] = znew[:, k] + hnew[:, k]
    tnew = np.zeros((n, nl))
    for k in range(nl):
        za, zb = znew[:, k], znew[:, k + 1]
        zh = np.zeros(n)
        for ko in range(nl):
            ov = np.maximum(np.minimum(zold[:, ko + 1], zb) - np.maximum(zold[:, ko], za), 0.0)
            zh = zh + tcur[:, ko] * ov
        tnew[:, k] = zh / np.maximum(hnew[:, k], eps)
    m = has_snow[:, None]
    return np.where(m, hnew, hcur), np.where(m, tnew, tcur)


def snowheat(tsno, tlsub, fhtop, xl, chl, fi, hsno, rhos, consno, hsoi1, consoi1, tsoi1,
             dtime, cice):
    """Implicit conduction through snow + buried vegetation (snow_physics.py:109-229).

    Returns tsno, tlsub, sflo (N, nsnolay+2).
    """
    n, ns = tsno.shape
    hfake, conbur = 0.01, 2.0
    nt = ns + 1
    temp = np.concatenate([tsno, tlsub[:, None]], axis=1)
    con = np.zeros((n, ns + 2))
    rs = 0.5 / consno
    for k in range(1, ns):
        con[:, k] = 1.0 / (np.maximum(hsno[:, k - 1], hfake) * rs
                           + np.maximum(hsno[:, k], hfake) * rs)
    con[:, ns] = 1.0 / (np.maximum(hsno[:, ns - 1], hfake) * rs + 0.5 * xl / conbur)
    con[:, ns + 1] = 1.0 / (xl * (0.5 / conbur) + 0.5 * hsoi1 / np.maximum(consoi1, 1e-12))
    d1 = np.zeros((n, nt))
    d2 = np.zeros((n, nt))
    d3 = np.zeros((n, nt))
    r = np.zeros((n, nt))
    rdt = dtime / (rhos * cice)
    for k in range(ns):
        km1, kp1 = max(k - 1, 0), min(k + 1, nt - 1)
        dt = rdt / np.maximum(hsno[:, k], hfake)
        d1[:, k] = -dt * 1.0 * con[:, k]
        d2[:, k] = 1.0 + dt * 1.0 * (con[:, k] + con[:, k + 1])
        d3[:, k] = -dt * 1.0 * con[:, k + 1]
        # (1 - RIMP) terms are exactly zero with RIMP = 1
        r[:, k] = temp[:, k] + dt * (0.0 * con[:, k] * (temp[:, km1] - temp[:, k])
                                     + 0.0 * con[:, k + 1] * (temp[:, kp1] - temp[:, k]))
        if k == 0:
            r[:, k] = r[:, k] + dt * fhtop
    k = ns
    dti = xl * chl / dtime
    d1[:, k] = -1.0 * con[:, k]
    d2[:, k] = dti + 1.0 * (con[:, k] + con[:, k + 1])
    r[:, k] = (dti * temp[:, k] + 0.0 * con[:, k] * (temp[:, k - 1] - temp[:, k])
               + con[:, k + 1] * (tsoi1 - 0.0 * temp[:, k]))
    tn = tridia(d1, d2, d3, r)
    sflo = np.zeros((n, ns + 2))
    sflo[:, 0] = fhtop
    rh = rhos * cice / dtime
    for k in range(ns):
        sflo[:, k + 1] = sflo[:, k] - rh * hsno[:, k] * (tn[:, k] - tsno[:, k])
    sflo[:, ns + 1] = sflo[:, ns] - xl * (chl / dtime) * (tn[:, ns] - tlsub)
    has = fi > 0.0
    return (np.where(has[:, None], tn[:, :ns], tsno), np.where(has, tn[:, ns], tlsub), sflo)


def snow_melt(hsno, tsno, fi, rhos, cice, hfus, tmelt):
    """Melt any layer above tmelt (snow_physics.py:236-256)."""
    xh = rhos * hsno * cice * np.maximum(tsno - tmelt, 0.0)
    xm = np.minimum(rhos * hsno, xh / hfus)
    hn = hsno - xm / rhos
    tn = np.minimum(tsno, tmelt)
    zm = fi[:, None] * xm
    zh = fi[:, None] * (xh - hfus * xm)
    zmelt, zheat = zm[:, 0], zh[:, 0]
    for k in range(1, hsno.shape[1]):
        zmelt = zmelt + zm[:, k]
        zheat = zheat + zh[:, k]
    return hn, tn, zmelt, zheat


def snow_coverage(hsno, fi, hsnomin, hsnoruf, fimax, eps):
    """Square-root cover from the pack volume (snow_physics.py:263-290)."""
    ht = hsno[:, 0] + hsno[:, 1] + hsno[:, 2]
    ht = np.where(ht < 1e-10, hsnomin, ht)
    zhh = 0.5 * hsnomin
    zdh = np.maximum(hsnoruf - hsnomin, eps)
    finew = (-zhh + np.sqrt(zhh * zhh + zdh * fi * ht)) / zdh
    finew = np.clip(finew, 0.0, fimax)
    scale = fi / np.maximum(finew, eps)
    return finew, hsno * scale[:, None]
