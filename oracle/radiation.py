"""Two-stream solar and infrared radiation (FP64), restating ``python-port/ibis/physics/radiation.py``.

Pinned by: the two-stream energy balance abvegd + refld + (transmitted absorbed by the
surface) = 1 for every band (``tests/test_radiation.py:105-127``); the IR balance
firu+firs+firl+firg+firi... against fira - firb (``tests/test_radiation.py:186-206``); zero
solar at night (``tests/test_radiation.py:332-342``); a bare surface returns the soil and
snow albedos; albedos in [0, 1].
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from ._num import ipow

OMEGASNO = (0.9, 0.7)      # radiation.py:62
BETADSNO = 0.5
BETAISNO = 0.5
EMISOIL = 0.95
EMISNOW = 0.95


def twoset(coszen, lai_v, sai_v, fwet_leaf, rliq_leaf, t_leaf, rho, tau, oriev, orieh,
           omegasno, tmelt, upper, sunlit, fwet_stem=None, rliq_stem=None, t_stem=None):
    """Two-stream optical parameters for one band and story (radiation.py:75-159)."""
    orand = 1.0 - oriev - orieh
    omega0 = rho + tau
    ztab = 0.5 - coszen / 3.0
    rw = 1.0 / max(omega0, 1e-12)
    betad = (oriev * 0.5 * (rho + tau) + orieh * rho
             + orand * ((1.0 - ztab) * rho + ztab * tau)) * rw
    betai = (oriev * 0.5 * (rho + tau) + orieh * rho
             + orand * ((2.0 / 3.0) * rho + (1.0 / 3.0) * tau)) * rw
    gdir = (oriev * (2.0 / 3.1415927) * np.sqrt(np.maximum(0.0, 1.0 - coszen * coszen))
            + orieh * coszen + orand * 0.5)
    avmu = np.ones_like(coszen)
    if upper:
        x = lai_v / np.maximum(lai_v + sai_v, 1e-7)
        y = x * (fwet_leaf * (1.0 - rliq_leaf)) + (1.0 - x) * (fwet_stem * (1.0 - rliq_stem))
        sl = np.clip((tmelt - t_leaf) / 0.1, 0.6, 1.0)
        ss = np.clip((tmelt - t_stem) / 0.1, 0.6, 1.0)
        o = (x * sl + (1.0 - x) * ss) * omegasno
    else:
        y = fwet_leaf * (1.0 - rliq_leaf)
        o = omegasno * (0.6 + 0.4 * np.clip((tmelt - t_leaf) / 1.0, 0.0, 1.0))
    sno = y * o
    omega = np.maximum((1.0 - y) * omega0 + sno, 1e-12)
    betad = ((1.0 - y) * omega0 * betad + sno * BETADSNO) / omega
    betai = ((1.0 - y) * omega0 * betai + sno * BETAISNO) / omega
    omega = np.where(sunlit, omega, 0.0)
    betad = np.where(sunlit, betad, 0.0)
    betai = np.where(sunlit, betai, 0.0)
    gdir = np.where(sunlit, gdir, 0.0)
    avmu = np.where(sunlit, avmu, 1.0)
    return omega, betad, betai, avmu, gdir


def twostr(coszen, lai_v, sai_v, asurd, asuri, omega, betad, betai, avmu, gdir, eps,
           sunlit, vis):
    """Two-stream solution for one vegetation layer (radiation.py:180-309)."""
    b = 1.0 - omega * (1.0 - betai)
    c = omega * betai
    tmp0 = b * b - c * c
    q = np.sqrt(np.maximum(0.0, tmp0))
    k = gdir / np.maximum(coszen, 0.01)
    p = avmu * k
    sgn = np.where(p >= q, 1.0, -1.0)
    p = np.where(np.abs(p - q) < 0.001 * p, (1.0 + sgn * 0.001) * p, p)
    c0 = omega * p
    d = c0 * betad
    f = c0 * (1.0 - betad)
    h = q / avmu
    sigma = p * p - tmp0
    ud1 = b - c / np.maximum(asurd, 1e-12)
    ui1 = b - c / np.maximum(asuri, 1e-12)
    ud2 = b - c * asurd
    ui2 = b - c * asuri
    ud3 = f + c * asurd
    xai = np.maximum(lai_v + sai_v, eps)
    s1 = np.exp(-h * xai)
    s2 = np.exp(-k * xai)
    p1, p2, p3, p4 = b + q, b - q, b + p, b - p
    s1s = np.maximum(s1, 1e-30)
    rw = 1.0 / s1s
    dd1 = p1 * (ud1 - q) * rw - p2 * (ud1 + q) * s1
    di1 = p1 * (ui1 - q) * rw - p2 * (ui1 + q) * s1
    dd2 = (ud2 + q) * rw - (ud2 - q) * s1
    di2 = (ui2 + q) * rw - (ui2 - q) * s1
    tiny = lambda v: np.where(np.abs(v) < 1e-30, 1e-30, v)
    dd1, di1, dd2, di2, sigma = tiny(dd1), tiny(di1), tiny(dd2), tiny(di2), tiny(sigma)
    h1 = -d * p4 - c * f
    rw2 = s2 * (d - c - h1 * (ud1 + p) / sigma)
    h2 = (1.0 / dd1) * ((d - h1 * p3 / sigma) * (ud1 - q) / s1s - p2 * rw2)
    h3 = -(1.0 / dd1) * ((d - h1 * p3 / sigma) * (ud1 + q) * s1 - p1 * rw2)
    h4 = -f * p3 - c * d
    rw3 = s2 * (ud3 - h4 * (ud2 - p) / sigma)
    h5 = -(1.0 / dd2) * (h4 * (ud2 + q) / (sigma * s1s) + rw3)
    h6 = (1.0 / dd2) * (h4 * s1 * (ud2 - q) / sigma + rw3)
    h7 = c * (ui1 - q) / (di1 * s1s)
    h8 = -c * s1 * (ui1 + q) / di1
    h9 = (ui2 + q) / (di2 * s1s)
    h10 = -s1 * (ui2 - q) / di2
    fbeldd = s2
    fbelid = h4 / sigma * s2 + h5 * s1 + h6 / s1s
    fbelii = h9 * s1 + h10 / s1s
    refld = h1 / sigma + h2 + h3
    refli = h7 + h8
    absurd = (1.0 - asurd) * fbeldd + (1.0 - asuri) * fbelid
    absuri = (1.0 - asuri) * fbelii
    abvegd = np.maximum(0.0, 1.0 - refld - absurd)
    abvegi = np.maximum(0.0, 1.0 - refli - absuri)
    noveg = xai < eps
    abvegd = np.where(noveg, 0.0, abvegd)
    abvegi = np.where(noveg, 0.0, abvegi)
    term = None
    if vis:
        term = [k * (1.0 + (h4 - h1) / sigma), h * (h5 - h2), h * (h9 - h7), h * (h3 - h6),
                h * (h8 - h10), k, h]
    z = lambda v: np.where(sunlit, v, 0.0)
    return SimpleNamespace(abvegd=z(abvegd), abvegi=z(abvegi), refld=z(refld), refli=z(refli),
                           fbeldd=z(fbeldd), fbelid=z(fbelid), fbelii=z(fbelii), term=term,
                           omega=omega)


def solsur(coszen, wsoi1, wisoi1, tsno1, albsav, albsan, tmelt, sunlit, ib):
    """Soil and snow albedos for one band (radiation.py:316-365)."""
    zw = wsoi1 * (1.0 - wisoi1)
    dinc = 1.0 + np.clip(1.0 - zw / 0.50, 0.0, 1.0)
    base = albsav if ib == 0 else albsan
    albsod = np.minimum(base * dinc, 0.80)
    albsoi = albsod
    a_hi, a_lo = (0.70, 0.90) if ib == 0 else (0.40, 0.60)
    t_shi, t_slo = tmelt, tmelt - 15.0
    x = (a_hi * (tsno1 - t_slo) + a_lo * (t_shi - tsno1)) / (t_shi - t_slo)
    x = np.clip(x, a_hi, a_lo)
    zfac = np.maximum(0.0, 1.5 / (1.0 + 4.0 * coszen) - 0.5)
    albsnd = np.minimum(0.99, x + (1.0 - x) * zfac)
    albsni = np.minimum(1.0, x)
    z = lambda v: np.where(sunlit, v, 0.0)
    return z(albsod), z(albsoi), z(albsnd), z(albsni)


def compute_radiation(coszen, solad, solai, fira, lai, sai, fu, fl, tu, ts, tl, tg, ti, fi,
                      tsno1, wsoi1, wisoi1, albsav, albsan, fw, vr, c):
    """All absorbed solar and IR fluxes for one hour (radiation.py:654-761).

    ``fw`` carries fwetu/fwets/fwetl/rliqu/rliqs/rliql; ``vr`` is VegRadParams.
    """
    eps, tmelt = c.epsilon, c.tmelt
    sunlit = coszen > 0.0
    n = coszen.shape[0]
    out = SimpleNamespace(solu=np.zeros(n), sols=np.zeros(n), soll=np.zeros(n),
                          solg=np.zeros(n), soli=np.zeros(n),
                          asurd=np.zeros((n, 2)), asuri=np.zeros((n, 2)),
                          topparu=np.zeros(n), topparl=np.zeros(n),
                          scalcoefu=np.zeros((n, 4)), scalcoefl=np.zeros((n, 4)),
                          dbeamu=np.zeros(n), dbeaml=np.zeros(n))
    for ib in range(2):
        albsod, albsoi, albsnd, albsni = solsur(coszen, wsoi1, wisoi1, tsno1, albsav, albsan,
                                                tmelt, sunlit, ib)
        om_l = twoset(coszen, lai[:, 0], sai[:, 0], fw.fwetl, fw.rliql, tl,
                      vr.rhoveg[ib, 0], vr.tauveg[ib, 0], vr.oriev[0], vr.orieh[0],
                      OMEGASNO[ib], tmelt, False, sunlit)
        lo = twostr(coszen, lai[:, 0], sai[:, 0], albsod, albsoi, *om_l, eps, sunlit, ib == 0)
        asurd = fl * (1.0 - fi) * lo.refld + (1.0 - fl) * (1.0 - fi) * albsod + fi * albsnd
        asuri = fl * (1.0 - fi) * lo.refli + (1.0 - fl) * (1.0 - fi) * albsoi + fi * albsni
        om_u = twoset(coszen, lai[:, 1], sai[:, 1], fw.fwetu, fw.rliqu, tu,
                      vr.rhoveg[ib, 1], vr.tauveg[ib, 1], vr.oriev[1], vr.orieh[1],
                      OMEGASNO[ib], tmelt, True, sunlit,
                      fwet_stem=fw.fwets, rliq_stem=fw.rliqs, t_stem=ts)
        up = twostr(coszen, lai[:, 1], sai[:, 1], asurd, asuri, *om_u, eps, sunlit, ib == 0)
        out.asurd[:, ib] = np.where(sunlit, fu * up.refld + (1.0 - fu) * asurd, 0.0)
        out.asuri[:, ib] = np.where(sunlit, fu * up.refli + (1.0 - fu) * asuri, 0.0)

        # solarf (radiation.py:450-536)
        sd, si = solad[:, ib], solai[:, ib]
        x = sd * up.abvegd + si * up.abvegi
        y = lai[:, 1] / np.maximum(lai[:, 1] + sai[:, 1], eps)
        solu = x * y
        sols = x * (1.0 - y)
        sol2d = sd * up.fbeldd
        sol2i = sd * up.fbelid + si * up.fbelii
        sol3d = fu * sol2d + (1.0 - fu) * sd
        sol3i = fu * sol2i + (1.0 - fu) * si
        soll = sol3d * lo.abvegd + sol3i * lo.abvegi
        xd = (fl * lo.fbeldd + 1.0 - fl) * sol3d
        xi = fl * (sol3d * lo.fbelid + sol3i * lo.fbelii) + (1.0 - fl) * sol3i
        solg = (1.0 - albsod) * xd + (1.0 - albsoi) * xi
        soli = (1.0 - albsnd) * sol3d + (1.0 - albsni) * sol3i
        z = lambda v: np.where(sunlit, v, 0.0)
        out.solu = out.solu + z(solu)
        out.sols = out.sols + z(sols)
        out.soll = out.soll + z(soll)
        out.solg = out.solg + z(solg)
        out.soli = out.soli + z(soli)
        if ib == 0:
            xaiu = np.maximum(lai[:, 1] + sai[:, 1], eps)
            tu_ = up.term
            sc0 = tu_[0] * sd
            sc1 = tu_[1] * sd + tu_[2] * si
            sc2 = tu_[3] * sd + tu_[4] * si
            sc3 = sc0 + sc1 + sc2
            out.scalcoefu = out.scalcoefu + np.stack([sc0, sc1, sc2, sc3], axis=1)
            out.topparu = out.topparu + sc3 * lai[:, 1] / xaiu
            out.dbeamu = out.dbeamu + (1.0 - up.omega) * tu_[5] * sd
            xail = np.maximum(lai[:, 0] + sai[:, 0], eps)
            tl_ = lo.term
            sl0 = tl_[0] * sol3d
            sl1 = tl_[1] * sol3d + tl_[2] * sol3i
            sl2 = tl_[3] * sol3d + tl_[4] * sol3i
            sl3 = sl0 + sl1 + sl2
            out.scalcoefl = out.scalcoefl + np.stack([sl0, sl1, sl2, sl3], axis=1)
            out.topparl = out.topparl + sl3 * lai[:, 0] / xail
            out.dbeaml = out.dbeaml + (1.0 - lo.omega) * tl_[5] * sol3d
            out.termu6, out.termu7 = tu_[5], tu_[6]
            out.terml6, out.terml7 = tl_[5], tl_[6]
    (out.firu, out.firs, out.firl, out.firg, out.firi, out.firb) = irrad(
        fira, lai, sai, fu, fl, fi, tu, ts, tl, tg, ti, c.stef)
    return out


def irrad(fira, lai, sai, fu, fl, fi, tu, ts, tl, tg, ti, stef):
    """Absorbed infrared fluxes (radiation.py:543-619)."""
    emu = 1.0 - np.exp(-lai[:, 1] / 1.0)
    ems = 1.0 - np.exp(-sai[:, 1] / 1.0)
    eml = 1.0 - np.exp(-(lai[:, 0] + sai[:, 0]) / 1.0)
    emg, emi = EMISOIL, EMISNOW
    tu4, ts4, tl4 = stef * ipow(tu, 4), stef * ipow(ts, 4), stef * ipow(tl, 4)
    tg4, ti4 = stef * ipow(tg, 4), stef * ipow(ti, 4)
    fdown = ((1.0 - fu) * fira
             + fu * ((1.0 - emu) * (1.0 - ems) * fira + emu * (1.0 - ems) * tu4 + ems * ts4))
    fdowng = (1.0 - eml) * fdown + eml * tl4
    fupg = (1.0 - emg) * fdowng + emg * tg4
    fupgb = (1.0 - emg) * fdown + emg * tg4
    fupi = (1.0 - emi) * fdown + emi * ti4
    fup = ((1.0 - fi) * (fl * (eml * tl4 + (1.0 - eml) * fupg) + (1.0 - fl) * fupgb)
           + fi * fupi)
    firb = ((1.0 - fu) * fup
            + fu * ((1.0 - emu) * (1.0 - ems) * fup + emu * tu4 + ems * (1.0 - emu) * ts4))
    firu = emu * ems * ts4 + emu * (1.0 - ems) * fup + emu * fira - 2.0 * emu * tu4
    firs = ems * emu * tu4 + ems * fup + ems * (1.0 - emu) * fira - 2.0 * ems * ts4
    firl = eml * fdown + eml * fupg - 2.0 * eml * tl4
    firg = fl * (fdowng - fupg) + (1.0 - fl) * (fdown - fupgb)
    firi = fdown - fupi
    return firu, firs, firl, firg, firi, firb
