"""Saturation thermodynamics (FP64), restating ``python-port/ibis/physics/thermo.py``.

Pinned by: esat(273.16 K) = 610.78 Pa exactly (liquid base constant, ``thermo.py:88``;
``tests/test_thermo.py:14-17``); monotonicity; desat against a central difference of esat
(the Lowe derivative polynomials approximate the derivative of the SVP polynomial to
<1 %); hvapf(tmelt, tmelt) = hvap (``tests/test_f77_verification.py:85-90``); qsat
against its closed form.
"""
from __future__ import annotations

import numpy as np

TMELT_POLY = 273.16

# Lowe (1977) polynomial coefficients, ascending powers (thermo.py:29-62).
ASAT = (6.1078000, 4.4365185e-1, 1.4289458e-2, 2.6506485e-4, 3.0312404e-6, 2.0340809e-8,
        6.1368209e-11)
BSAT = (6.1091780, 5.0346990e-1, 1.8860134e-2, 4.1762237e-4, 5.8247203e-6, 4.8388032e-8,
        1.8388269e-10)
CSAT = (4.4381000e-1, 2.8570026e-2, 7.9380540e-4, 1.2152151e-5, 1.0365614e-7, 3.5324218e-10,
        -7.0902448e-13)
DSAT = (5.0303052e-1, 3.7732550e-2, 1.2679954e-3, 2.4775631e-5, 3.0056931e-7, 2.1585425e-9,
        7.1310977e-12)


def _horner_tail(x, c):
    # x*(c1 + x*(c2 + ... + x*c6)), the order of thermo.py:85-90
    acc = c[6]
    for j in (5, 4, 3, 2, 1):
        acc = c[j] + x * acc
    return x * acc


def _split(t):
    tl = np.clip(t - TMELT_POLY, 0.0, 100.0)
    ti = np.clip(t - TMELT_POLY, -60.0, 0.0)
    return tl, ti


def esat(t):
    """Saturation vapour pressure, Pa (thermo.py:65-97): both branches summed, one is zero."""
    t = np.asarray(t, dtype=np.float64)
    tl, ti = _split(t)
    a0 = np.where(t >= TMELT_POLY, ASAT[0], BSAT[0])
    return 100.0 * (a0 + _horner_tail(tl, ASAT) + _horner_tail(ti, BSAT))


def desat(t):
    """d esat / dT, Pa/K, from Lowe's own derivative polynomials (thermo.py:100-126)."""
    t = np.asarray(t, dtype=np.float64)
    tl, ti = _split(t)
    c0 = np.where(t >= TMELT_POLY, CSAT[0], DSAT[0])
    return 100.0 * (c0 + _horner_tail(tl, CSAT) + _horner_tail(ti, DSAT))


def qsat(e, p):
    """Saturation specific humidity (thermo.py:129-149)."""
    return 0.622 * e / np.maximum(p - (1.0 - 0.622) * e, 0.622 * e)


def dqsat(t, q1, e=None):
    """d qsat / dT (thermo.py:152-174)."""
    if e is None:
        e = esat(t)
    return desat(t) * q1 * (1.0 + q1 * (1.0 / 0.622 - 1.0)) / e


def hvapf(t, tair, c):
    """Latent heat of vaporisation at surface temperature t (thermo.py:177-199)."""
    return c.hvap + c.cvap * (tair - c.tmelt) - c.ch2o * (t - c.tmelt)


def hsubf(t, tair, c):
    """Latent heat of sublimation at surface temperature t (thermo.py:202-224)."""
    return c.hsub + c.cvap * (tair - c.tmelt) - c.cice * (t - c.tmelt)
