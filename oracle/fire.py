"""Per-step fire-weather accumulators (FP64), restating ``python-port/ibis/physics/fire.py``.

Pinned by the F77 formulas the reference transcribes (``fortran/ibis_2.6b4/stats.f:1174-1180``):
fueldry in [0, 1], the running mean reproducing ``fueldry`` at nytimes = 1, and the VPD
from ``esat`` and the vapour pressure of qa.
"""
from __future__ import annotations

import numpy as np

from .thermo import esat


def fueldry(wsoi1, wisoi1, swilt1):
    """Fuel dryness from top-layer water (fire.py:88-110)."""
    water = wisoi1 + (1.0 - wisoi1) * wsoi1
    frac = (water - swilt1) / (1.0 - swilt1)
    return np.minimum(1.0, np.maximum(0.0, -2.0 * (frac - 0.5)))


def update_firefac(firefac, nytimes, wsoi1, wisoi1, swilt1):
    """Running mean of fuel dryness over the year's steps (fire.py:113-142)."""
    n = max(float(nytimes), 1.0)
    return ((n - 1.0) * firefac + fueldry(wsoi1, wisoi1, swilt1)) / n


def update_accumulators(nfire, sdry, wsoi1, wisoi1, swilt1, d_crit):
    """Count steps above d_crit and sum their dryness (fire.py:340-364)."""
    fd = fueldry(wsoi1, wisoi1, swilt1)
    hit = (fd > d_crit).astype(np.float64)
    return nfire + hit, sdry + fd * hit


def update_accumulators_vpd(nfire, sdry, ta, qa, psurf, fp):
    """VPD flammability counters (fire.py:465-512, 583-603)."""
    eact = qa * psurf / (0.622 + 0.378 * qa)
    vpd = np.maximum(esat(ta) - eact, 0.0) / 1000.0
    f = np.clip(vpd / fp.vpd_ref, 0.0, 1.0)
    hit = (f > fp.d_crit_vpd).astype(np.float64)
    return nfire + hit, sdry + f * hit
