"""GSWP3 daily observations -> daily weather of the diurnal generator (FP64) -- TEST
INFRASTRUCTURE: the checker of ``ibis_gswp3_days`` (include/ibis_driver.h) and
of the observed-longwave forcing.  Restates python-port/ibis/io/gswp3_forcing.py:

* ``daily_mean_coszen`` -- ``daily_mean_coszen`` (gswp3_forcing.py:85-116): the 24-hour mean
  of max(0, cos zenith) from the sunset hour angle, Hartmann declination (weather.py:519-527).
* ``rsds_to_cloud``     -- ``rsds_to_cloud`` (:119-165): inverts 0.251 + 0.509 (1 - cloud)
  against the effective solar constant (weather.py:448-455) x the mean cos zenith; 0.5 where
  that mean is below 0.01.
* ``observed_shortwave`` -- ``observed_shortwave`` (weather.py:458-480): the tower's hourly
  total split by the Spitters et al. (1986) diffuse fraction on the clearness index and into
  the two bands; the generator's values where the mask is off.
* ``day_weather``       -- ``gswp3_day_to_daily_weather`` (:172-230): NaN -> 0, precipitation
  x 86400, cloud from rsds; plus the day's event from ``precip_timing`` (weather.py:395-441,
  restated in ``oracle.driver``) and the observed longwave.

Pins (tests/test_gswp3.py): the shortwave split conserves the observed total, is all
diffuse with the sun down or under heavy cloud (kt <= 0.22) and leaves the fraction
R = 0.847 - 1.61 cos z + 1.04 cos^2 z diffuse under a clear sky; the mean cos zenith against a brute-force quadrature of the hourly
cos zenith over the day at every latitude band, and its closed forms at the equator, in
polar day and in polar night; the cloud inversion against the generator's forward
transmissivity (cloud -> rsds -> cloud), its two clips and the polar-night default.
"""
from __future__ import annotations

import numpy as np

from .driver import precip_timing

PI = np.pi
VARS = ("tas", "tasmax", "tasmin", "pr", "rsds", "rlds", "huss", "wind", "ps")
DAILY_OBS = ("td", "tmax", "tmin", "qd", "ud", "precip", "cloud", "psurf", "plens", "startp",
             "endp", "rlds")


def declination(jday):
    o = 2.0 * PI * jday / 365.2425
    return (0.006918 - 0.399912 * np.cos(o) + 0.070257 * np.sin(o) - 0.006758 * np.cos(2 * o)
            + 0.000907 * np.sin(2 * o) - 0.002697 * np.cos(3 * o) + 0.001480 * np.sin(3 * o))


def solar_constant(jday):
    o = 2.0 * PI * jday / 365.2425
    return 1370.0 * (1.000110 + 0.034221 * np.cos(o) + 0.001280 * np.sin(o)
                     + 0.000719 * np.cos(2 * o) + 0.000077 * np.sin(2 * o))


def daily_mean_coszen(lat_rad, jday):
    """gswp3_forcing.py:85-116."""
    d = declination(jday)
    hs = np.arccos(np.clip(-np.tan(lat_rad) * np.tan(d), -1.0, 1.0))
    m = (np.sin(lat_rad) * np.sin(d) * hs + np.cos(lat_rad) * np.cos(d) * np.sin(hs)) / PI
    return np.maximum(0.0, m)


def rsds_to_cloud(rsds, lat_rad, jday, min_coszen=0.01):
    """gswp3_forcing.py:119-165."""
    mcz = daily_mean_coszen(lat_rad, jday)
    night = mcz < min_coszen
    trans = np.clip(rsds / np.where(night, 1.0, solar_constant(jday) * mcz), 0.0, 1.0)
    cloud = np.clip(1.0 - (trans - 0.251) / 0.509, 0.0, 1.0)
    return np.where(night, 0.5, cloud)


def day_weather(obs: dict, lat_deg, jday, dtime, u_len, u_start, store=np.float64) -> dict:
    """One day: ``obs`` maps the VARS names to (N,) arrays.  ``store`` is the storage type the
    event length sees the precipitation in (gswp3_forcing.py:206-216)."""
    g = {k: np.where(np.isnan(np.asarray(obs[k], np.float64)), 0.0,
                     np.asarray(obs[k], np.float64)) for k in VARS}
    precip = np.asarray(g["pr"] * 86400.0, store).astype(np.float64)
    plens, startp, endp = precip_timing(precip, dtime, u_len, u_start)
    return dict(td=g["tas"], tmax=g["tasmax"], tmin=g["tasmin"], qd=g["huss"], ud=g["wind"],
                precip=precip, cloud=rsds_to_cloud(g["rsds"], np.deg2rad(lat_deg), jday),
                psurf=g["ps"], plens=plens, startp=startp, endp=endp, rlds=g["rlds"])


def observed_shortwave(sw, use, coszen, jday, solad_gen, solai_gen):
    """weather.py:458-480; band arrays (N, 2)."""
    kt = np.clip(sw / np.maximum(solar_constant(jday) * coszen, 1.0), 0.0, 1.0)
    r = 0.847 - 1.61 * coszen + 1.04 * coszen ** 2
    kk = (1.47 - r) / 1.66
    fd = np.where(kt <= 0.22, 1.0, np.where(kt <= 0.35, 1.0 - 6.4 * (kt - 0.22) ** 2,
                                            np.where(kt <= kk, 1.47 - 1.66 * kt, r)))
    fd = np.where(coszen <= 0.0, 1.0, np.clip(fd, 0.0, 1.0))
    frac = np.array([0.46, 0.54])
    s = np.maximum(sw, 0.0)
    solad = s[:, None] * frac[None] * (1.0 - fd)[:, None]
    solai = s[:, None] * frac[None] * fd[:, None]
    u = np.asarray(use, bool)[:, None]
    return np.where(u, solad, solad_gen), np.where(u, solai, solai_gen)

