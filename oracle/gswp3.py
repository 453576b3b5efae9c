"""GSWP3 daily observations -> daily weather of the diurnal generator (FP64) -- TEST
INFRASTRUCTURE: the checker of ``ibis_gswp3_days`` (include/ibis_driver.h) and
of the observed-longwave forcing.  Restates python-port/ibis/io/gswp3_forcing.py:

* ``daily_mean_coszen`` -- ``daily_mean_coszen`` (gswp3_forcing.py:85-116): the 24-hour mean
  of max(0, cos zenith) from the sunset hour angle, Hartmann declination (weather.py:519-527).
* ``rsds_to_cloud``     -- ``rsds_to_cloud`` (:119-165): inverts 0.251 + 0.509 (1 - cloud)
  against the effective solar constant (weather.py:448-455) x the mean cos zenith; 0.5 where
  that mean is below 0.01.
* ``day_weather``       -- ``gswp3_day_to_daily_weather`` (:172-230): NaN -> 0, precipitation
  x 86400, cloud from rsds; plus the day's event from ``precip_timing`` (weather.py:395-441,
  restated in ``oracle.driver``) and the observed longwave.

Pins (tests/test_gswp3.py): the mean cos zenith against a brute-force quadrature of the hourly
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
