"""GSWP3 transient forcing: the decade-file reader and the month's daily weather on the GPU.

``GSWP3Reader`` mirrors the reference's reader (python-port/ibis/io/gswp3_reader.py): decade
files ``{var}_gswp3_{start}_{end}.nc4`` with (time, lat, lon) variables, a daily time axis
from January 1 of the decade's first year, lat N->S and lon W->E on the 0.5 degree grid;
``read_month`` returns the month of each variable on the land cells, float32 (ndays, npoi),
with values at or above half the fill value as NaN, and falls back to tas +- 2 K for an
unreadable tasmax / tasmin and to NaN for any other unreadable variable, as the reference
does (:166-251).

File format.  The reference opens the files with the netCDF4 library.  This image has no
netCDF-4 / HDF5 library, so the reader opens the netCDF classic and 64-bit-offset formats
(``scipy.io.netcdf_file``, memory-mapped: a month reads only its own records) and raises
``GSWP3FormatError`` for a netCDF-4 (HDF5) file, naming the conversion
(``nccopy -k 64-bit-offset``), instead of taking the unreadable-file fallback: a format the
build cannot read is not a transient failure.  ``scale_factor`` / ``add_offset`` are applied
as the netCDF4 library does by default.

``MonthWeather`` uploads a month to the GPU and converts every day in one launch of
``ibis_gswp3_days`` (include/ibis_driver.h) into the daily weather of the diurnal generator
with the observed longwave (gswp3_forcing.py:168-269, loops_gswp3.py:650-686); it is the
``weather`` argument of ``ibis_amd.daily.run_days`` for the month, whose hourly forcing then
takes fira from the observed rlds (loops_gswp3.py:131-148), and, where the month carries the
tower's hourly shortwave (``rsds_hourly``, ``rsds_hourly_ok``: site runs), solad and solai
from it (``ibis_forcing_observed_sw``, loops_gswp3.py:126-129).
"""
from __future__ import annotations

import ctypes as C
import logging
import os
from datetime import date
from typing import Optional

import numpy as np

GSWP3_VARS = ("tas", "tasmax", "tasmin", "pr", "rsds", "rlds", "huss", "wind", "ps", "rhs")
#: the variables ibis_gswp3_days reads, in the order of its observation planes
OBS_VARS = ("tas", "tasmax", "tasmin", "pr", "rsds", "rlds", "huss", "wind", "ps")
NGSWP3 = 9
NDAILY_OBS = 12
FORCING_OBSERVED_LW = 0x100

_DAYS_IN_MONTH = (31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31)
_DAYS_IN_MONTH_LEAP = (31, 29, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31)

log = logging.getLogger(__name__)


class GSWP3FormatError(RuntimeError):
    """A file in a format this reader does not open (netCDF-4 / HDF5)."""


def _is_leap(year: int) -> bool:
    return (year % 4 == 0 and year % 100 != 0) or year % 400 == 0


def _days_in_month(year: int, month: int) -> int:
    """Days in ``month`` (0-based) of ``year`` (gswp3_reader.py:56-58)."""
    return (_DAYS_IN_MONTH_LEAP if _is_leap(year) else _DAYS_IN_MONTH)[month]


def _decade_bounds(year: int) -> tuple[int, int]:
    """First and last year of the decade file holding ``year``: 1901-1910, 1911-1920, ...
    (gswp3_reader.py:61-65)."""
    start = 1901 + ((year - 1901) // 10) * 10
    return start, start + 9


def _time_index(year: int, month: int, day: int) -> int:
    """Record of (year, 0-based month, 0-based day) in its decade file (:68-81)."""
    return (date(year, month + 1, day + 1) - date(_decade_bounds(year)[0], 1, 1)).days


def _month_time_slice(year: int, month: int) -> tuple[int, int]:
    """Records [i0, i1) of a whole month (:84-88)."""
    i0 = _time_index(year, month, 0)
    return i0, i0 + _days_in_month(year, month)


def julian_day(month: int, day: int, leap: bool = False) -> int:
    """1-based day of the year of 0-based (month, day) (timestepping/calendar.py:38-46)."""
    table = _DAYS_IN_MONTH_LEAP if leap else _DAYS_IN_MONTH
    return int(sum(table[:month])) + day + 1


def _open(path: str):
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] == b"CDF" and magic[3:4] in (b"\x01", b"\x02"):
        from scipy.io import netcdf_file
        return netcdf_file(path, "r", mmap=True, maskandscale=False)
    if magic == b"\x89HDF":
        raise GSWP3FormatError(
            f"{path}: netCDF-4 (HDF5) file; this build reads the netCDF classic and 64-bit-offset "
            f"formats (no HDF5 library in the image): convert it with "
            f"`nccopy -k 64-bit-offset {os.path.basename(path)} out.nc`")
    raise GSWP3FormatError(f"{path}: not a netCDF file (magic {magic!r})")


def _values(var, i0: Optional[int] = None, i1: Optional[int] = None) -> np.ndarray:
    """Records [i0, i1) of a variable as float32, scaled as the netCDF4 library does."""
    raw = var[i0:i1] if i0 is not None else var[:]
    a = np.array(raw, dtype=np.float64)
    scale = getattr(var, "scale_factor", None)
    offset = getattr(var, "add_offset", None)
    if scale is not None:
        a = a * float(np.asarray(scale).ravel()[0])
    if offset is not None:
        a = a + float(np.asarray(offset).ravel()[0])
    return a.astype(np.float32)


class GSWP3Reader:
    """Reader of GSWP3 decade files (gswp3_reader.py:95-343); handles stay open until
    ``close`` (or the end of a ``with`` block)."""

    def __init__(self, data_dir: str, variables: tuple[str, ...] = GSWP3_VARS):
        self._handles: dict[tuple[str, int], object] = {}
        if not os.path.isdir(data_dir):
            raise FileNotFoundError(f"GSWP3 data directory not found: {data_dir}")
        self.data_dir = data_dir
        self.variables = tuple(variables)

    def _decade_path(self, var: str, year: int) -> str:
        s, e = _decade_bounds(year)
        return os.path.join(self.data_dir, f"{var}_gswp3_{s}_{e}.nc4")

    def _get_handle(self, var: str, year: int):
        key = (var, _decade_bounds(year)[0])
        if key not in self._handles:
            path = self._decade_path(var, year)
            if not os.path.exists(path):
                raise FileNotFoundError(f"GSWP3 file not found: {path}")
            self._handles[key] = _open(path)
        return self._handles[key]

    def _drop_handle(self, var: str, year: int):
        ds = self._handles.pop((var, _decade_bounds(year)[0]), None)
        if ds is not None:
            try:
                ds.close()
            except Exception:          # a stale handle may not close cleanly
                pass

    def close(self) -> None:
        for key in list(self._handles):
            self._drop_handle(key[0], key[1])

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *args):
        self.close()

    def _check_files(self, year: int) -> list[str]:
        return [p for p in (self._decade_path(v, year) for v in self.variables)
                if not os.path.exists(p)]

    def read_month(self, year: int, month: int, latindex, lonindex,
                   fill_value: float = 1e20) -> dict:
        """One month (0-based) on the land points (latindex[k], lonindex[k]): each variable
        (ndays, npoi) float32, plus ``ndays``, ``year`` and ``month`` (gswp3_reader.py:166-251).
        A failed read is retried once on a fresh handle (stale network handles)."""
        latindex = np.asarray(latindex)
        lonindex = np.asarray(lonindex)
        i0, i1 = _month_time_slice(year, month)
        ndays = i1 - i0
        out: dict = {}
        for var in self.variables:
            land = None
            for attempt in range(2):
                try:
                    ds = self._get_handle(var, year)
                    grid = _values(ds.variables[var], i0, i1)           # (ndays, nlat, nlon)
                    land = grid[:, latindex, lonindex]                 # paired indexing
                    del grid
                    break
                except GSWP3FormatError:
                    raise
                except Exception as exc:
                    self._drop_handle(var, year)
                    if attempt == 1:
                        if var in ("tasmax", "tasmin") and "tas" in out:
                            log.warning("Failed reading %s year %d month %d: %s. Using tas as "
                                        "fallback.", var, year, month, exc)
                            land = out["tas"] + (2.0 if var == "tasmax" else -2.0)
                        else:
                            log.warning("Failed reading %s year %d month %d: %s. Filling with "
                                        "NaN.", var, year, month, exc)
                            land = np.full((ndays, latindex.size), np.nan, np.float32)
            land = np.array(land, dtype=np.float32)
            land[land >= fill_value * 0.5] = np.nan
            out[var] = land
        out["ndays"], out["year"], out["month"] = ndays, year, month
        return out

    def read_year(self, year: int, latindex, lonindex) -> dict:
        """Twelve months concatenated along time (gswp3_reader.py:253-278)."""
        months = [self.read_month(year, m, latindex, lonindex) for m in range(12)]
        out = {v: np.concatenate([m[v] for m in months], axis=0) for v in self.variables}
        out["year"] = year
        return out

    def get_grid(self) -> dict:
        """lat (N->S) and lon (W->E) of the first decade file found (:280-300)."""
        for var in self.variables:
            for year in (1980, 1990, 2000, 1901):
                if os.path.exists(self._decade_path(var, year)):
                    ds = self._get_handle(var, year)
                    return {"lat": _values(ds.variables["lat"]).astype(np.float64),
                            "lon": _values(ds.variables["lon"]).astype(np.float64)}
        raise RuntimeError("Cannot find any GSWP3 file to read grid from.")

    def build_land_indices(self, land_mask, lat=None, lon=None):
        """Row and column of every land cell of a (nlat, nlon) mask on the GSWP3 grid, in
        row-major order, and their latitude and longitude (:302-343)."""
        if lat is None or lon is None:
            g = self.get_grid()
            lat, lon = g["lat"], g["lon"]
        ij = np.argwhere(np.asarray(land_mask) > 0)
        li = ij[:, 0].astype(np.int32)
        lj = ij[:, 1].astype(np.int32)
        return li, lj, np.asarray(lat)[li], np.asarray(lon)[lj]


def _lib():
    from . import driver
    return driver._lib_driver()


class MonthWeather:
    """A GSWP3 month as the GPU daily weather of the diurnal generator.

    ``month``: the dict of ``GSWP3Reader.read_month``; ``lat_deg``: (npoi,) latitudes.  The
    observations are copied to the device once and converted in one ``ibis_gswp3_days``
    launch into ``self.days``, (ndays, NDAILY_OBS, pitch) in ``dtype``.  ``seed`` keys the
    precipitation events' draws (per cell and day of the run: ``day0`` is the month's first
    day counted from the run's start)."""

    observed_lw = True

    def __init__(self, month: dict, lat_deg, seed: int, pitch: Optional[int] = None,
                 dtime: float = 3600.0, day0: int = 0, cell0: int = 0, device="cuda",
                 dtype=None):
        import torch
        from . import abi
        dtype = torch.float32 if dtype is None else dtype
        self.lat = np.asarray(lat_deg, dtype=np.float64)
        n = self.lat.size
        self.ndays = int(month["ndays"])
        year, mon = int(month["year"]), int(month["month"])
        self.jday0 = julian_day(mon, 0, _is_leap(year))
        self.seed, self.dtime = int(seed), float(dtime)
        pitch = n if pitch is None else int(pitch)
        obs = np.zeros((NGSWP3, self.ndays, n), np.float32)
        for v, name in enumerate(OBS_VARS):
            a = np.asarray(month[name], np.float32)
            if a.shape != (self.ndays, n):
                raise ValueError(f"{name}: shape {a.shape}, expected {(self.ndays, n)}")
            obs[v] = a
        self.obs = torch.as_tensor(obs, device=device)
        self.lat_dev = torch.as_tensor(self.lat, device=device)
        self.days = torch.zeros((self.ndays, NDAILY_OBS, pitch), dtype=dtype, device=device)
        self._code = abi.PREC_F64 if dtype == torch.float64 else abi.PREC_MIXED
        self._arena = abi.IbisArena(self.days.data_ptr(), pitch)
        self.n, self.cell0, self.day0 = n, int(cell0), int(day0)
        self.convert()
        # the tower's observed hourly shortwave where the month carries one (site runs,
        # loops_gswp3.py:677-686): (ndays, steps per day, npoi) W m-2 and an (npoi,) mask
        self.sw_hourly = self.sw_use = None
        if month.get("rsds_hourly") is not None:
            nsteps = int(86400.0 / self.dtime)
            sw = np.asarray(month["rsds_hourly"], np.float64)
            if sw.shape != (self.ndays, nsteps, n):
                raise ValueError(f"rsds_hourly: shape {sw.shape}, expected {(self.ndays, nsteps, n)}")
            use = month.get("rsds_hourly_ok")
            use = np.ones(n, bool) if use is None else np.asarray(use, bool)
            if use.shape != (n,):
                raise ValueError(f"rsds_hourly_ok: shape {use.shape}, expected {(n,)}")
            pad = np.zeros((self.ndays, nsteps, pitch))
            pad[:, :, :n] = sw
            self.sw_hourly = torch.as_tensor(pad, device=device).to(dtype)
            mask = np.zeros(pitch, np.uint8)
            mask[:n] = use
            self.sw_use = torch.as_tensor(mask, device=device)

    def convert(self, stream=None):
        """(Re)build ``self.days`` from the observations on the device: one launch."""
        import torch
        from . import _lib as L
        st = stream if stream is not None else torch.cuda.current_stream(self.days.device)
        rc = _lib().ibis_gswp3_days(C.c_void_p(self.obs.data_ptr()), C.c_int64(self.n),
                                    C.c_int(self.ndays), C.c_int(self.jday0),
                                    C.c_void_p(self.lat_dev.data_ptr()), C.c_int64(self.n),
                                    C.c_double(self.dtime), C.c_uint64(self.seed),
                                    C.c_int64(self.cell0), C.c_int64(self.day0),
                                    C.byref(self._arena), C.c_int(self._code),
                                    C.c_void_p(st.cuda_stream))
        L.check(rc, "ibis_gswp3_days")

    def device_days(self, day0: int, days: int, state) -> list:
        """Daily-weather blocks of days [day0, day0 + days) of the month, for run_days."""
        if day0 < 0 or day0 + days > self.ndays:
            raise ValueError(f"days {day0}..{day0 + days - 1} outside the month's {self.ndays}")
        if self.days.dtype != state.dtype or self.days.shape[2] != state.pitch:
            raise ValueError("month weather and state must share dtype and pitch")
        return [self.days[d] for d in range(day0, day0 + days)]


# ---- measurement ---------------------------------------------------------------------------

#: algorithmic HBM bytes per cell-day of ibis_gswp3_days with float32 storage: 9 float32
#: observations read, 12 daily planes written (the latitude, 8 B per cell, is read once per day
#: block and stays in L2)
BYTES_PER_CELL_DAY = 4 * NGSWP3 + 4 * NDAILY_OBS


def synthetic_month(n: int, year: int, month: int, seed: int) -> dict:
    """A month of plausible GSWP3 values on ``n`` cells (float32, native units), as
    ``GSWP3Reader.read_month`` returns them."""
    nd = _days_in_month(year, month)
    rng = np.random.default_rng(seed)
    tas = rng.uniform(230.0, 310.0, (nd, n))
    m = {"tas": tas, "tasmax": tas + rng.uniform(1, 9, (nd, n)),
         "tasmin": tas - rng.uniform(1, 9, (nd, n)), "rsds": rng.uniform(0, 350, (nd, n)),
         "rlds": rng.uniform(150, 450, (nd, n)), "huss": rng.uniform(1e-4, 0.02, (nd, n)),
         "wind": rng.uniform(0.5, 12, (nd, n)), "ps": rng.uniform(6e4, 1.04e5, (nd, n)),
         "rhs": rng.uniform(20, 100, (nd, n)),
         "pr": np.where(rng.random((nd, n)) < 0.4, rng.exponential(8.0 / 86400.0, (nd, n)), 0.0)}
    m = {k: v.astype(np.float32) for k, v in m.items()}
    m.update(ndays=nd, year=year, month=month)
    return m


def main(argv=None):
    """Time ibis_gswp3_days on one month of ``--cells`` cells: HIP events on the launch
    stream around ``--reps`` launches; prints one JSON line."""
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--cells", type=int, default=95333)
    ap.add_argument("--year", type=int, default=1980)
    ap.add_argument("--month", type=int, default=0)
    ap.add_argument("--reps", type=int, default=200)
    a = ap.parse_args(argv)
    lat = np.linspace(-56.0, 83.0, a.cells)
    w = MonthWeather(synthetic_month(a.cells, a.year, a.month, 1), lat, seed=7)
    st = torch.cuda.current_stream()
    for _ in range(10):
        w.convert(st)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(a.reps):
        w.convert(st)
    e1.record(st)
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / a.reps
    traffic = BYTES_PER_CELL_DAY * a.cells * w.ndays
    gbs = traffic / (us * 1e-6) / 1e9
    print(json.dumps({"kernel": "ibis_gswp3_days", "cells": a.cells, "days": w.ndays,
                      "us_per_launch": round(us, 2), "bytes_per_launch": traffic,
                      "achieved_GBps": round(gbs, 1), "peak_GBps": 8000.0,
                      "frac": round(gbs / 8000.0, 3),
                      "finite": bool(torch.isfinite(w.days).all())}), flush=True)


if __name__ == "__main__":
    main()
