"""GSWP3 forcing on the CPU: the oracle's solar geometry and cloud inversion pinned against
closed forms and brute force; the per-cell conversion of ``ibis_gswp3_days`` (built for the
host) against the oracle; the decade-file reader (python-port/ibis/io/gswp3_reader.py) on
synthetic netCDF files written here.  No reference test covers this path: the fixtures are
synthetic, and the pins are the mathematics and the date arithmetic."""
import ctypes as C
import logging
import os
from datetime import date

import numpy as np
import pytest

from ibis_amd import abi, gswp3 as G, synth
from oracle import gswp3 as OG

from _helpers import HOSTCHECK


# ---- oracle pins ------------------------------------------------------------------------

@pytest.mark.parametrize("jday", [1, 80, 172, 264, 355])
def test_mean_coszen_is_the_days_mean_of_the_clipped_cosine(jday):
    lat = np.deg2rad(np.arange(-89.75, 90.0, 2.5))
    d = OG.declination(jday)
    h = (np.arange(40000) + 0.5) / 40000 * 2 * np.pi - np.pi      # hour angle, midpoints
    cz = np.sin(lat)[:, None] * np.sin(d) + np.cos(lat)[:, None] * np.cos(d) * np.cos(h)[None]
    brute = np.maximum(cz, 0.0).mean(axis=1)
    np.testing.assert_allclose(OG.daily_mean_coszen(lat, jday), brute, atol=2e-8)


def test_mean_coszen_closed_forms():
    d = OG.declination(172)
    assert OG.daily_mean_coszen(np.array([0.0]), 172)[0] == pytest.approx(np.cos(d) / np.pi,
                                                                          rel=1e-14)
    polar = np.deg2rad(85.0)                    # sun never sets: the mean of the whole cosine
    assert OG.daily_mean_coszen(np.array([polar]), 172)[0] == pytest.approx(
        np.sin(polar) * np.sin(d), rel=1e-14)
    assert OG.daily_mean_coszen(np.array([polar]), 355)[0] == 0.0      # polar night
    assert OG.rsds_to_cloud(np.array([50.0]), np.array([polar]), 355)[0] == 0.5


def test_cloud_inverts_the_generators_transmissivity():
    lat = np.deg2rad(np.linspace(-60, 60, 25))
    for jday in (15, 172, 300):
        cloud = np.linspace(0.0, 1.0, 25)
        rsds = OG.solar_constant(jday) * OG.daily_mean_coszen(lat, jday) * \
            (0.251 + 0.509 * (1.0 - cloud))
        np.testing.assert_allclose(OG.rsds_to_cloud(rsds, lat, jday), cloud, atol=1e-12)
        # the two clips: no light reads as full cloud, more than the top of the atmosphere as none
        assert (OG.rsds_to_cloud(np.zeros_like(lat), lat, jday) == 1.0).all()
        top = 2.0 * OG.solar_constant(jday) * OG.daily_mean_coszen(lat, jday)
        assert (OG.rsds_to_cloud(top, lat, jday) == 0.0).all()


def test_event_length_covers_heavy_days_and_stays_in_the_day():
    u = np.full(3, 0.0)
    plens, startp, endp = OG.precip_timing(np.array([0.0, 100.0, 1000.0]), 3600.0, u, u)
    np.testing.assert_array_equal(plens, [4 * 3600.0, 7 * 3600.0, 24 * 3600.0])
    assert (endp <= 86400.0).all() and (startp == 0.0).all()


# ---- the kernel's per-cell code (host build) against the oracle ---------------------------

def _obs(n, ndays, seed):
    rng = np.random.default_rng(seed)
    o = {"tas": rng.uniform(230, 310, (ndays, n)), "rsds": rng.uniform(0, 400, (ndays, n)),
         "rlds": rng.uniform(150, 450, (ndays, n)), "huss": rng.uniform(1e-4, 0.02, (ndays, n)),
         "wind": rng.uniform(0.5, 12, (ndays, n)), "ps": rng.uniform(6e4, 1.04e5, (ndays, n)),
         "pr": np.where(rng.random((ndays, n)) < 0.4,
                        rng.exponential(8.0 / 86400.0, (ndays, n)), 0.0)}
    o["tasmax"] = o["tas"] + rng.uniform(1, 9, (ndays, n))
    o["tasmin"] = o["tas"] - rng.uniform(1, 9, (ndays, n))
    o["pr"][0, :4] = [0.0, 15.0 / 86400.0, 400.0 / 86400.0, np.nan]   # event-length edges, a gap
    o["rsds"][0, 4] = np.nan
    return {k: v.astype(np.float32) for k, v in o.items()}


def _host_days(obs, lat, jday0, seed, day0, prec, dtime=3600.0):
    n = lat.size
    ndays = obs["tas"].shape[0]
    planes = np.ascontiguousarray(np.stack([obs[k] for k in G.OBS_VARS]), np.float32)
    dt = np.float64 if prec == abi.PREC_F64 else np.float32
    out = np.zeros((ndays, G.NDAILY_OBS, n), dt)
    rc = C.CDLL(HOSTCHECK).ibis_hostcheck_gswp3_days(
        planes.ctypes.data_as(C.c_void_p), C.c_int64(n), C.c_int(ndays), C.c_int(jday0),
        lat.ctypes.data_as(C.c_void_p), C.c_int64(n), C.c_double(dtime), C.c_uint64(seed),
        C.c_int64(0), C.c_int64(day0), out.ctypes.data_as(C.c_void_p), C.c_int64(n),
        C.c_int(prec))
    assert rc == 0
    return out


def oracle_days(obs, lat, jday0, seed, day0, store, dtime=3600.0):
    n = lat.size
    cells = np.arange(n, dtype=np.uint64)
    out = []
    for d in range(obs["tas"].shape[0]):
        ul = synth.uniform(seed, 28, cells, day0 + d)
        us = synth.uniform(seed, 29, cells, day0 + d)
        w = OG.day_weather({k: obs[k][d] for k in OG.VARS}, lat, jday0 + d, dtime, ul, us, store)
        out.append(np.stack([w[k] for k in OG.DAILY_OBS]))
    return np.stack(out)


@pytest.mark.parametrize("prec", [abi.PREC_F64, abi.PREC_MIXED])
def test_host_build_matches_the_oracle(prec):
    n, ndays = 700, 5
    lat = np.linspace(-89.75, 89.75, n)
    obs = _obs(n, ndays, 5)
    store = np.float64 if prec == abi.PREC_F64 else np.float32
    got = _host_days(obs, lat, 60, 9, 31, prec)
    ref = oracle_days(obs, lat, 60, 9, 31, store)
    for j, k in enumerate(OG.DAILY_OBS):
        r = ref[:, j]
        if prec == abi.PREC_F64:
            np.testing.assert_allclose(got[:, j], r, rtol=1e-13, atol=1e-300, err_msg=k)
        else:
            np.testing.assert_allclose(got[:, j], r.astype(np.float32), rtol=2.5e-7, atol=0,
                                       err_msg=k)
    assert np.isfinite(got).all()
    assert got[0, 5, 3] == 0.0 and got[0, 6, 4] == 1.0          # NaN precipitation / shortwave
    assert got[0, 8, 2] == 86400.0                              # 400 mm: the whole day


# ---- the decade-file reader on synthetic netCDF files -------------------------------------

NLAT, NLON = 4, 6


def _code(t, i, j):
    """The value written at record t, row i, column j (exact in float32)."""
    return t + 10000.0 * i + 100000.0 * j


def _write(path, var, start, scale=None, fill_at=None):
    from scipy.io import netcdf_file
    nt = (date(start + 10, 1, 1) - date(start, 1, 1)).days
    t = np.arange(nt)[:, None, None]
    i = np.arange(NLAT)[None, :, None]
    j = np.arange(NLON)[None, None, :]
    data = _code(t, i, j).astype(np.float32)
    if fill_at is not None:
        data[fill_at] = 1e20
    with netcdf_file(path, "w") as f:
        f.createDimension("time", None)
        f.createDimension("lat", NLAT)
        f.createDimension("lon", NLON)
        la = f.createVariable("lat", "f4", ("lat",))
        la[:] = 89.75 - 0.5 * np.arange(NLAT)
        lo = f.createVariable("lon", "f4", ("lon",))
        lo[:] = -179.75 + 0.5 * np.arange(NLON)
        if scale is None:
            v = f.createVariable(var, "f4", ("time", "lat", "lon"))
            v[:] = data
        else:                         # packed: stored = (value - offset) / scale
            v = f.createVariable(var, "f8", ("time", "lat", "lon"))
            v.scale_factor, v.add_offset = scale
            v[:] = (data.astype(np.float64) - scale[1]) / scale[0]


@pytest.fixture(scope="module")
def gdir(tmp_path_factory):
    d = tmp_path_factory.mktemp("gswp3")
    for var in G.GSWP3_VARS:
        if var == "tasmax":
            continue                                  # missing: the tas + 2 K fallback
        _write(os.path.join(d, f"{var}_gswp3_1971_1980.nc4"), var, 1971,
               scale=(0.5, 1000.0) if var == "ps" else None,
               fill_at=(3347 + 2, 1, 2) if var == "tas" else None)
    with open(os.path.join(d, "tas_gswp3_1981_1990.nc4"), "wb") as fh:
        fh.write(b"\x89HDF\r\n\x1a\n" + bytes(64))      # a netCDF-4 (HDF5) file
    return str(d)


def test_decade_and_record_arithmetic():
    assert G._decade_bounds(1901) == (1901, 1910) and G._decade_bounds(1980) == (1971, 1980)
    assert G._decade_bounds(1981) == (1981, 1990)
    assert G._time_index(1980, 2, 0) == (date(1980, 3, 1) - date(1971, 1, 1)).days
    assert G._month_time_slice(1980, 1) == (3318, 3347)                    # leap February
    assert G._month_time_slice(1979, 1)[1] - G._month_time_slice(1979, 1)[0] == 28
    assert [G.julian_day(m, 0, True) for m in (0, 2, 11)] == [1, 61, 336]
    assert G.julian_day(11, 30, False) == 365


def test_read_month_gathers_the_land_points(gdir, caplog):
    li = np.array([0, 3, 1, 2, 1])
    lj = np.array([0, 5, 2, 4, 2])
    with G.GSWP3Reader(gdir) as r, caplog.at_level(logging.WARNING):
        m = r.read_month(1980, 2, li, lj)
    i0 = G._time_index(1980, 2, 0)
    assert m["ndays"] == 31 and m["year"] == 1980 and m["month"] == 2
    want = _code(np.arange(i0, i0 + 31)[:, None], li[None], lj[None]).astype(np.float32)
    np.testing.assert_array_equal(m["pr"], want)
    assert m["pr"].dtype == np.float32 and m["pr"].shape == (31, 5)
    np.testing.assert_allclose(m["ps"], want, rtol=1e-6)                 # scale and offset
    tas = want.copy()
    tas[2, [2, 4]] = np.nan                       # the fill value at record i0 + 2, (1, 2)
    np.testing.assert_array_equal(m["tas"], tas)
    np.testing.assert_array_equal(m["tasmax"], tas + 2.0)                 # missing file
    assert any("tasmax" in rec.message for rec in caplog.records)


def test_read_year_and_grid(gdir):
    r = G.GSWP3Reader(gdir, variables=("pr", "rlds"))
    y = r.read_year(1980, np.array([0]), np.array([1]))
    assert y["pr"].shape == (366, 1) and y["year"] == 1980
    i0 = G._time_index(1980, 0, 0)
    np.testing.assert_array_equal(y["rlds"][:, 0], _code(np.arange(i0, i0 + 366), 0, 1))
    g = r.get_grid()
    np.testing.assert_allclose(g["lat"], 89.75 - 0.5 * np.arange(NLAT))
    mask = np.zeros((NLAT, NLON))
    mask[1, 3] = mask[2, 0] = 1
    li, lj, la, lo = r.build_land_indices(mask)
    assert list(li) == [1, 2] and list(lj) == [3, 0]
    np.testing.assert_allclose(la, [89.25, 88.75])
    np.testing.assert_allclose(lo, [-178.25, -179.75])
    r.close()


def test_unreadable_variable_is_nan_and_hdf5_is_refused(gdir):
    r = G.GSWP3Reader(gdir, variables=("pr", "rhs"))
    os.rename(os.path.join(gdir, "rhs_gswp3_1971_1980.nc4"), os.path.join(gdir, "rhs.bak"))
    try:
        m = r.read_month(1975, 5, np.array([0, 1]), np.array([0, 1]))
    finally:
        os.rename(os.path.join(gdir, "rhs.bak"), os.path.join(gdir, "rhs_gswp3_1971_1980.nc4"))
    assert np.isnan(m["rhs"]).all() and m["rhs"].shape == (30, 2)
    with pytest.raises(G.GSWP3FormatError, match="nccopy"):
        G.GSWP3Reader(gdir, variables=("tas",)).read_month(1985, 0, np.array([0]), np.array([0]))
    with pytest.raises(FileNotFoundError):
        G.GSWP3Reader(os.path.join(gdir, "nowhere"))


# ---- the tower's observed hourly shortwave -------------------------------------------------

def test_observed_shortwave_pins():
    n = 400
    rng = np.random.default_rng(2)
    cz = np.concatenate([np.zeros(50), rng.uniform(0.05, 1.0, n - 50)])
    sw = rng.uniform(-5.0, 1100.0, n)
    gen_d, gen_i = rng.random((n, 2)), rng.random((n, 2))
    use = rng.random(n) < 0.8
    d, i = OG.observed_shortwave(sw, use, cz, 172, gen_d, gen_i)
    np.testing.assert_allclose((d + i).sum(axis=1)[use], np.maximum(sw, 0.0)[use], rtol=1e-14)
    np.testing.assert_array_equal(d[~use], gen_d[~use])
    np.testing.assert_array_equal(i[~use], gen_i[~use])
    assert (d[:50][use[:50]] == 0.0).all()                          # sun down: all diffuse
    kt = sw / np.maximum(OG.solar_constant(172) * cz, 1.0)
    cloudy = use & (kt <= 0.22)
    assert (d[cloudy] == 0.0).all()
    r = 0.847 - 1.61 * cz + 1.04 * cz ** 2
    clear = use & (cz > 0) & (kt > (1.47 - r) / 1.66) & (kt < 1.0) & (r >= 0) & (r <= 1)
    assert clear.sum() > 10
    np.testing.assert_allclose(i[clear].sum(axis=1), (sw * r)[clear], rtol=1e-13)


@pytest.mark.parametrize("prec", [abi.PREC_F64, abi.PREC_MIXED])
def test_observed_shortwave_host_build_matches_the_oracle(prec):
    from ibis_amd import layout as L
    n = 600
    rng = np.random.default_rng(4)
    dt = np.float64 if prec == abi.PREC_F64 else np.float32
    a = rng.random((L.NPLANE, n)).astype(dt)
    cz = np.concatenate([np.zeros(60), rng.uniform(0.02, 1.0, n - 60)]).astype(dt)
    a[L.PLANE_OFFSET["atm.coszen"]] = cz
    sw = rng.uniform(-5.0, 1100.0, n).astype(dt)
    use = (rng.random(n) < 0.7).astype(np.uint8)
    od, oi = L.PLANE_OFFSET["atm.solad"], L.PLANE_OFFSET["atm.solai"]
    gen_d, gen_i = a[od:od + 2].T.astype(np.float64), a[oi:oi + 2].T.astype(np.float64)
    before = a.copy()
    rc = C.CDLL(HOSTCHECK).ibis_hostcheck_forcing_observed_sw(
        a.ctypes.data_as(C.c_void_p), C.c_int64(n), sw.ctypes.data_as(C.c_void_p),
        use.ctypes.data_as(C.c_void_p), C.c_int64(n), C.c_int(200), C.c_int(prec))
    assert rc == 0
    d, i = OG.observed_shortwave(sw.astype(np.float64), use, cz.astype(np.float64), 200,
                                 gen_d, gen_i)
    tol = dict(rtol=1e-13, atol=1e-300) if prec == abi.PREC_F64 else dict(rtol=2.5e-7, atol=0)
    np.testing.assert_allclose(a[od:od + 2].T, d.astype(dt), **tol)
    np.testing.assert_allclose(a[oi:oi + 2].T, i.astype(dt), **tol)
    other = np.ones(L.NPLANE, bool)
    other[[od, od + 1, oi, oi + 1]] = False
    np.testing.assert_array_equal(a[other], before[other])
