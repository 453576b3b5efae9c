"""Parameter-estimation step on the CPU: the host build of the same per-thread bodies the
GPU kernels run (libibis_hostcheck.so), in values and in dual numbers.

The forward-mode derivative is pinned by central differences of the FP64 value run (the
mathematics fixes it: for a smooth loss, (L(p + h) - L(p - h)) / 2h -> dL/dp with O(h^2)
error), and the value path by the plain step on the same cells and parameters."""
import ctypes as C

import numpy as np
import pytest

from ibis_amd import abi, adjoint as AD, layout as L, synth
from _helpers import HOSTCHECK

FLAGS_F64 = abi.PREC_F64 | abi.FLAG_BEDROCK
FLAGS_MIXED = abi.PREC_MIXED | abi.FLAG_BEDROCK


def _lib():
    lib = C.CDLL(HOSTCHECK)
    lib.ibis_hostcheck_tuned_hour.restype = C.c_int
    return lib


class HostRun:
    """Hourly value / tangent runs of the host build over numpy arrays."""

    def __init__(self, nsite, seed=3, site0=0):
        self.sites = AD.Sites(nsite, seed, site0=site0)
        self.base, _ = AD.base_params()
        self.rows = AD.row_pft()

    def run(self, tune, hours, prec="f64", tangent=False, ndir=1, dir0=-1, target=None,
            record=False):
        n = self.sites.nsite
        npair = tune.shape[1]
        nt = ndir * npair
        dt = np.float64 if prec == "f64" else np.float32
        k = 2 if tangent else 1
        a = np.zeros((L.NPLANE, nt, k), dtype=dt)
        a[:, :, 0] = np.tile(self.sites.state64.astype(dt), (1, nt // n))
        tune = np.ascontiguousarray(tune, dtype=np.float64)
        loss = np.zeros((2, nt))
        flux = np.zeros((NF := AD.NFLUX, nt), dtype=np.float32) if record else None
        rec = []
        run = AD.IbisTuned()
        run.nsite, run.npair, run.ndir, run.dir0 = n, npair, ndir, dir0
        run.jday0, run.seed, run.dtime = self.sites.jday0, self.sites.seed, 3600.0
        run.site0 = self.sites.site0
        run.flux_scale[:] = list(AD.FLUX_SCALE)
        flags = (FLAGS_F64 if prec == "f64" else FLAGS_MIXED) | (AD.FLAG_TANGENT if tangent else 0)
        ptr = lambda x: x.ctypes.data_as(C.c_void_p) if x is not None else None
        lib = _lib()
        for h in range(hours):
            run.hour, run.nytimes = h, float(h)
            tg = None if target is None else np.ascontiguousarray(target[h], dtype=np.float32)
            rc = lib.ibis_hostcheck_tuned_hour(ptr(a), C.c_int64(nt), ptr(self.sites.slot),
                                               ptr(self.sites.lat), ptr(tune), ptr(self.rows),
                                               C.byref(self.base), C.byref(run), ptr(tg),
                                               ptr(flux), ptr(loss), C.c_int(flags))
            assert rc == 0
            if record:
                rec.append(flux.copy())
        return a, loss, (np.stack(rec) if record else None)


@pytest.fixture(scope="module")
def case():
    hr = HostRun(4, seed=3)
    hours = 14                                     # night into the morning at most sites
    tgt_tune = AD.draw_tune(4, seed=91)
    _, _, target = hr.run(tgt_tune, hours, prec="mixed", record=True)
    tune = AD.draw_tune(4, seed=17)
    return hr, hours, tune, target


def test_tunable_defaults_follow_the_parameter_set():
    d = AD.default_tune()
    p, sim = AD.base_params()
    assert d.shape == (AD.NTUNE,)
    assert d[AD.TUNE_NAMES.index("alpha3")] == p.alpha3
    assert d[AD.TUNE_NAMES.index("coefm[3]")] == p.coefm[3]
    assert d[AD.TUNE_NAMES.index("bed_csoi")] == p.bed_csoi
    np.testing.assert_array_equal(d[AD.TUNE_VMAX:], sim.stomata_params.vmax_pft)
    t = AD.draw_tune(1000, seed=5)
    r = t / d[:, None]
    assert r.min() >= 0.8 and r.max() <= 1.2 and r.min() < 0.81 and r.max() > 1.19


def test_default_tuned_run_is_the_plain_step(case):
    """Value mode with every pair at the defaults performs the plain step's arithmetic."""
    hr, _, _, _ = case
    from _helpers import host_step
    n = hr.sites.nsite
    tune = np.repeat(AD.default_tune()[:, None], n, axis=1)
    a, _, _ = hr.run(tune, 1, prec="f64")
    f, slot, lat = synth.make_cells(n, hr.sites.seed)
    f.update(synth.forcing(n, hr.sites.seed, 0, lat, np.asarray(f["atm.psurf"])))
    ref = host_step(f, slot, "f64")
    got = abi.unpack(a[:, :, 0], n)
    for key in ("soil.tsoi", "soil.wsoi", "canopy.tu", "canopy.ag_upper", "hydrology.gtrans",
                "diag.fsena", "snow.fi"):
        np.testing.assert_allclose(got[key], ref[key], rtol=1e-13, atol=1e-300, err_msg=key)


def test_value_part_of_the_tangent_run_is_the_value_run(case):
    hr, hours, tune, target = case
    for prec in ("f64", "mixed"):
        a0, l0, _ = hr.run(tune, hours, prec=prec, target=target)
        a1, l1, _ = hr.run(tune, hours, prec=prec, tangent=True, ndir=3, dir0=5, target=target)
        n = tune.shape[1]
        for k in range(3):
            np.testing.assert_array_equal(a1[:, k * n:(k + 1) * n, 0], a0[:, :, 0])
            np.testing.assert_array_equal(l1[0, k * n:(k + 1) * n], l0[0])


def test_tangent_matches_central_differences_f64(case):
    """Every dL/dtheta of the FP64 tangent sweep against FP64 central differences at relative
    step 1e-4: the median component agrees to 1e-6, 95 % to 1e-4."""
    hr, hours, tune, target = case
    n = tune.shape[1]
    _, lt, _ = hr.run(tune, hours, prec="f64", tangent=True, ndir=AD.NTUNE, dir0=0,
                      target=target)
    g_ad = lt[1].reshape(AD.NTUNE, n) / tune
    step = 1e-4
    var = np.tile(tune, (1, 2 * AD.NTUNE))
    h = step * np.abs(tune)
    for k in range(AD.NTUNE):
        var[k, 2 * k * n:(2 * k + 1) * n] += h[k]
        var[k, (2 * k + 1) * n:(2 * k + 2) * n] -= h[k]
    _, lv, _ = hr.run(var, hours, prec="f64", target=target)
    lv = lv[0].reshape(AD.NTUNE, 2, n)
    g_fd = (lv[:, 0] - lv[:, 1]) / (2 * h)
    assert np.isfinite(g_ad).all()
    scale = np.abs(g_fd * tune).max(axis=0) + 1e-300          # per pair, in loss units
    err = np.abs(g_ad - g_fd) * np.abs(tune) / scale
    assert np.median(err) < 1e-6, np.median(err)
    assert np.quantile(err, 0.95) < 1e-4, np.quantile(err, 0.95)
    # the derivative is not trivially zero: the stomatal slope and Vmax move the fluxes
    assert np.abs(g_ad[AD.TUNE_NAMES.index("coefm[0]")]).max() > 0
    assert np.abs(g_ad[AD.TUNE_VMAX:]).max() > 0


def test_tuned_hour_rejects_bad_arguments():
    from ibis_amd import _lib as LIB
    lib = AD._lib_tuned()
    p, _ = AD.base_params()
    run = AD.IbisTuned()
    run.nsite, run.npair, run.ndir, run.dir0, run.dtime = 4, 4, 1, -1, 3600.0
    run.flux_scale[:] = list(AD.FLUX_SCALE)
    fake = C.c_void_p(256)
    ar = abi.IbisArena(C.c_void_p(256), 256)
    nb = lib.ibis_tuned_workspace_bytes(C.c_int64(256), C.c_int(abi.PREC_MIXED))
    assert nb > 0
    call = lambda flags, r=run, wsb=nb: lib.ibis_tuned_hour(
        C.byref(ar), fake, C.c_int64(wsb), fake, fake, fake, fake, C.byref(p), C.byref(r),
        None, None, None, C.c_int(flags), None)
    assert call(abi.PREC_F32) == -2                           # precision not offered
    assert call(abi.PREC_MIXED | 0x1000) == -2                # unknown flag bit
    assert call(abi.PREC_MIXED, wsb=nb - 1) == -1             # small workspace
    bad = AD.IbisTuned.from_buffer_copy(run)
    bad.npair = 6                                             # not a multiple of nsite
    assert call(abi.PREC_MIXED, r=bad) == -1
    bad = AD.IbisTuned.from_buffer_copy(run)
    bad.ndir, bad.dir0 = 3, 40                                # 40 + 3 > IBIS_NTUNE
    assert call(abi.PREC_MIXED | AD.FLAG_TANGENT, r=bad) == -1
    bad = AD.IbisTuned.from_buffer_copy(run)
    bad.flux_scale[1] = 0.0
    assert call(abi.PREC_MIXED, r=bad) == -1
    assert b"flux scales" in lib.ibis_last_error()
    assert LIB.load() is lib
