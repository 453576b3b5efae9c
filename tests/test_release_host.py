"""The optional StomataParams switches in the kernel source, built for the host, against the
oracle (no GPU needed): ``ibis_step_ext`` (include/ibis_step.h) with the release parameter
file's switches (sunlit/shaded canopies, composition-weighted Vmax and slopes, acclimation
per plant type, per-slot tables, a per-cell stressfac) and the other switches (capacity
scaling, the leaf-only denominator, the light-rate factor, the frost ramp)."""
import copy

import numpy as np
import pytest

from _helpers import (assert_parity, capscale_params, crop_params, host_step, host_step_ext,
                      oracle_with_band, release_params, rounded, seasonal, workload)
from ibis_amd import abi
from ibis_amd.params import default_sim_params


def test_no_switch_is_the_plain_step_bit_for_bit():
    for prec in ("f64", "mixed"):
        f, slot = workload(500, seed=51, hour=13, prec=prec)
        sim = default_sim_params()
        a = host_step(f, slot, prec, params=sim)
        b = host_step_ext(f, slot, prec, sim)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{prec} {k}")


def switch_config(config, n, prec):
    sim = {"release": lambda: release_params(n), "capscale": capscale_params,
           "crop": lambda: crop_params(n)}[config]()
    if prec != "f64":                            # the per-cell planes as the FP32 build sees them
        if np.ndim(sim.stomata_params.stressfac):
            sim.stomata_params.stressfac = np.asarray(sim.stomata_params.stressfac,
                                                      np.float32).astype(np.float64)
        if sim.cropfrac is not None:
            sim.cropfrac = np.asarray(sim.cropfrac, np.float32).astype(np.float64)
    return sim


@pytest.mark.parametrize("prec", ["f64", "mixed"])
@pytest.mark.parametrize("config", ["release", "capscale", "crop"])
def test_switches_match_the_oracle(prec, config):
    n = 1500
    f, slot = workload(n, seed=61, hour=13, prec=prec)
    sim = switch_config(config, n, prec)
    aux = rounded(seasonal(n), prec)
    fo = dict(f, **aux)
    band, refs, keys, tol = oracle_with_band(fo, slot, prec, params=sim)
    got = host_step_ext(f, slot, prec, sim, aux)
    assert_parity(got, refs, band, keys, prec, f"host {config}")


def test_switches_change_the_canopy():
    n = 800
    f, slot = workload(n, seed=71, hour=13, prec="f64")
    base = host_step(f, slot, "f64")
    aux = seasonal(n)
    for config in ("release", "capscale", "crop"):
        sim = switch_config(config, n, "f64")
        got = host_step_ext(f, slot, "f64", sim, aux)
        day = base["canopy.ag_upper"][:, 0] > 0
        assert day.sum() > 50
        assert not np.allclose(got["canopy.ag_upper"][day], base["canopy.ag_upper"][day]), config


def test_acclimation_needs_the_seasonal_state():
    """acclim_on without a seasonal state keeps the F77 Vmax (timestepping/loops.py:294)."""
    n = 300
    f, slot = workload(n, seed=72, hour=13, prec="f64")
    sim = default_sim_params()
    sim.stomata_params.acclim_on = True
    a = host_step(f, slot, "f64", params=sim)
    b = host_step_ext(f, slot, "f64", sim)                 # no seasonal.acclim_x
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    c = host_step_ext(f, slot, "f64", sim, {"seasonal.acclim_x": np.full(n, -20.0)})
    conifer = f["vegetation.exist"][:, 3] > 0.5
    assert not np.allclose(c["canopy.ag_upper"][:, 1], a["canopy.ag_upper"][:, 1])
    assert (c["canopy.ag_upper"][:, 1][conifer] <= a["canopy.ag_upper"][:, 1][conifer] + 1e-12).all()


def test_ext_entry_point_rejects_bad_switch_blocks():
    import ctypes as C
    from ibis_amd import _lib
    lib = _lib.load()
    ar = abi.IbisArena(C.c_void_p(256), 256)
    p = abi.IbisParams()
    for k in range(3):
        p.bed_hsoi[k] = 1.0
    x = abi.IbisStomataExt()
    x.type_tables = 256
    call = lambda xx, flags=abi.PREC_MIXED: lib.ibis_step_ext(
        C.byref(ar), C.byref(ar), C.c_void_p(256), C.c_int64(1 << 30), None, C.c_int64(0),
        C.byref(p), C.byref(xx) if xx is not None else None, C.c_double(3600.0), C.c_double(0.0),
        C.c_int64(4), C.c_int(flags), None)
    assert call(None) == -1
    bad = abi.IbisStomataExt.from_buffer_copy(x)
    bad.switches = 0x1000
    assert call(bad) == -2 and b"switch" in lib.ibis_last_error()
    bad.switches = abi.SW_CROP
    assert call(bad) == -1 and b"cropland" in lib.ibis_last_error()
    bad = abi.IbisStomataExt.from_buffer_copy(x)
    bad.switches = abi.SW_ACCLIM
    assert call(bad) == -1 and b"acclimation state" in lib.ibis_last_error()
    bad.acclim_x = 256
    bad.switches = abi.SW_ACCLIM | abi.SW_FROST
    assert call(bad) == -1 and b"daily mean temperature" in lib.ibis_last_error()
    bad = abi.IbisStomataExt.from_buffer_copy(x)
    bad.acclim_mask = 1 << 12
    assert call(bad) == -1
    bad = abi.IbisStomataExt.from_buffer_copy(x)
    bad.type_tables = None
    assert call(bad) == -1
    assert call(x, flags=abi.PREC_MIXED | 0x1000) == -2
    ok = abi.IbisStomataExt.from_buffer_copy(x)
    assert lib.ibis_step_ext(C.byref(ar), C.byref(ar), C.c_void_p(256), C.c_int64(1 << 30), None,
                             C.c_int64(0), C.byref(p), C.byref(ok), C.c_double(3600.0),
                             C.c_double(0.0), C.c_int64(0), C.c_int(abi.PREC_MIXED), None) == 0


def test_switch_bits_follow_the_parameter_names():
    sim = release_params(4)
    x = abi.stomata_ext(sim.stomata_params)
    want = (abi.SW_SUNSHADE_U | abi.SW_SUNSHADE_L | abi.SW_VMAX_MIX | abi.SW_COEFM_MIX
            | abi.SW_ACCLIM | abi.SW_ACCLIM_PER_PFT)
    assert x.switches == want
    assert x.acclim_mask == (1 << 3) | (1 << 5) | (1 << 8)
    assert default_sim_params().stomata_params.switches_on() is False
    assert sim.stomata_params.switches_on() is True
    sp = copy.copy(default_sim_params().stomata_params)
    sp.stressfac = np.full(3, -5.0)
    assert sp.switches_on() is True                      # a per-cell stressfac takes the ext path
