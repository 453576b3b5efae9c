"""The step on a real IBIS state: the reference's saved 171-cell site state
(``verify_out_stage1/site_transient_1996.npz``, extracted by ``tests/golden/make_site_fixture.py``).

Its stored forcing is a night hour; a daytime variant swaps in seeded synthetic forcing on the
same state.  The kernel source (host build) is checked against the oracle, and the oracle and
kernel both close the soil-column water balance the reference tests
(``tests/test_water_balance.py:35-47``) on the snow-free cells, with the ``wadjust`` mass
correction (``hydrology.gadjust``) counted as a source.  The same cases run on the GPU in
``test_gpu_step.py``.
"""
import numpy as np
import pytest

from _helpers import assert_parity, host_step, oracle_with_band, site_case, water_residual


def _f32(f):
    return {k: np.asarray(v, np.float32).astype(np.float64) for k, v in f.items()}


@pytest.mark.parametrize("prec", ["f64", "mixed"])
@pytest.mark.parametrize("day", [None, 13])
def test_site_state_matches_oracle(prec, day):
    f, slot, glob = site_case(day)
    if prec != "f64":
        f = _f32(f)
    band, refs, keys, tol = oracle_with_band(f, slot, prec, glob=glob)
    got = host_step(f, slot, prec, glob=glob)
    assert_parity(got, refs, band, keys, prec, f"site {'night' if day is None else 'day'}")


@pytest.mark.parametrize("day", [None, 13])
def test_site_state_closes_the_snow_free_water_balance(day):
    f, slot, glob = site_case(day)
    band, refs, keys, tol = oracle_with_band(f, slot, "f64", glob=glob)
    got = host_step(f, slot, "f64", glob=glob)
    free = f["snow.fi"] == 0.0
    assert free.sum() > 50
    for out in (refs[0], got):
        resid, scale = water_residual(f, out, glob["hsoi"], 3600.0)
        resid = resid - out["hydrology.gadjust"] * 3600.0
        # the reference's bound is 0.02 x flux scale + 0.05 kg m-2; FP64 closes to rounding
        assert np.all(np.abs(resid[free]) <= 0.02 * scale[free] + 0.05)
        assert np.abs(resid[free]).max() < 1e-8
