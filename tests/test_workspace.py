"""The step's workspace needs no initialisation (include/ibis_step.h): the host build of the
kernel source, run with every workspace byte NaN (IBIS_HOSTCHECK_WS_FILL=nan), gives bit
for bit the results of a zero-filled workspace -- for the plain step and for the dual-number
tangent step, whose tangents would pick up any unwritten value a zero weight hides from the
value."""
import os

import numpy as np
import pytest

from ibis_amd import adjoint as AD, synth
from _helpers import host_step
from test_adjoint import HostRun


def _with_fill(fill, fn):
    old = os.environ.get("IBIS_HOSTCHECK_WS_FILL")
    try:
        if fill:
            os.environ["IBIS_HOSTCHECK_WS_FILL"] = fill
        else:
            os.environ.pop("IBIS_HOSTCHECK_WS_FILL", None)
        return fn()
    finally:
        if old is None:
            os.environ.pop("IBIS_HOSTCHECK_WS_FILL", None)
        else:
            os.environ["IBIS_HOSTCHECK_WS_FILL"] = old


@pytest.mark.parametrize("prec", ["mixed", "f64"])
@pytest.mark.parametrize("hour", [3, 13])
def test_step_never_reads_unwritten_workspace(prec, hour):
    f, slot, lat = synth.make_cells(400, 17, 0.1, 0.1)
    f.update(synth.forcing(400, 17, hour, lat, np.asarray(f["atm.psurf"])))
    zero = _with_fill(None, lambda: host_step(f, slot, prec))
    nan = _with_fill("nan", lambda: host_step(f, slot, prec))
    for k in zero:
        np.testing.assert_array_equal(nan[k], zero[k], err_msg=k)


def test_tangent_step_never_reads_unwritten_workspace():
    hr = HostRun(8, seed=3)
    _, _, target = hr.run(AD.draw_tune(8, seed=91), 6, prec="mixed", record=True)
    tune = AD.draw_tune(8, seed=17)
    run = lambda: hr.run(tune, 6, prec="f64", tangent=True, ndir=AD.NTUNE, dir0=0,
                         target=target)
    a0, l0, _ = _with_fill(None, run)
    a1, l1, _ = _with_fill("nan", run)
    np.testing.assert_array_equal(a1, a0)
    np.testing.assert_array_equal(l1, l0)
