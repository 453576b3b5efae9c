"""The host CPU path (``ibis_step_cpu``, include/ibis_cpu.h) against the oracle and against the
host build of the kernel source (no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

from _helpers import assert_parity, host_step, oracle_with_band, workload
from ibis_amd import _lib, abi
from ibis_amd.cpu import HostStep


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_cpu_path_matches_the_oracle(prec):
    """FP64 to 1e-6 and mixed to the 1e-4 bar; FP32 throughout (the precision of the
    1,024-cell CPU case) is measured, not held to the bar, as on the GPU."""
    n = 1024
    f, slot = workload(n, seed=91, hour=13, prec=prec)
    band, refs, keys, tol = oracle_with_band(f, slot, prec)
    hs = HostStep(f, slot, precision=prec, nthreads=3)
    hs.step()
    got = hs.fields()
    assert_parity(got, refs, band, keys, prec, "cpu path")


def test_cpu_path_is_the_host_build_and_independent_of_the_thread_count():
    n = 700
    f, slot = workload(n, seed=92, hour=11, prec="f64")
    ref = host_step(f, slot, "f64")
    for threads in (1, 4):
        hs = HostStep(f, slot, precision="f64", nthreads=threads)
        hs.step()
        got = hs.fields()
        for k in ref:
            np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=1e-300, err_msg=k)


def test_cpu_path_rejects_bad_arguments():
    lib = _lib.load()
    a = abi.IbisArena(C.c_void_p(256), 4)
    p = abi.IbisParams()
    call = lambda n, flags, table=256: lib.ibis_step_cpu(
        C.byref(a), C.byref(a), None, C.c_int64(0), C.c_void_p(table), C.byref(p),
        C.c_double(3600.0), C.c_double(0.0), C.c_int64(n), C.c_int(flags), C.c_int(1))
    assert call(5, abi.PREC_F32) == -1                       # pitch < ncell
    assert call(4, abi.PREC_F32 | 0x1000) == -2
    assert call(4, abi.PREC_F32, table=0) == -1
    assert call(4, abi.PREC_F32 | abi.FLAG_BEDROCK) == -4    # zero bedrock thicknesses
    assert call(0, abi.PREC_F32) == 0
