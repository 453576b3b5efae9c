"""The hourly step on the host CPU (``include/ibis_cpu.h``): the kernels' per-cell source built
with g++ and run over host threads, for the 1,024-cell benchmark case.

``step_cpu(fields, slot, params, ...)`` steps a dict of reference-shaped NumPy arrays one hour
and returns the new fields; ``python -m ibis_amd.cpu`` times 1,024 seeded cells x 24 h (FP32)
and prints one JSON line.  Replaces timestep_body (python-port/ibis/timestepping/loops.py:
234-1284) on the reference's CPU path.  It is an explicit entry point: the GPU path never
falls back to it.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib
from . import abi
from . import synth
from .params import SnowConstants, default_sim_params, get_constants, root_fractions

PRECISIONS = {"mixed": abi.PREC_MIXED, "f32": abi.PREC_F32, "f64": abi.PREC_F64}


class HostStep:
    """Packed host arenas of ``n`` cells stepped in place by ``ibis_step_cpu``."""

    def __init__(self, fields: dict, slot, params=None, consts=None, precision="f32",
                 hsoi=None, froot=None, co2conc=380.0e-6, snow=None, nthreads=0):
        self.n = len(np.asarray(fields["soil.tg"]))
        self.code = PRECISIONS[precision]
        dt = np.float64 if self.code == abi.PREC_F64 else np.float32
        self.params = params or default_sim_params()
        hsoi = synth.HSOI if hsoi is None else hsoi
        froot = root_fractions() if froot is None else froot
        self.cp = abi.params_from(consts or get_constants(), self.params, snow or SnowConstants(),
                                  hsoi, froot, co2conc)
        self.arena = np.ascontiguousarray(abi.pack(fields, self.n, dtype=dt))
        self.slot = np.ascontiguousarray(abi.pack_slots(slot, self.n))
        self.table = synth.type_table(self.params.stomata_params.vmax_pft,
                                      self.params.vmax_type_overrides)
        self.flags = self.code | (abi.FLAG_BEDROCK if self.params.bedrock is not None else 0)
        self.nthreads = int(nthreads)
        self.nytimes = 0.0

    def set(self, fields: dict):
        """Overwrite some fields (e.g. the hour's forcing)."""
        part = abi.pack(fields, self.n, dtype=self.arena.dtype)
        from . import layout as L
        for k in fields:
            o, w = L.PLANE_OFFSET[k], L.FIELDS[L.FIELD_INDEX[k]].width
            self.arena[o:o + w] = part[o:o + w]

    def step(self, dtime: float = 3600.0):
        lib = _lib.load()
        a = abi.IbisArena(self.arena.ctypes.data, self.n)
        rc = lib.ibis_step_cpu(C.byref(a), C.byref(a), C.c_void_p(self.slot.ctypes.data),
                               C.c_int64(self.n), C.c_void_p(self.table.ctypes.data),
                               C.byref(self.cp), C.c_double(dtime), C.c_double(self.nytimes),
                               C.c_int64(self.n), C.c_int(self.flags), C.c_int(self.nthreads))
        _lib.check(rc, "ibis_step_cpu")
        self.nytimes += 1.0

    def fields(self) -> dict:
        return abi.unpack(self.arena, self.n)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--hours", type=int, default=24)
    ap.add_argument("--precision", default="f32", choices=list(PRECISIONS))
    ap.add_argument("--threads", type=int, default=0, help="0: every core of this process")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args(argv)
    n = args.cells
    f, slot, lat = synth.make_cells(n, args.seed)
    threads = args.threads or len(os.sched_getaffinity(0))
    hs = HostStep(f, slot, precision=args.precision, nthreads=threads)
    forc = [synth.forcing(n, args.seed, h, lat, f["atm.psurf"]) for h in range(args.hours + 1)]
    hs.set(forc[0])
    hs.step()                                      # warm-up hour
    t = 0.0
    for h in range(1, args.hours + 1):
        hs.set(forc[h])
        t0 = time.perf_counter()
        hs.step()
        t += time.perf_counter() - t0
    out = hs.fields()
    print(json.dumps({
        "metric": "cell-hours/s", "value": n * args.hours / t, "unit": "cell-hours/s",
        "path": "host CPU (ibis_step_cpu)", "cells": n, "hours": args.hours,
        "precision": args.precision, "threads": threads,
        "finite_fraction": float(np.isfinite(out["canopy.tu"]).mean())}), flush=True)


if __name__ == "__main__":
    main()
