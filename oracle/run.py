"""Convenience driver of the oracle step -- TEST INFRASTRUCTURE.

``run_step`` applies the per-cell plant-type gather the reference does on the host
(``python-port/ibis/io/pft_types.py:170-189``: a slot holding zone type t takes the Vmax of
t's parent PFT unless overridden; a negative id takes the position's own value) and calls
``oracle.step.timestep_body`` with reference-named parameter objects.
"""
from __future__ import annotations

import copy
from types import SimpleNamespace

import numpy as np

from . import params as OP
from .step import timestep_body

#: parent IBIS PFT (0-based) of the 40 zone types, python-port/ibis/io/pft_types.py:51-92
TYPE_PARENT = np.array([0, 1, 1, 2, 1, 2, 3, 4, 2, 3, 4, 4, 3, 2, 4, 3, 5, 6, 7, 3,
                        8, 9, 8, 9, 8, 9, 8, 10, 10, 10, 11, 11, 11, 11, 11,
                        10, 11, 11, 11, 10])


def gather_vmax(slot_type, vmax_base, overrides=None):
    """(N, 12) per-cell Vmax from slot type ids (pft_types.gather)."""
    base = np.asarray(vmax_base, dtype=np.float64)
    val = base[TYPE_PARENT].copy()
    for t, v in (overrides or {}).items():
        val[int(t)] = float(v)
    s = np.asarray(slot_type, dtype=np.int64)
    own = np.broadcast_to(base, s.shape)
    return np.where(s >= 0, val[np.clip(s, 0, 39)], own)


def run_step(fields, slot_type, consts=None, params=None, dtime=3600.0, hsoi=None, froot=None,
             co2conc=380.0e-6, nytimes=0.0, snow=None, bedrock=True, knife_sides=None):
    consts = consts or OP.physical_constants()
    params = params or OP.sim_params()
    p = copy.copy(params)
    sp = copy.copy(params.stomata_params)
    sp.vmax_pft = gather_vmax(slot_type, params.stomata_params.vmax_pft,
                              getattr(params, "vmax_type_overrides", None))
    p.stomata_params = sp
    if not bedrock:
        p.bedrock = None
    glob = SimpleNamespace(hsoi=OP.HSOI_LAYERS if hsoi is None else hsoi,
                           froot=OP.root_fractions() if froot is None else froot,
                           co2conc=co2conc, nytimes=nytimes,
                           snow=snow or OP.snow_constants(), knife_sides=knife_sides)
    f = dict(fields)
    if not bedrock:
        f.pop("soil.tsoi_deep", None)
    with np.errstate(all="ignore"):
        return timestep_body(f, consts, p, dtime, glob)
