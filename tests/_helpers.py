"""Shared builders for the step tests: synthetic cells, parameters, oracle runs."""
import numpy as np

from ibis_amd import layout as L, synth
from ibis_amd.params import default_sim_params, get_constants
from oracle import compare, params as OP, run as orun

OUT_KEYS = [f.key for f in L.FIELDS if f.role != "in"]
#: relative perturbation of the delta band and tolerance per precision: the FP64 kernel is
#: held to 1e-6 (the oracle and the kernel differ only by libm rounding), the mixed build to
#: the 1e-4 per-field bar, against FP32-rounded inputs
PREC = {"f64": (1e-12, 1e-6), "mixed": (1e-7, 1e-4), "f32": (1e-7, 1e-4)}


def workload(n, seed=11, hour=13, prec="f64", frozen=0.0, saturated=0.0):
    f, slot, lat = synth.make_cells(n, seed, frozen, saturated)
    f.update(synth.forcing(n, seed, hour, lat, f["atm.psurf"]))
    if prec != "f64":
        f = {k: np.asarray(v, np.float32).astype(np.float64) for k, v in f.items()}
    return f, slot


def oracle_with_band(f, slot, prec, bedrock=True, consts=None, params=None):
    """Oracle outputs and delta band for a workload.

    Returns ``(band, refs, keys, tol)``: ``refs[0]`` is the plain oracle run, the rest are
    its snow-skin branch variants (empty when no cell's weight is fractional); a cell
    passes if it matches any of them."""
    consts = consts or get_constants()
    params = params or default_sim_params(bedrock=bedrock)
    rel, tol = PREC[prec]

    def run(fl, sides=None):
        return orun.run_step(fl, slot, consts, params, hsoi=synth.HSOI,
                             froot=OP.root_fractions(), bedrock=bedrock, knife_sides=sides)
    keys = [k for k in OUT_KEYS if bedrock or k != "soil.tsoi_deep"]
    band, ref = compare.delta_band(run, f, keys, rel, tol)
    variants = compare.branch_variants(lambda sides: run(f, sides), ref)
    return band, [ref] + variants, keys, tol


def failing_cells(got, refs, keys, tol):
    return compare.failing(got, refs, keys, tol)


def cell_errors(got, refs, keys):
    """Per cell, the largest field error against the closest of ``refs``."""
    best = None
    for ref in refs:
        e = np.max([compare.rel_errors(got[k], ref[k]) for k in keys], axis=0)
        best = e if best is None else np.minimum(best, e)
    return best


def assert_parity(got, refs, band, keys, prec, label=""):
    """FP64: every cell outside the delta band within 1e-6.  Mixed / FP32 (the 1e-4 bar):
    outside the band at least 99.9 % of cells within 1e-4 per field and every cell within
    1e-3.  Either way at most 0.5 % of all cells, band included, over the tolerance.
    Knife-edge cells are checked against the branch variants, not excluded.  Prints the
    counts."""
    rel, tol = PREC[prec]
    bad = failing_cells(got, refs, keys, tol)
    out = bad & ~band
    n = len(bad)
    knife = np.asarray(refs[0]["_knife_edge"]).any(axis=1)
    err = cell_errors(got, refs, keys)
    worst = float(err[~band].max(initial=0.0))
    print(f"\n{label} {prec}: cells {n}, delta band {int(band.sum())}, snow-skin knife edge "
          f"{int(knife.sum())} (checked against {len(refs) - 1} branch variants), over {tol:g}: "
          f"{int(bad.sum())} ({int(out.sum())} outside band), worst outside band {worst:.2e}")
    if prec == "f64":
        assert not out.any(), np.where(out)[0][:20]
    else:
        assert out.sum() <= max(1, int(0.001 * n)), np.where(out)[0][:20]
        assert not (failing_cells(got, refs, keys, 1e-3) & ~band).any()
    assert bad.mean() < 0.005, bad.mean()
    return bad
