"""Benchmark of the hourly land-surface step: cell-hours/s on one MI355X.

Workload: 95,333 land cells (0.5-degree global land count) stepped hour by hour, one
kernel launch per hour, mixed precision (FP32 state and bulk arithmetic, FP64 canopy and
soil solves), synthetic seeded cells and forcing with the benchmark distributions
(``ibis_amd.synth``).  A "step" is one hour over all cells; ``--steps`` hours are timed
after ``--warmup`` untimed hours.  The forcing of every hour is generated before the timed
region and kept in HBM; each timed hour copies its 12 forcing planes into the state (as the
reference driver writes ``state.atm`` before each step, ``loops_gswp3.py:142-152``) and
launches the step.  By default the timed hours are captured once into a HIP graph and the
graph is replayed (``--launch eager`` enqueues each hour from the host instead).

Reported beside the throughput:

* ``roofline``: the step's achieved algorithmic bandwidth -- 1,448 B per cell-hour (179 fp32
  words read, 180 written, 12 int8 slot ids) x cells / mean duration of one ``ibis_step``
  call (its surface, canopy and ground kernels) measured with HIP events on the launch
  stream (with a graph, around the replay, so the hour's forcing copy is included) --
  against MI355X's 8 TB/s.  ``traffic`` is the HBM bytes one call moves, from the
  calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE summary committed in ``profiles/`` for this
  workload (null when no summary matches the run);
* ``cpu_baseline``: the oracle (NumPy FP64 restatement of the reference step, one core) on a
  bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
PROFILE = os.path.join(ROOT, "profiles", "step_{precision}_{cells}.md")


def profiled_traffic(precision, cells):
    """HBM bytes per step from the committed counter summary of this workload, or None."""
    path = PROFILE.format(precision=precision, cells=cells)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("traffic_bytes_per_step:"):
                return int(line.split(":", 1)[1].strip())
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=95333)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "f32", "f64"])
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-hours", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--launch", default="graph", choices=["eager", "graph"],
                    help="graph: the timed hours (forcing copy and step of each) captured once "
                         "into a HIP graph and replayed; eager: each hour enqueued from the "
                         "host, with HIP events around each step")
    ap.add_argument("--stomata", default="f77", choices=["f77", "release"],
                    help="f77: every optional StomataParams switch off (the default parameter "
                         "file); release: the switches of the release parameter file, a "
                         "per-cell stressfac and a seasonal acclimation state (ibis_step_ext)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the hours, copy a buffer of this many MiB once (a kernel of "
                         "known traffic that calibrates the HBM byte counters of a profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box of this build has a single MI355X; cells "
              "shard across GPUs with no exchange, so N GPUs run N independent shards.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from ibis_amd import layout as L, synth
    from ibis_amd.params import default_sim_params, get_constants, root_fractions
    from ibis_amd.state import IBISState
    from ibis_amd import step as S

    n = args.cells
    hours = args.warmup + args.steps
    dev = torch.device("cuda:0")
    dtype = torch.float64 if args.precision == "f64" else torch.float32
    pitch = (n + 255) // 256 * 256
    fields, slot, lat = synth.make_cells(n, args.seed)
    consts, params = get_constants(), default_sim_params(bedrock=True)
    if args.stomata == "release":
        release_inputs(params, fields, n, args.seed)
    st = IBISState.from_numpy(fields, slot, synth.HSOI, root_fractions(), device=dev,
                              dtype=dtype, pitch=pitch)
    cparams = S.ibis_params(st, consts, params)
    cext = S.ext_inputs(st, params) if S.uses_ext(params) else None

    atm0 = L.PLANE_OFFSET["atm.ta"]
    natm = sum(f.width for f in L.fields_of("atm"))
    forc = torch.zeros((hours, natm, pitch), dtype=dtype, device=dev)
    for h in range(hours):
        fh = synth.forcing(n, args.seed, h, lat, fields["atm.psurf"])
        host = np.zeros((natm, pitch))
        for f in L.fields_of("atm"):
            off = L.PLANE_OFFSET[f.key] - atm0
            v = np.asarray(fh[f.key])
            if f.width == 1:
                host[off, :n] = v
            else:
                host[off:off + f.width, :n] = v.T
        forc[h] = torch.from_numpy(host).to(dev, dtype)
    atm_view = st.arena[atm0:atm0 + natm]
    stream = torch.cuda.current_stream(dev)

    def hour(h, ev=None, s=stream):
        atm_view.copy_(forc[h])
        if ev is not None:
            ev[0].record(s)
        S.timestep_(st, consts, params, 3600.0, precision=args.precision, stream=s,
                    cparams=cparams, cext=cext)
        if ev is not None:
            ev[1].record(s)

    if args.copy_probe_mib > 0:
        words = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(words, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)          # elementwise kernel: reads and writes the size once
        torch.cuda.synchronize()
        del src, dst
    if args.launch == "graph":
        # the timed hours as one HIP graph on a capture stream of default priority; the
        # warm-up hours run on that stream first, which also allocates its workspace
        cs = torch.cuda.Stream(dev)
        with torch.cuda.stream(cs):
            for h in range(args.warmup):
                hour(h, s=cs)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cs):
            for k in range(args.steps):
                hour(args.warmup + k, s=cs)
        torch.cuda.synchronize()
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        t0 = time.perf_counter()
        with torch.cuda.stream(cs):
            ev[0].record(cs)
            graph.replay()
            ev[1].record(cs)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # per hour: the forcing copy and the three kernels, from events around the replay
        kms = [ev[0].elapsed_time(ev[1]) / args.steps]
    else:
        for h in range(args.warmup):
            hour(h)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        t0 = time.perf_counter()
        for k in range(args.steps):
            hour(args.warmup + k, evs[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = [a.elapsed_time(b) for a, b in evs]
    kmean = float(np.mean(kms)) * 1e-3
    out = st.to_numpy(keys=("soil.tsoi", "canopy.tu"))
    finite = float(np.mean(np.isfinite(out["soil.tsoi"]).all(1) & np.isfinite(out["canopy.tu"])))

    value = n * args.steps / wall
    per_cell = L.ALGORITHMIC_BYTES_PER_CELL_HOUR
    achieved = per_cell * n / kmean
    res = {
        "metric": "cell-hours/s", "value": value, "unit": "cell-hours/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "f32+f64", "f32": "f32", "f64": "f64"}[args.precision],
        "data": "synthetic (seeded cells and hourly forcing, benchmark distributions)",
        "config": {"workload": f"plain hourly step, {n:,} cells x {args.steps} h, one launch per hour",
                   "cells": n, "hours": args.steps, "precision": args.precision,
                   "soil": "4 m soil + 8 m bedrock", "plant_type_slots": 12,
                   "stomata": args.stomata, "launch": args.launch},
        "kernel_ms_mean": kmean * 1e3, "kernel_ms_min": float(np.min(kms)),
        "finite_fraction": finite,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": (profiled_traffic(args.precision, n)
                                 if args.stomata == "f77" else None),
                     "bytes_per_cell_hour": per_cell,
                     "scope": ("one ibis_step call" if args.stomata == "f77" else
                               "one ibis_step_ext call") + ": surface + canopy + ground kernels"
                              + ("" if args.launch == "eager" else
                                 ", timed per hour over the graph replay with the forcing copy")},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(n, args.cpu_hours, args.seed, args.stomata)
    print(json.dumps(res), flush=True)


def release_inputs(params, fields, n, seed):
    """The release switches on ``params``, with a per-cell stressfac and a seasonal
    acclimation state drawn for the cells (added to ``fields``)."""
    from ibis_amd.params import release_switches
    release_switches(params)
    rng = np.random.default_rng(seed + 1)
    params.stomata_params.stressfac = rng.uniform(-9.0, -3.0, n)
    fields["seasonal.acclim_x"] = rng.uniform(-15.0, 25.0, n)


def cpu_baseline(n, hours, seed, stomata="f77"):
    """The oracle (NumPy FP64, one core) on the first ``hours`` hours of the same cells."""
    from ibis_amd import synth
    from ibis_amd.params import default_sim_params, get_constants
    from oracle import params as OP, run as orun
    fields, slot, lat = synth.make_cells(n, seed)
    consts, params = get_constants(), default_sim_params(bedrock=True)
    if stomata == "release":
        release_inputs(params, fields, n, seed)
    st = dict(fields)
    t = 0.0
    for h in range(hours):
        st.update(synth.forcing(n, seed, h, lat, fields["atm.psurf"]))
        t0 = time.perf_counter()
        out = orun.run_step(st, slot, consts, params, hsoi=synth.HSOI, froot=OP.root_fractions(),
                            nytimes=float(h))
        t += time.perf_counter() - t0
        for k, v in out.items():
            if not k.startswith("_"):
                st[k] = v
    return {"value": n * hours / t, "unit": "cell-hours/s", "cores": 1, "kind": "port",
            "sample": f"{n:,} cells x {hours} h, NumPy FP64 oracle, one thread",
            "cpu_model": _cpu_model()}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()
