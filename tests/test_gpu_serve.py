"""Served step on the GPU: each request's result is the same alone and in a mix, a graph
replay equals eager launches, and a request with default parameters matches the plain
hourly driver (forcing, step, accumulators) on the same cells."""
import numpy as np
import pytest
import torch

from ibis_amd import driver as DR, serve as SV, synth
from ibis_amd.params import default_sim_params, get_constants, root_fractions
from ibis_amd.state import IBISState

pytestmark = pytest.mark.gpu


def _reqs():
    return [SV.Request(0, 100, 30, seed=11), SV.Request(1, 700, 45, seed=12),
            SV.Request(2, 64, 24, seed=13), SV.Request(3, 1, 37, seed=14),
            SV.Request(4, 3000, 26, seed=15)]


def test_each_request_is_the_same_alone_and_in_the_mix():
    alone, mixed = SV.run_alone_and_mixed(_reqs(), capacity=8192)
    for rid, a in alone.items():
        m = mixed[rid]
        assert a.shape == m.shape and np.isfinite(a).all()
        assert np.array_equal(a, m), (rid, np.abs(a - m).max())


def test_graph_replay_equals_eager_launches():
    g_alone, g_mixed = SV.run_alone_and_mixed(_reqs()[:3], capacity=4096, use_graph=True)
    e_alone, e_mixed = SV.run_alone_and_mixed(_reqs()[:3], capacity=4096, use_graph=False)
    for rid in g_mixed:
        assert np.array_equal(g_mixed[rid], e_mixed[rid]), rid
        assert np.array_equal(g_alone[rid], e_alone[rid]), rid


def test_default_request_matches_the_plain_driver():
    r = SV.Request(0, 500, 30, seed=21, perturb=False)
    _, mixed = SV.run_alone_and_mixed([r], capacity=1024)
    got = mixed[0]
    f, slot, lat = synth.make_cells(r.ncell, r.seed)
    st = IBISState.from_numpy(f, slot, synth.HSOI, root_fractions(), device="cuda")
    acc = DR.Accumulators(st)
    DR.run_hours(st, r.hours, DR.SyntheticForcing(lat, r.seed), get_constants(),
                 default_sim_params(), acc=acc)
    torch.cuda.synchronize()
    out = st.to_numpy(keys=SV.ServeEngine.RESULT_KEYS)
    ref = np.concatenate([np.asarray(out[k]).reshape(r.ncell, -1).T for k in SV.ServeEngine.RESULT_KEYS]
                         + [np.stack([acc.values()[k] for k in DR.ACC_NAMES])])
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    rel = np.where(np.abs(ref) < 1e-20, np.abs(got - ref), rel)
    print("\nserved vs plain driver: max rel diff %.3e" % rel.max())
    assert rel.max() < 1e-5


def test_requests_admitted_mid_run_into_freed_blocks_match_alone():
    """Continuous batching: with a 4,096-cell arena, the third request does not fit until
    the first retires; it is admitted into the freed blocks while the second is still
    running, and every result equals the request stepped alone."""
    reqs = [SV.Request(0, 3000, 20, seed=31), SV.Request(1, 700, 45, seed=32),
            SV.Request(2, 2000, 24, seed=33), SV.Request(3, 5, 12, seed=34)]
    arrive = {0: 0, 1: 3, 2: 5, 3: 30}
    eng = SV.ServeEngine(4096, max_requests=8)
    waiting, out, admitted_at, tick = [], {}, {}, 0
    pending = sorted(reqs, key=lambda r: arrive[r.rid])
    while pending or waiting or eng.active or eng.retiring:
        while pending and arrive[pending[0].rid] <= tick:
            waiting.append(pending.pop(0))
        while waiting and eng.admit(waiting[0]):
            admitted_at[waiting.pop(0).rid] = tick
        if eng.active:
            eng.tick()
        tick += 1
        eng.retire_finished()
        for r in eng.collect(lambda: 0.0, block=not eng.active):
            out[r.rid] = r.result["state_acc"]
    assert admitted_at[2] == 20 and admitted_at[2] > arrive[2], admitted_at
    assert reqs[2].start == 0          # the first request's blocks, reused
    alone, _ = SV.run_alone_and_mixed(
        [SV.Request(r.rid, r.ncell, r.hours, r.seed) for r in reqs], capacity=8192)
    for rid, a in alone.items():
        assert np.array_equal(a, out[rid]), rid
