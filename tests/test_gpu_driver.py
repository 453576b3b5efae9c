"""The hourly driver on the GPU (include/ibis_driver.h): forcing kernels against the host
generator and the oracle, the accumulator kernel against the oracle, and the 30-day
mixed-precision run against a full-FP64 GPU run on the 95,333-cell workload with frozen and
saturated cells."""
import numpy as np
import pytest
import torch

from ibis_amd import driver as DR, synth
from ibis_amd.params import default_sim_params, get_constants
from ibis_amd.state import IBISState
from ibis_amd.step import timestep_
from oracle import driver as OD, params as OP
from _helpers import workload

from test_driver import _acc_case, _daily

pytestmark = pytest.mark.gpu


def _state(f, slot, dtype, pitch=None):
    return IBISState.from_numpy(f, slot, synth.HSOI, OP.root_fractions(), device="cuda",
                                dtype=dtype, pitch=pitch)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_synthetic_forcing_kernel_matches_the_generator(dtype):
    n, seed = 5000, 2024
    f, slot, lat = synth.make_cells(n, seed)
    st = _state(f, slot, dtype, pitch=n + 11)
    gen = DR.SyntheticForcing(lat, seed)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    for hour in (0, 13, 30, 700):
        gen.write(st, hour)
        torch.cuda.synchronize()
        got = st.to_numpy()
        ps = np.asarray(f["atm.psurf"], npdt).astype(np.float64)
        for k, v in synth.forcing(n, seed, hour, lat, ps).items():
            r = np.asarray(v, npdt).astype(np.float64)
            # FP64: last-bit differences of the device sin/cos, amplified by 1 - fd where the
            # diffuse fraction is near 1; FP32: one rounding of the stored value
            tol = 1e-11 if dtype == torch.float64 else 2.5e-7
            np.testing.assert_allclose(got[k], r, rtol=tol, atol=1e-300, err_msg=k)


def test_diurnal_kernel_matches_the_oracle():
    n = 4000
    dw = _daily(n)
    lat = np.random.default_rng(9).uniform(-56.0, 84.0, n)
    f, slot, _ = synth.make_cells(n, 1)
    st = _state(f, slot, torch.float64)
    gen = DR.DiurnalForcing(dw, lat, jday=172, seed=77, dtype=torch.float64)
    total = np.zeros(n)
    for step in range(24):
        gen.write(st, step)
        torch.cuda.synchronize()
        got = st.to_numpy()
        rn = synth.uniform(77, 27, np.arange(n, dtype=np.uint64), step)
        ref = OD.diurnal(step * 3600.0, 172.0, dw, lat, 3600.0, rn)
        for k, v in ref.items():
            np.testing.assert_allclose(got[k], v, rtol=1e-12, atol=1e-300, err_msg=k)
        total += (got["atm.raina"] + got["atm.snowa"]) * 3600.0
    np.testing.assert_allclose(total, dw["precip"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_accumulate_kernel_matches_the_oracle(dtype, split):
    n = 5000
    stf, vx, op = _acc_case(n)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    stf = {k: np.asarray(v, npdt).astype(np.float64) for k, v in stf.items()}
    vx = {k: np.asarray(v, npdt).astype(np.float64) for k, v in vx.items()}
    slot = np.full((n, 12), -1, np.int8)
    params = None
    if split:                                  # gpp_split_by_capacity over random slot types
        slot = np.random.default_rng(4).integers(-1, 40, (n, 12)).astype(np.int8)
        params = default_sim_params()
        params.stomata_params.gpp_split_by_capacity = True
        params.vmax_type_overrides = {3: 90e-6, 17: 5e-6}
        table = synth.type_table(params.stomata_params.vmax_pft, params.vmax_type_overrides)
        op.vmax = table[np.where(slot >= 0, slot, 40 + np.arange(12)[None, :])]
    st = _state(stf, slot, dtype)
    acc = DR.Accumulators(st, vegx=vx, params=params)
    for _ in range(5):
        acc.add(st)
    torch.cuda.synchronize()
    got = acc.values()
    ref = 5.0 * OD.acc_increments(stf, vx, op)
    tol = 1e-12 if dtype == torch.float64 else 3e-5
    for j, name in enumerate(DR.ACC_NAMES):
        scale = max(np.abs(ref[j]).max(), 1e-30)
        assert np.abs(got[name] - ref[j]).max() <= tol * scale, name


def test_thirty_days_mixed_precision_against_an_fp64_gpu_run():
    """95,333 cells x 720 h, 10 % frozen and 10 % saturated: the mixed build against the full
    FP64 build, both on the GPU, from the same FP32-representable initial state and forcing.
    Prints the per-field error of the final state and of the 30-day budget sums."""
    n, hours, seed = 95333, 720, 2024
    f, slot, lat = synth.make_cells(n, seed, frozen_frac=0.1, saturated_frac=0.1)
    f = {k: np.asarray(v, np.float32).astype(np.float64) for k, v in f.items()}
    c, p = get_constants(), default_sim_params()
    out, sums = {}, {}
    for name, dtype, prec in (("f64", torch.float64, "f64"), ("mixed", torch.float32, "mixed")):
        st = _state(f, slot, dtype, pitch=(n + 255) // 256 * 256)
        acc = DR.Accumulators(st)
        DR.run_hours(st, hours, DR.SyntheticForcing(lat, seed), c, p, precision=prec, acc=acc)
        torch.cuda.synchronize()
        out[name] = st.to_numpy()
        sums[name] = acc.values()
    ref, got = out["f64"], out["mixed"]
    keys = ("soil.tsoi", "soil.wsoi", "soil.wisoi", "soil.tg", "soil.tsoi_deep", "snow.fi",
            "snow.hsno", "canopy.tu", "canopy.tl", "canopy.t12", "canopy.q12", "soil.wpud",
            "vegetation.firefac")
    print("\nfield after 720 h           median     p99        max        finite")
    for k in keys:
        r = np.asarray(ref[k]).reshape(n, -1)
        g = np.asarray(got[k]).reshape(n, -1)
        e = np.abs(g - r) / np.maximum(np.abs(r), 1e-2 * np.sqrt(np.mean(r * r, axis=0)) + 1e-30)
        e = e.max(axis=1)
        print(f"{k:26s} {np.median(e):.2e}  {np.percentile(e, 99):.2e}  {e.max():.2e}  "
              f"{np.isfinite(g).all()}")
        assert np.isfinite(g).all(), k
    print("30-day sum                   median     p99        max")
    for k in ("aygpptot", "aynpptot", "adaet", "et_trans", "et_soilev", "aytrunof",
              "aytdrain", "adrain", "adsnow"):
        r, g = sums["f64"][k], sums["mixed"][k]
        e = np.abs(g - r) / np.maximum(np.abs(r), 1e-2 * np.sqrt(np.mean(r * r)) + 1e-30)
        print(f"{k:26s} {np.median(e):.2e}  {np.percentile(e, 99):.2e}  {e.max():.2e}")
        if k in ("adrain", "adsnow"):
            assert e.max() < 1e-6, k          # forcing only: FP32 storage of the same draws
        else:
            assert np.median(e) < 1e-3, k


def test_nonfinite_counter_counts_cells_with_a_bad_written_plane():
    n = 3000
    f, slot = workload(n, seed=7, hour=12, prec="mixed")
    st = IBISState.from_numpy(f, slot, synth.HSOI, OP.root_fractions(), device="cuda")
    new = timestep_(st, get_constants(), default_sim_params(), 3600.0)
    c = DR.count_nonfinite(new)
    torch.cuda.synchronize()
    assert int(c.item()) == 0
    new.canopy.tu[[5, 17]] = float("nan")                 # written planes: counted per cell
    new.soil.tsoi[17, 3] = float("inf")
    new.hydrology.gtrans[2999] = float("-inf")
    new.soil.sand[40, 0] = float("nan")                   # a read-only plane: not counted
    c = DR.count_nonfinite(new)
    DR.count_nonfinite(new, counter=c)                    # adds to the given counter
    torch.cuda.synchronize()
    assert int(c.item()) == 2 * 3
