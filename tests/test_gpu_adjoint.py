"""Parameter-estimation step on the GPU: value runs against the host build, the tangent
sweep against the host build's tangent and against FP64 central differences, and the
autograd wrapper."""
import numpy as np
import pytest
import torch

from ibis_amd import adjoint as AD
from test_adjoint import HostRun

pytestmark = pytest.mark.gpu
NS, HOURS = 64, 14


@pytest.fixture(scope="module")
def gpu_case():
    hr = HostRun(NS, seed=3)
    _, _, target = hr.run(AD.draw_tune(NS, seed=91), HOURS, prec="mixed", record=True)
    run = AD.TunedRunner(AD.Sites(NS, seed=3))
    tune = AD.draw_tune(NS, seed=17)
    return hr, run, tune, target


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_value_run_matches_host_build(gpu_case):
    hr, run, tune, target = gpu_case
    for prec, tol in (("f64", 1e-9), ("mixed", 2e-3)):
        _, lh, _ = hr.run(tune, HOURS, prec=prec, target=target)
        lg, _ = run.value(_dev(tune), _dev(target), prec=prec)
        lg = lg.cpu().numpy()
        rel = np.abs(lg - lh[0]) / np.maximum(np.abs(lh[0]), 1e-12)
        print(f"\n{prec} loss GPU vs host: max rel {rel.max():.3e}")
        assert rel.max() < tol, (prec, rel.max())


def test_tangent_matches_host_build_f64(gpu_case):
    hr, run, tune, target = gpu_case
    _, lt, _ = hr.run(tune, HOURS, prec="f64", tangent=True, ndir=AD.NTUNE, dir0=0,
                      target=target)
    gh = lt[1].reshape(AD.NTUNE, NS)
    _, gg = run.tangent(_dev(tune), _dev(target), prec="f64")
    e = AD.gradient_errors(gg.cpu().numpy(), gh, tune)
    print("\nf64 tangent GPU vs host:", e)
    assert e["pair_p99"] < 1e-7 and e["rel_p99"] < 1e-5


def test_mixed_tangent_matches_f64_central_differences(gpu_case):
    _, run, tune, target = gpu_case
    t, tg = _dev(tune), _dev(target)
    _, g = run.tangent(t, tg, prec="mixed")
    g_fd = run.central_differences(t, tg, step=1e-3, prec="f64")
    e = AD.gradient_errors(g.cpu().numpy(), g_fd.cpu().numpy(), tune)
    print("\nmixed tangent vs FP64 central differences:", e)
    assert e["pairs_within_1e-3"] >= 0.95 and e["within_1e-3"] >= 0.95
    assert e["pair_median"] < 1e-4


def test_value_part_of_the_tangent_run_follows_the_value_run(gpu_case):
    _, run, tune, target = gpu_case
    t, tg = _dev(tune), _dev(target)
    lv, _ = run.value(t, tg, prec="mixed")
    lt, _ = run.tangent(t, tg, prec="mixed", dirs=(0, 4))
    rel = (lt - lv).abs() / lv.abs().clamp_min(1e-12)
    print(f"\nloss of the tangent run vs value run: max rel {rel.max().item():.3e}")
    assert rel.max().item() < 1e-5


def test_autograd_backward_is_the_tangent_sweep(gpu_case):
    _, run, tune, target = gpu_case
    t = _dev(tune).requires_grad_(True)
    tg = _dev(target)
    w = torch.linspace(0.5, 1.5, NS, dtype=torch.float64, device="cuda")
    loss = AD.FluxMisfit.apply(t, run, tg, "mixed", 6)
    (loss * w).sum().backward()
    _, g = run.tangent(t.detach(), tg, prec="mixed")
    # checkpoint restarts move the linearisation point by at most FP32 rounding
    e = AD.gradient_errors((t.grad / w).cpu().numpy(), g.cpu().numpy(), tune)
    print("\nautograd (checkpointed) vs plain tangent sweep:", e)
    assert e["pair_p99"] < 1e-4
    lv, _ = run.value(t.detach(), tg, prec="mixed")
    assert torch.equal(loss.detach(), lv)
