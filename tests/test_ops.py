"""The step's registered PyTorch operators: schema and fake (meta) behaviour on the CPU."""
import ctypes as C

import pytest
import torch

from ibis_amd import abi, layout as L, ops  # noqa: F401
from ibis_amd.params import default_sim_params


def test_operators_are_registered_with_the_step_signature():
    s = str(torch.ops.ibis.step_.default._schema)
    assert s.startswith("ibis::step_(Tensor(a0!) arena, Tensor(a1!) workspace, Tensor? slot_type, "
                        "Tensor type_table, Tensor params, float dtime, float nytimes")
    assert s.endswith("-> ()")
    f = str(torch.ops.ibis.step.default._schema)
    assert f.startswith("ibis::step(Tensor arena, Tensor(a1!) workspace") and f.endswith("-> Tensor")


def test_fake_step_keeps_the_arena_shape():
    from torch._subclasses.fake_tensor import FakeTensorMode
    with FakeTensorMode():
        a = torch.empty((L.NPLANE, 256), dtype=torch.float32)
        ws = torch.empty(1024, dtype=torch.uint8)
        tab = torch.empty(52, dtype=torch.float64)
        p = torch.empty(C.sizeof(abi.IbisParams), dtype=torch.uint8)
        out = torch.ops.ibis.step(a, ws, None, tab, p, 3600.0, 0.0, 200, abi.PREC_MIXED)
        assert out.shape == a.shape and out.dtype == a.dtype


def test_step_refuses_host_tensors():
    a = torch.zeros((L.NPLANE, 8))
    p = abi.params_bytes(abi.IbisParams())
    with pytest.raises(RuntimeError, match="GPU"):
        torch.ops.ibis.step_(a, torch.zeros(8, dtype=torch.uint8), None,
                             torch.zeros(52, dtype=torch.float64), p, 3600.0, 0.0, 8, 0)


def test_switch_variant_is_registered_and_checks_its_switch_block():
    s = str(torch.ops.ibis.step_ext.default._schema)
    assert s.startswith("ibis::step_ext(Tensor arena, Tensor(a1!) workspace, Tensor? slot_type, "
                        "Tensor tables, Tensor params, Tensor ext, Tensor? acclim_x, Tensor? td, "
                        "Tensor? stressfac, Tensor? cropfrac, float dtime")
    assert s.endswith("-> Tensor")
    x = abi.stomata_ext(default_sim_params().stomata_params)
    blk = ops.ext_block(x)
    assert blk.dtype == torch.float64 and blk.numel() == 10 and blk[0] == 0 and blk[2] == 0.15
    from torch._subclasses.fake_tensor import FakeTensorMode
    with FakeTensorMode():
        a = torch.empty((L.NPLANE, 64), dtype=torch.float32)
        out = torch.ops.ibis.step_ext(a, torch.empty(8, dtype=torch.uint8), None,
                                      torch.empty(4 * 52, dtype=torch.float64),
                                      torch.empty(C.sizeof(abi.IbisParams), dtype=torch.uint8),
                                      torch.empty(10, dtype=torch.float64), None, None, None,
                                      None, 3600.0, 0.0, 64, abi.PREC_MIXED)
        assert out.shape == a.shape
