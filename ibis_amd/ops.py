"""The hourly step as PyTorch custom operators over the C ABI (``include/ibis_step.h``).

    torch.ops.ibis.step(arena, workspace, slot_type, type_table, params, dtime, nytimes,
                        ncell, flags) -> Tensor        # functional: a new arena
    torch.ops.ibis.step_(arena, workspace, slot_type, type_table, params, dtime, nytimes,
                         ncell, flags) -> ()           # in place

    torch.ops.ibis.step_ext(arena, workspace, slot_type, tables, params, ext, acclim_x, td,
                            stressfac, cropfrac, dtime, nytimes, ncell, flags) -> Tensor

``arena`` is the plane-major state (``(IBIS_NPLANE, pitch)``, float32 or float64, on the
GPU), ``workspace`` a uint8 device tensor of ``ibis_step_workspace_bytes`` bytes,
``slot_type`` the int8 ``(12, pitch)`` plant-type slots or None, ``type_table`` the 52
float64 Vmax rows on the device, ``params`` the ``IbisParams`` block as a CPU uint8 tensor
(``ibis_amd.abi.params_bytes``), ``flags`` the precision and bedrock bits of ``ibis_step``.
``step_ext`` is ``ibis_step_ext``: ``tables`` the IBIS_NTABLE x 52 per-slot tables on the
device, ``ext`` the switch block of ``ext_block`` (CPU float64), ``acclim_x`` / ``td`` /
``stressfac`` / ``cropfrac`` optional (pitch,) device planes of the arena's dtype.
Both run on the current stream of the arena's device and synchronise nothing.  They are
registered with ``torch.library`` (the Python form of TORCH_LIBRARY), with fake
implementations for tracing.  Replaces ``timestep_body``
(python-port/ibis/timestepping/loops.py:234-1284).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from . import _lib
from . import abi

_NS = "ibis"


def _run(arena_in: torch.Tensor, arena_out: torch.Tensor, workspace: torch.Tensor,
         slot_type: Optional[torch.Tensor], type_table: torch.Tensor, params: torch.Tensor,
         dtime: float, nytimes: float, ncell: int, flags: int) -> None:
    if arena_in.device.type != "cuda":
        raise RuntimeError("ibis::step runs on the GPU")
    if params.device.type != "cpu" or params.dtype != torch.uint8 or \
            params.numel() != C.sizeof(abi.IbisParams):
        raise ValueError("params must be the IbisParams block as a CPU uint8 tensor")
    for t in (arena_in, arena_out, workspace, type_table):
        if not t.is_contiguous():
            raise ValueError("ibis::step needs contiguous tensors")
    p = abi.IbisParams.from_buffer_copy(params.numpy().tobytes())
    lib = _lib.load()
    pitch = arena_in.shape[1]
    ain = abi.IbisArena(arena_in.data_ptr(), pitch)
    aout = abi.IbisArena(arena_out.data_ptr(), pitch)
    stream = torch.cuda.current_stream(arena_in.device)
    rc = lib.ibis_step(C.byref(ain), C.byref(aout), C.c_void_p(workspace.data_ptr()),
                       C.c_int64(workspace.numel()),
                       C.c_void_p(slot_type.data_ptr() if slot_type is not None else 0),
                       C.c_int64(pitch), C.c_void_p(type_table.data_ptr()), C.byref(p),
                       C.c_double(dtime), C.c_double(nytimes), C.c_int64(ncell),
                       C.c_int(flags), C.c_void_p(stream.cuda_stream))
    _lib.check(rc, "ibis::step")


@torch.library.custom_op(f"{_NS}::step_", mutates_args=("arena", "workspace"))
def step_(arena: torch.Tensor, workspace: torch.Tensor, slot_type: Optional[torch.Tensor],
          type_table: torch.Tensor, params: torch.Tensor, dtime: float, nytimes: float,
          ncell: int, flags: int) -> None:
    _run(arena, arena, workspace, slot_type, type_table, params, dtime, nytimes, ncell, flags)


@torch.library.custom_op(f"{_NS}::step", mutates_args=("workspace",))
def step(arena: torch.Tensor, workspace: torch.Tensor, slot_type: Optional[torch.Tensor],
         type_table: torch.Tensor, params: torch.Tensor, dtime: float, nytimes: float,
         ncell: int, flags: int) -> torch.Tensor:
    out = torch.empty_like(arena)
    _run(arena, out, workspace, slot_type, type_table, params, dtime, nytimes, ncell,
         flags | abi.FLAG_COPY_IN)
    return out


@step_.register_fake
def _(arena, workspace, slot_type, type_table, params, dtime, nytimes, ncell, flags):
    return None


@step.register_fake
def _(arena, workspace, slot_type, type_table, params, dtime, nytimes, ncell, flags):
    return torch.empty_like(arena)


_EXT_SCALARS = ("switches", "acclim_mask", "cap_kn", "acclim_s0", "acclim_smax", "frost_t0",
                "frost_t1", "crop_c4_share", "crop_vmax_scale_c3", "crop_vmax_scale_c4")


def ext_block(x: "abi.IbisStomataExt") -> torch.Tensor:
    """The scalar part of an IbisStomataExt as the CPU float64 ``ext`` argument of step_ext."""
    return torch.tensor([float(getattr(x, k)) for k in _EXT_SCALARS], dtype=torch.float64)


def _run_ext(arena_in, arena_out, workspace, slot_type, tables, params, ext, acclim_x, td,
             stressfac, cropfrac, dtime, nytimes, ncell, flags):
    if arena_in.device.type != "cuda":
        raise RuntimeError("ibis::step_ext runs on the GPU")
    if params.device.type != "cpu" or params.dtype != torch.uint8 or \
            params.numel() != C.sizeof(abi.IbisParams):
        raise ValueError("params must be the IbisParams block as a CPU uint8 tensor")
    if ext.device.type != "cpu" or ext.dtype != torch.float64 or ext.numel() != len(_EXT_SCALARS):
        raise ValueError("ext must be the switch block of ibis_amd.ops.ext_block")
    if tables.dtype != torch.float64 or tables.numel() != abi.NTABLE * abi.TYPE_ROWS:
        raise ValueError("tables must hold IBIS_NTABLE x 52 float64 rows")
    pitch = arena_in.shape[1]
    for name, t in (("acclim_x", acclim_x), ("td", td), ("stressfac", stressfac),
                    ("cropfrac", cropfrac)):
        if t is not None and (t.shape != (pitch,) or t.dtype != arena_in.dtype or
                              t.device != arena_in.device):
            raise ValueError(f"{name} must be a (pitch,) device plane of the arena's dtype")
    for t in (arena_in, arena_out, workspace, tables):
        if not t.is_contiguous():
            raise ValueError("ibis::step_ext needs contiguous tensors")
    p = abi.IbisParams.from_buffer_copy(params.numpy().tobytes())
    v = ext.tolist()
    x = abi.IbisStomataExt()
    x.switches, x.acclim_mask = int(v[0]), int(v[1])
    (x.cap_kn, x.acclim_s0, x.acclim_smax, x.frost_t0, x.frost_t1, x.crop_c4_share,
     x.crop_vmax_scale_c3, x.crop_vmax_scale_c4) = v[2:]
    ptr = lambda t: t.data_ptr() if t is not None else 0
    x.type_tables, x.acclim_x, x.td, x.stressfac, x.cropfrac = (
        tables.data_ptr(), ptr(acclim_x), ptr(td), ptr(stressfac), ptr(cropfrac))
    lib = _lib.load()
    ain = abi.IbisArena(arena_in.data_ptr(), pitch)
    aout = abi.IbisArena(arena_out.data_ptr(), pitch)
    stream = torch.cuda.current_stream(arena_in.device)
    rc = lib.ibis_step_ext(C.byref(ain), C.byref(aout), C.c_void_p(workspace.data_ptr()),
                           C.c_int64(workspace.numel()),
                           C.c_void_p(slot_type.data_ptr() if slot_type is not None else 0),
                           C.c_int64(pitch), C.byref(p), C.byref(x), C.c_double(dtime),
                           C.c_double(nytimes), C.c_int64(ncell), C.c_int(flags),
                           C.c_void_p(stream.cuda_stream))
    _lib.check(rc, "ibis::step_ext")


@torch.library.custom_op(f"{_NS}::step_ext", mutates_args=("workspace",))
def step_ext(arena: torch.Tensor, workspace: torch.Tensor, slot_type: Optional[torch.Tensor],
             tables: torch.Tensor, params: torch.Tensor, ext: torch.Tensor,
             acclim_x: Optional[torch.Tensor], td: Optional[torch.Tensor],
             stressfac: Optional[torch.Tensor], cropfrac: Optional[torch.Tensor], dtime: float,
             nytimes: float, ncell: int, flags: int) -> torch.Tensor:
    out = torch.empty_like(arena)
    _run_ext(arena, out, workspace, slot_type, tables, params, ext, acclim_x, td, stressfac,
             cropfrac, dtime, nytimes, ncell, flags | abi.FLAG_COPY_IN)
    return out


@step_ext.register_fake
def _(arena, workspace, slot_type, tables, params, ext, acclim_x, td, stressfac, cropfrac, dtime,
      nytimes, ncell, flags):
    return torch.empty_like(arena)
