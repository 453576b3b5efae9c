"""``IBISState`` over a device arena, with the reference's groups, field names and shapes.

The reference state is an Equinox pytree (``python-port/ibis/state/model_state.py:28-63``)
of sub-modules ``soil``, ``snow``, ``canopy``, ``atm``, ``vegetation``, ``hydrology``.  Here
the leaves the hourly step touches live in one plane-major torch tensor (``arena``,
``(NPLANE, pitch)``); ``state.soil.tsoi`` returns the ``(npoi, 6)`` view of its planes, so
a caller reads and writes fields by the reference's names without copies.  The per-cell
plant-type slot ids live in ``slot_type`` (``(12, pitch)`` int8).  Leaves that are not
per cell -- ``soil.hsoi`` (7,), ``vegetation.froot`` (6, 2), ``atm.co2conc``,
``vegetation.nytimes`` and the snow constants -- are host values on the state object.
"""
from __future__ import annotations

import numpy as np
import torch

from . import layout as L
from .params import SnowConstants

_GROUPS = ("soil", "snow", "canopy", "atm", "vegetation", "hydrology", "diag", "seasonal")
#: optional per-cell leaves kept beside the arena (read by the optional StomataParams
#: switches): the SeasonalState (python-port/ibis/state/seasonal.py:36-43) and the day's mean
#: air temperature AtmState.td
AUX_KEYS = ("seasonal.acclim_x", "seasonal.ssen_u", "seasonal.ssen_l", "atm.td", "stress.aystressu",
            "stress.aystressl", "stress.nday")


class _Group:
    """Attribute view of one sub-state: reads return tensor views, writes copy in."""

    def __init__(self, state, group):
        object.__setattr__(self, "_state", state)
        object.__setattr__(self, "_group", group)

    def __getattr__(self, name):
        key = f"{self._group}.{name}"
        if key in L.FIELD_INDEX:
            return self._state.field(key)
        st = self._state
        if key in st.aux:
            return st.aux[key][:st.n]
        if key == "soil.hsoi":
            return st.hsoi
        if key == "vegetation.froot":
            return st.froot
        if key == "vegetation.nytimes":
            return st.nytimes
        if key == "atm.co2conc":
            return st.co2conc
        if self._group == "snow" and hasattr(st.snow_const, name):
            return getattr(st.snow_const, name)
        raise AttributeError(f"{key} is not a leaf of the hourly step's state")

    def __setattr__(self, name, value):
        key = f"{self._group}.{name}"
        st = self._state
        if key in AUX_KEYS:
            if key not in st.aux:
                st.aux[key] = torch.zeros(st.pitch, dtype=st.dtype, device=st.device)
            st.aux[key][:st.n].copy_(torch.as_tensor(value, dtype=st.dtype, device=st.device).reshape(st.n))
            return
        if key not in L.FIELD_INDEX:
            raise AttributeError(f"{key} is not a per-cell leaf of the hourly step's state")
        view = self._state.field(key)
        view.copy_(torch.as_tensor(value, dtype=view.dtype, device=view.device).reshape(view.shape))


class IBISState:
    def __init__(self, arena: torch.Tensor, slot_type: torch.Tensor | None, n: int, hsoi, froot,
                 co2conc: float = 380.0e-6, nytimes: float = 0.0,
                 snow_const: SnowConstants | None = None, bedrock: bool = True,
                 aux: dict | None = None):
        if arena.dim() != 2 or arena.shape[0] != L.NPLANE or arena.shape[1] < n:
            raise ValueError(f"arena must be ({L.NPLANE}, pitch >= {n}), got {tuple(arena.shape)}")
        if slot_type is not None and (slot_type.shape[0] != L.NPFT or slot_type.shape[1] != arena.shape[1]
                                      or slot_type.dtype != torch.int8):
            raise ValueError("slot_type must be int8 of shape (12, pitch)")
        self.arena = arena
        self.slot_type = slot_type
        self.n = int(n)
        self.hsoi = np.asarray(hsoi, dtype=np.float64)
        self.froot = np.asarray(froot, dtype=np.float64)
        self.co2conc = float(co2conc)
        self.nytimes = float(nytimes)
        self.snow_const = snow_const or SnowConstants()
        self.bedrock = bool(bedrock)
        self.aux = dict(aux or {})
        for k, t in self.aux.items():
            if k not in AUX_KEYS or t.shape != (arena.shape[1],) or t.dtype != arena.dtype:
                raise ValueError(f"aux leaf {k!r} must be one of {AUX_KEYS}, shaped (pitch,) "
                                 "in the arena's dtype")
        self._work = {}

    # -- sub-states -------------------------------------------------------------------
    def __getattr__(self, name):
        if name in _GROUPS:
            return _Group(self, name)
        raise AttributeError(name)

    @property
    def pitch(self) -> int:
        return int(self.arena.shape[1])

    @property
    def device(self):
        return self.arena.device

    @property
    def dtype(self):
        return self.arena.dtype

    def field(self, key: str) -> torch.Tensor:
        f = L.FIELDS[L.FIELD_INDEX[key]]
        off = L.PLANE_OFFSET[key]
        if f.width == 1:
            return self.arena[off, :self.n]
        return self.arena[off:off + f.width, :self.n].t()

    def like(self, arena: torch.Tensor) -> "IBISState":
        """A state sharing everything but the arena."""
        return IBISState(arena, self.slot_type, self.n, self.hsoi, self.froot, self.co2conc,
                         self.nytimes, self.snow_const, self.bedrock, self.aux)

    def clone(self) -> "IBISState":
        new = self.like(self.arena.clone())
        new.aux = {k: t.clone() for k, t in self.aux.items()}
        return new

    # -- host conversion ----------------------------------------------------------------
    @classmethod
    def from_numpy(cls, fields: dict, slot_type, hsoi, froot, *, device="cuda",
                   dtype=torch.float32, pitch: int | None = None, co2conc=380.0e-6,
                   nytimes=0.0, snow_const=None, bedrock=True) -> "IBISState":
        n = int(np.asarray(next(iter(fields.values()))).shape[0])
        pitch = n if pitch is None else int(pitch)
        np_dt = np.float64 if dtype == torch.float64 else np.float32
        a = np.zeros((L.NPLANE, pitch), dtype=np_dt)
        for f in L.FIELDS:
            if f.key in fields:
                v = np.asarray(fields[f.key], dtype=np.float64)
                off = L.PLANE_OFFSET[f.key]
                if f.width == 1:
                    a[off, :n] = v.reshape(n)
                else:
                    a[off:off + f.width, :n] = v.reshape(n, f.width).T
        arena = torch.from_numpy(a).to(device)
        aux = {}
        for k in AUX_KEYS:
            if k in fields:
                v = np.zeros(pitch, dtype=np_dt)
                v[:n] = np.asarray(fields[k], dtype=np.float64).reshape(n)
                aux[k] = torch.from_numpy(v).to(device)
        st = None
        if slot_type is not None:
            s = np.full((L.NPFT, pitch), -1, dtype=np.int8)
            s[:, :n] = np.asarray(slot_type, dtype=np.int8).T
            st = torch.from_numpy(s).to(device)
        return cls(arena, st, n, hsoi, froot, co2conc, nytimes, snow_const, bedrock, aux)

    def to_numpy(self, keys=None) -> dict:
        a = self.arena[:, :self.n].detach().cpu().numpy().astype(np.float64)
        out = {}
        for f in L.FIELDS:
            if keys is not None and f.key not in keys:
                continue
            off = L.PLANE_OFFSET[f.key]
            out[f.key] = a[off] if f.width == 1 else a[off:off + f.width].T.copy()
        return out
