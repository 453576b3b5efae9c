"""Restart files in the reference's on-disk format, so states move between its runs and ours.

The reference writes a checkpoint as one ``.npz`` (python-port/ibis/io/restart.py:19-55):
``treedef_repr`` (the ``repr`` of the state pytree's structure), ``num_leaves``, one
``leaf_<i>`` array per leaf of the flattened ``IBISState`` (``leaf_<i>_is_none`` for a None
leaf) and ``meta_<key>`` values.  Leaves are flattened in the field order of the state
dataclasses: ``soil``, ``snow``, ``canopy``, ``atm``, ``vegetation``, ``carbon``,
``hydrology``, ``summary``, then the optional modules (python-port/ibis/state/
model_state.py:28-63), each sub-state's leaves in its own field order (``state/soil.py``,
``snow.py``, ``canopy.py``, ``atmosphere.py``, ``vegetation.py``, ``hydrology.py``).

``Restart`` reads such a file without executing anything from it (``numpy.load`` with
``allow_pickle=False``), checks the structure recorded in ``treedef_repr`` against the field
lists below, and maps the leaves the hourly step uses onto an ``IBISState`` arena.  Every
other leaf (carbon pools, summaries, phenology, optional modules) is carried through
untouched, so ``Restart.update(state)`` followed by ``save`` writes a file the reference's
``load_state`` reads as it reads its own: same structure, same leaf dtypes, same metadata.

Optional trailing modules.  The reference's optional modules form one nested chain
(StressAccum 4 leaves -- 3 in files written before its smoothed-stress leaf existed --,
ETParts 4, SeasonalState 3, NEEParts 2) after an optional wetland state of 3 leaves; its
loader aligns a checkpoint anywhere on that chain with a template
(python-port/ibis/io/restart.py:57-170).  The leaves this module maps all lie before the
chain, so a checkpoint loads here at any chain position; ``trailing`` reports which modules
it carries, and a structure off the chain is refused as the reference refuses it.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import layout as L
from .params import SnowConstants

#: field order of each sub-state dataclass (python-port/ibis/state/*.py)
FIELD_ORDER = {
    "soil": "tsoi wsoi wisoi tg wpud wipud consoi csoi rhosoi hflo qglif heatg hvasug z0soi "
            "stresstl stresstu stressl stressu upsoil upsoiu porosflo poros hydraul suction "
            "bex sfield swilt sand clay hsoi albsav albsan tsoi_deep".split(),
    "snow": "fi tsno hsno ti tlsub heati hvasui z0sno rhos consno hsnotop hsnomin fimin "
            "fimax".split(),
    "canopy": "tu ts tl t12 t34 q12 q34 wliqu wliqs wliql wsnou wsnos wsnol ci_upper gs_upper "
              "cs_upper an_upper ag_upper ci_lower gs_lower cs_lower an_lower ag_lower fwetu "
              "fwets fwetl rliqu rliqs rliql solu sols soll solg soli firu firs firl firg firi "
              "firb asurd asuri topparu topparl scalcoefu scalcoefl fsenu fsens fsenl fseng "
              "fseni fvapg fvapi raing snowg a10scalparamu a10daylightu a10scalparaml "
              "a10daylightl rhoa".split(),
    "atm": "ta qa ua psurf coszen solad solai fira raina snowa td tmax tmin qd ud precip cloud "
           "rh xstore iwet co2conc o2conc".split(),
    "vegetation": "cbiol cbior cbiow plai frac biomass exist tnpp tgpp lai sai ztop zbot fu fl "
                  "froot tempu templ dropu dropls dropl3 dropl4 agddu agddl tc tw gdd0 gdd5 "
                  "tcmin tnpptot tgpptot tco2mic tco2root tneetot sapfrac firefac nytimes "
                  "nfirestep sdry nfirestep_vpd sdry_vpd totlaiu totlail totbiou totbiol_agg "
                  "disturbf disturbo vegtype0 tnmin gdd0this gdd5this tcthis twthis specla "
                  "aleaf aroot awood".split(),
    "hydrology": "ginvap gsuvap gtrans gtransu gtransl grunof gdrain gadjust wtot raing_avg "
                 "traing_avg fhtop_soil rainl snowl trainl tsnowl sflo_bot fhtop_snow".split(),
}
NODE = {"soil": "SoilState", "snow": "SnowState", "canopy": "CanopyState", "atm": "AtmState",
        "vegetation": "VegetationState", "carbon": "CarbonState", "hydrology": "HydrologyDiag",
        "summary": "SummaryState"}
#: IBISState's sub-states in field order; the optional modules follow
#: (permafrost, wetland, fire, bvoc, stress, etparts, seasonal, neeparts)
SUBSTATES = ("soil", "snow", "canopy", "atm", "vegetation", "carbon", "hydrology", "summary")
#: SeasonalState's leaves in field order (python-port/ibis/state/seasonal.py:36-43)
SEASONAL_FIELDS = ("acclim_x", "ssen_u", "ssen_l")
#: the optional modules whose leading leaves map to the state's aux leaves: module name ->
#: (aux prefix, leaves in field order); StressAccum's fourth leaf, u_smooth, is the annual
#: driver's (python-port/ibis/state/stress_accum.py:15-21)
AUX_MODULES = {"SeasonalState": ("seasonal", SEASONAL_FIELDS),
               "StressAccum": ("stress", ("aystressu", "aystressl", "nday"))}
#: leaf counts of the optional trailing chain, and of its first module in old files
CHAIN_LEAVES = (4, 4, 3, 2)
OLD_STRESS_LEAVES = 3
WETLAND_LEAVES = 3


# ---- the pytree structure string ------------------------------------------------------------

def parse_treedef(text: str):
    """Parse ``repr(treedef)``: nodes are ("node", name, children), ("leaf",) or ("none",)."""
    s = text.strip()
    pos = 0

    def expect(tok):
        nonlocal pos
        if not s.startswith(tok, pos):
            raise ValueError(f"treedef_repr: expected {tok!r} at {pos}: {s[pos:pos + 40]!r}")
        pos += len(tok)

    def skip_ws():
        nonlocal pos
        while pos < len(s) and s[pos] in " ,":
            pos += 1

    def node():
        nonlocal pos
        skip_ws()
        if s.startswith("*", pos):
            pos += 1
            return ("leaf",)
        if s.startswith("None", pos):
            pos += 4
            return ("none",)
        expect("CustomNode(")
        j = s.index("[", pos)
        name = s[pos:j]
        depth, pos = 0, j
        while True:                                   # the auxiliary data, balanced brackets
            c = s[pos]
            depth += (c in "[(") - (c in "])")
            pos += 1
            if depth == 0:
                break
        skip_ws()
        expect("[")
        kids = []
        skip_ws()
        while not s.startswith("]", pos):
            kids.append(node())
            skip_ws()
        expect("]")
        expect(")")
        return ("node", name, kids)

    if s.startswith("PyTreeDef("):
        pos = len("PyTreeDef(")
    root = node()
    return root


def count_leaves(n) -> int:
    if n[0] == "leaf":
        return 1
    if n[0] == "none":
        return 0
    return sum(count_leaves(k) for k in n[2])


# ---- the file ---------------------------------------------------------------------------------

class Restart:
    """One restart file: its leaves, structure and metadata, and the step's fields in it."""

    def __init__(self, leaves, treedef_repr, meta=None):
        self.leaves = list(leaves)
        self.treedef_repr = treedef_repr
        self.meta = dict(meta or {})
        root = parse_treedef(treedef_repr)
        if root[0] != "node" or root[1] != "IBISState":
            raise ValueError("restart: the structure's root is not an IBISState")
        if count_leaves(root) != len(self.leaves):
            raise ValueError(f"restart: {len(self.leaves)} leaves but the structure has "
                             f"{count_leaves(root)}")
        kids = root[2]
        if len(kids) < len(SUBSTATES):
            raise ValueError("restart: fewer sub-states than IBISState declares")
        self.offset, start = {}, 0
        for name, k in zip(SUBSTATES, kids):
            if k[0] != "node" or k[1] != NODE[name]:
                raise ValueError(f"restart: sub-state {name} is {k[:2]}, expected {NODE[name]}")
            n = count_leaves(k)
            if name in FIELD_ORDER and n != len(FIELD_ORDER[name]):
                raise ValueError(f"restart: {NODE[name]} has {n} leaves, expected "
                                 f"{len(FIELD_ORDER[name])}")
            self.offset[name] = start
            start += n
        self.trailing = [(k[1] if k[0] == "node" else None, count_leaves(k))
                         for k in kids[len(SUBSTATES):]]
        self.module_offset = {}
        for name, n in self.trailing:
            if name is not None:
                self.module_offset[name] = start
            start += n
        self._check_chain()
        self.ncell = int(np.asarray(self.leaf("soil.tg")).shape[0])

    # -- optional modules ---------------------------------------------------------------------
    def _check_chain(self):
        counts = [c for _, c in self.trailing if c]
        total = sum(counts)
        chains = set()
        for first in (CHAIN_LEAVES[0], OLD_STRESS_LEAVES):
            acc = 0
            chains.add(0)
            for c in (first,) + CHAIN_LEAVES[1:]:
                acc += c
                chains.add(acc)
        ok = total in chains or (total - WETLAND_LEAVES) in chains
        if not ok:
            raise ValueError(f"restart: optional modules carry {total} leaves, not a position "
                             f"of the optional-module chain")

    # -- io -----------------------------------------------------------------------------------
    @classmethod
    def load(cls, path: str) -> "Restart":
        if not path.endswith(".npz"):
            path = path + ".npz"
        with np.load(path, allow_pickle=False) as z:
            num = int(z["num_leaves"])
            leaves = [None if f"leaf_{i}_is_none" in z.files else np.array(z[f"leaf_{i}"])
                      for i in range(num)]
            meta = {}
            for key in z.files:
                if key.startswith("meta_"):
                    v = np.array(z[key])
                    meta[key[5:]] = v.item() if v.ndim == 0 else v
            tree = str(z["treedef_repr"])
        return cls(leaves, tree, meta)

    def save(self, path: str, metadata: dict | None = None) -> None:
        """Write in the reference's layout (np.savez_compressed, as restart.py:53)."""
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        out = {"treedef_repr": self.treedef_repr, "num_leaves": len(self.leaves)}
        for i, leaf in enumerate(self.leaves):
            if leaf is None:
                out[f"leaf_{i}_is_none"] = True
            else:
                out[f"leaf_{i}"] = np.asarray(leaf)
        meta = dict(self.meta)
        meta.update(metadata or {})
        for k, v in meta.items():
            out[f"meta_{k}"] = np.array(v)
        np.savez_compressed(path, **out)

    # -- leaves -------------------------------------------------------------------------------
    def index(self, key: str) -> int:
        group, name = key.split(".")
        return self.offset[group] + FIELD_ORDER[group].index(name)

    def leaf(self, key: str):
        return self.leaves[self.index(key)]

    def step_keys(self):
        """Per-cell fields of the step's arena that the file carries."""
        return [f.key for f in L.FIELDS if f.group in FIELD_ORDER
                and f.key.split(".")[1] in FIELD_ORDER[f.group]]

    def fields(self) -> dict:
        """The step's per-cell fields, reference-shaped float64."""
        out = {}
        for key in self.step_keys():
            f = L.FIELDS[L.FIELD_INDEX[key]]
            a = np.asarray(self.leaf(key), dtype=np.float64)
            want = (self.ncell,) if f.width == 1 else (self.ncell, f.width)
            if a.shape != want:
                raise ValueError(f"restart: {key} has shape {a.shape}, expected {want}")
            out[key] = a
        return out

    def aux_fields(self) -> dict:
        """The optional per-cell leaves the step's switches read, where the file has them:
        the day's mean temperature ``atm.td`` and, in a file carrying the SeasonalState module
        (state/seasonal.py:36-43), ``seasonal.acclim_x``, ``ssen_u``, ``ssen_l``; with the
        StressAccum module (state/stress_accum.py:15-21), ``stress.aystressu``, ``aystressl``,
        ``nday``."""
        out = {"atm.td": np.asarray(self.leaf("atm.td"), dtype=np.float64).reshape(self.ncell)}
        for mod, (pre, names) in AUX_MODULES.items():
            o = self.module_offset.get(mod)
            if o is not None:
                for j, k in enumerate(names):
                    out[f"{pre}.{k}"] = np.asarray(self.leaves[o + j],
                                                   dtype=np.float64).reshape(self.ncell)
        return out

    def to_state(self, slot_type=None, *, device="cuda", dtype=torch.float32, pitch=None,
                 bedrock=None):
        """An ``IBISState`` holding the file's step fields.  The snow constants come from the
        defaults, not the file, as the reference's loader takes them from its template
        (restart.py:185-195)."""
        from .state import IBISState
        hsoi = np.asarray(self.leaf("soil.hsoi"), dtype=np.float64)
        froot = np.asarray(self.leaf("vegetation.froot"), dtype=np.float64)
        co2 = np.asarray(self.leaf("atm.co2conc"), dtype=np.float64).reshape(-1)[0]
        nyt = np.asarray(self.leaf("vegetation.nytimes"), dtype=np.float64).reshape(-1)[0]
        if bedrock is None:
            bedrock = self.leaf("soil.tsoi_deep") is not None
        return IBISState.from_numpy({**self.fields(), **self.aux_fields()}, slot_type, hsoi,
                                    froot, device=device,
                                    dtype=dtype, pitch=pitch, co2conc=float(co2),
                                    nytimes=float(nyt), snow_const=SnowConstants(),
                                    bedrock=bool(bedrock))

    def update(self, state) -> None:
        """Write the state's step fields (and ``nytimes``) back into the leaves, each in its
        leaf's dtype and shape; every other leaf is left as it is."""
        if state.n != self.ncell:
            raise ValueError(f"restart: state has {state.n} cells, file {self.ncell}")
        vals = state.to_numpy(keys=self.step_keys())
        for key in self.step_keys():
            i = self.index(key)
            old = self.leaves[i]
            self.leaves[i] = np.asarray(vals[key]).reshape(old.shape).astype(old.dtype)
        i = self.index("vegetation.nytimes")
        old = np.asarray(self.leaves[i])
        self.leaves[i] = np.full(old.shape, state.nytimes, dtype=old.dtype)
        for mod, (pre, names) in AUX_MODULES.items():
            o = self.module_offset.get(mod)
            for j, k in enumerate(names):
                t = state.aux.get(f"{pre}.{k}")
                if o is not None and t is not None:
                    old = self.leaves[o + j]
                    v = t[:state.n].detach().cpu().numpy()
                    self.leaves[o + j] = np.asarray(v).reshape(np.shape(old)).astype(old.dtype)
