"""Parameter objects of the hourly step, with the reference's names and defaults.

Mirrors (F77 path, every optional StomataParams switch off):

* ``PhysicalConstants``  -- ``python-port/ibis/constants.py:16-79``
* ``StomataParams``      -- ``python-port/ibis/physics/stomata.py:33-161``
* ``CanopyParams``       -- ``python-port/ibis/physics/canopy_exchange.py:68-84``
* ``InterceptionParams`` -- ``python-port/ibis/physics/interception.py:31-45``
* ``VegRadParams``       -- ``python-port/ibis/physics/radiation.py:32-56``
* ``BedrockParams``      -- ``python-port/ibis/physics/bedrock.py:35-66``
* ``FireParams``         -- the per-step thresholds of ``python-port/ibis/physics/fire.py:203-270``
* ``SimParams``          -- ``python-port/ibis/timestepping/loops.py:72-134``

Plain dataclasses of Python floats and NumPy arrays: they are the host-side description;
``ibis_amd.abi.params_from`` packs them into the by-value C block.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class PhysicalConstants:
    stef: float = 5.67e-8
    vonk: float = 0.4
    grav: float = 9.80616
    tmelt: float = 273.16
    hvap: float = 2.5104e6
    hfus: float = 0.3336e6
    hsub: float = 2.5104e6 + 0.3336e6
    ch2o: float = 4.218e3
    cice: float = 2.106e3
    cair: float = 1.00464e3
    cvap: float = 1.81e3
    rair: float = 287.04
    rvap: float = 461.0
    cappa: float = 287.04 / 1.00464e3
    rhow: float = 1.0e3
    epsilon: float = 1.0e-7
    pi: float = 3.1415927


def get_constants() -> PhysicalConstants:
    return PhysicalConstants()


#: Vmax per IBIS PFT, python-port/ibis/io/physics_params.py:89-90
DEFAULT_VMAX = (65e-6, 65e-6, 40e-6, 30e-6, 30e-6, 25e-6, 30e-6, 30e-6, 27.5e-6, 27.5e-6,
                15e-6, 25e-6)


#: the boolean StomataParams switches the hourly step reads (stomata.py:93-161)
SWITCH_NAMES = ("vmax_mix", "scale_leaf_only", "acclim_on", "cap_scale_u", "cap_scale_l",
                "sunshade_u", "sunshade_l", "coefm_mix", "acclim_per_pft", "acclim_light",
                "frost_on")


@dataclass
class StomataParams:
    tau15: float = 4500.0
    kc15: float = 1.5e-4
    ko15: float = 0.25
    cimax: float = 2000.0e-6
    o2conc: float = 0.209
    alpha3: float = 0.080
    theta3: float = 0.950
    beta3: float = 0.990
    alpha4: float = 0.050
    theta4: float = 0.970
    beta4: float = 0.800
    # Ball-Berry groups: broadleaf, conifer, shrub, C4 grass, C3 grass
    gamma: np.ndarray = field(default_factory=lambda: np.array([0.015, 0.015, 0.015, 0.030, 0.015]))
    coefm: np.ndarray = field(default_factory=lambda: np.array([10.0, 6.0, 9.0, 4.0, 9.0]))
    coefb: np.ndarray = field(default_factory=lambda: np.array([0.010, 0.010, 0.010, 0.040, 0.010]))
    gsmin: np.ndarray = field(default_factory=lambda: np.array([1e-5] * 5))
    vmax_pft: np.ndarray = field(default_factory=lambda: np.array(DEFAULT_VMAX))
    #: scalar, or (npoi,) per cell (the per-zone value of python-port/ibis/io/pft_types.py:218-230)
    stressfac: float = -5.0
    # optional switches (python-port/ibis/physics/stomata.py:93-161); all off is the F77 path
    vmax_mix: bool = False
    scale_leaf_only: bool = False
    acclim_on: bool = False
    acclim_pfts: tuple = (3, 5, 8)
    acclim_tau: float = 10.93
    acclim_s0: float = -3.063
    acclim_smax: float = 17.72
    cap_scale_u: bool = False
    cap_scale_l: bool = False
    cap_kn: float = 0.15
    sunshade_u: bool = False
    sunshade_l: bool = False
    coefm_mix: bool = False
    coefm_pft: np.ndarray = field(default_factory=lambda: np.array(
        [10.0, 10.0, 10.0, 6.0, 10.0, 6.0, 10.0, 6.0, 9.0, 9.0, 4.0, 9.0]))
    acclim_per_pft: bool = False
    acclim_light: bool = False
    frost_on: bool = False
    frost_t0: float = -5.0
    frost_t1: float = 5.0
    gpp_split_by_capacity: bool = False
    acclim_s0_pft: np.ndarray = field(default_factory=lambda: np.full(12, -3.063))
    acclim_smax_pft: np.ndarray = field(default_factory=lambda: np.full(12, 17.72))

    def switches_on(self) -> bool:
        """Whether any optional switch (or a per-cell stressfac) leaves the F77 path."""
        return (any(bool(getattr(self, k)) for k in SWITCH_NAMES)
                or np.ndim(self.stressfac) != 0)


@dataclass
class CanopyParams:
    cleaf: float = 0.01
    cstem: float = 0.01
    cgrass: float = 0.01
    dleaf: tuple = (0.10, 0.10)
    dstem: tuple = (0.10, 0.10)
    alaiml: float = 8.0
    alaimu: float = 8.0
    chu: float = 8436.0
    chs: float = 2.109e5
    chl: float = 8436.0
    soil_beta_power: float = 1.0


@dataclass
class InterceptionParams:
    wliqumax: float = 0.20
    wliqsmax: float = 0.40
    wliqlmax: float = 0.20
    wsnoumax: float = 2.00
    wsnosmax: float = 4.00
    wsnolmax: float = 2.00
    tdripu: float = 7200.0
    tdrips: float = 7200.0
    tdripl: float = 7200.0
    tblowu: float = 43200.0
    tblows: float = 43200.0
    tblowl: float = 43200.0
    fmax: float = 0.25


@dataclass
class VegRadParams:
    rhoveg: np.ndarray = field(default_factory=lambda: np.array([[0.10, 0.10], [0.60, 0.40]]))
    tauveg: np.ndarray = field(default_factory=lambda: np.array([[0.07, 0.05], [0.25, 0.20]]))
    oriev: np.ndarray = field(default_factory=lambda: np.array([0.5, 0.0]))
    orieh: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0]))


@dataclass
class BedrockParams:
    hsoi: tuple = (2.0, 2.0, 4.0)
    consoi: float = 3.0
    rhosoi: float = 2700.0
    csoi: float = 740.74074074074076

    @property
    def nlayer(self) -> int:
        return len(self.hsoi)


@dataclass
class FireParams:
    d_crit: float = 0.5
    vpd_ref: float = 3.0
    d_crit_vpd: float = 0.33


@dataclass
class CropParams:
    """Cropland as managed grass (python-port/ibis/io/parameters.py:262-275): the grasses on a
    cell's cropland share photosynthesise with Vmax x vmax_scale_c3 (PFT 11) / _c4 (PFT 10);
    c4_share in 0..1 mixes the two scales on both grasses, -1 keeps them apart."""
    harvest_index: float = 0.45
    vmax_scale_c3: float = 1.0
    vmax_scale_c4: float = 1.0
    c4_share: float = -1.0


@dataclass
class SimParams:
    veg_rad: VegRadParams = field(default_factory=VegRadParams)
    canopy: CanopyParams = field(default_factory=CanopyParams)
    interception: InterceptionParams = field(default_factory=InterceptionParams)
    stomata_params: StomataParams = field(default_factory=StomataParams)
    fire: FireParams = field(default_factory=FireParams)
    wpudmax: float = 10.0
    zwpmax: float = 0.5
    bperm: float = 0.04
    h2o_consistent_weights: bool = False
    rhosoi_val: float = 2636.5
    csoi_mineral_val: float = 880.5
    bedrock: BedrockParams | None = None
    #: per-type overrides {type index 0..39: value} of Vmax, the stomatal slope and the
    #: acclimation constants, IBIS 3.0 zone types (python-port/ibis/io/pft_types.py:170-212)
    vmax_type_overrides: dict = field(default_factory=dict)
    #: cropland share per cell (SimParams.cropfrac, None: no cropland) and its parameters
    cropfrac: np.ndarray | None = None
    crop: CropParams = field(default_factory=CropParams)
    coefm_type_overrides: dict = field(default_factory=dict)
    acclim_s0_type_overrides: dict = field(default_factory=dict)
    acclim_smax_type_overrides: dict = field(default_factory=dict)


def default_sim_params(bedrock: bool = True) -> SimParams:
    return SimParams(bedrock=BedrockParams() if bedrock else None)


@dataclass
class SnowConstants:
    """Static snow parameters carried by SnowState (python-port/ibis/state/snow.py:34-40)."""
    z0sno: float = 0.0005
    rhos: float = 250.0
    consno: float = 0.20
    hsnotop: float = 0.05
    hsnomin: float = 0.15
    fimin: float = 0.00004
    fimax: float = 1.00


#: the reference's default soil layering above bedrock, m (python-port/ibis/io/soil_init.py:17)
SOIL_LAYERS = (0.10, 0.15, 0.25, 0.50, 1.00, 2.00)


def root_fractions(hsoi=SOIL_LAYERS, beta_lower: float = 0.950, beta_upper: float = 0.975):
    """Root fraction per soil layer, (nsoilay, 2): column 0 lower canopy, 1 upper canopy.

    Jackson et al. (1996) asymptotic profile 1 - beta^d at the layer bottoms (d in cm),
    normalised to the column (python-port/ibis/io/veg_init.py:78-103)."""
    depth = np.cumsum(np.asarray(hsoi, dtype=np.float64)) * 100.0
    out = np.zeros((len(depth), 2))
    for col, beta in ((0, beta_lower), (1, beta_upper)):
        cum = 1.0 - beta ** depth
        prev = np.concatenate([[0.0], cum[:-1]])
        out[:, col] = (cum - prev) / (1.0 - beta ** depth[-1])
    return out


# ---- the reference's parameter files ---------------------------------------------------

#: IBIS PFT names, long and short spellings (python-port/ibis/io/physics_params.py:39-60)
PFT_NAMES = ("tropical_broadleaf_evergreen", "tropical_broadleaf_deciduous",
             "warm_temperate_broadleaf_evergreen", "temperate_conifer_evergreen",
             "temperate_broadleaf_deciduous", "boreal_conifer_evergreen",
             "boreal_broadleaf_deciduous", "boreal_conifer_deciduous", "evergreen_shrub",
             "deciduous_shrub", "c4_grass", "c3_grass")
PFT_SHORT_NAMES = ("trop_bl_ever", "trop_bl_dec", "warm_bl_ever", "temp_con_ever",
                   "temp_bl_dec", "boreal_con_ever", "boreal_bl_dec", "boreal_con_dec",
                   "ever_shrub", "dec_shrub", "c4_grass", "c3_grass")
_PFT_INDEX = {**{n: i for i, n in enumerate(PFT_NAMES)},
              **{n: i for i, n in enumerate(PFT_SHORT_NAMES)}}
#: Ball-Berry groups of the yaml 'ball_berry' table, in IbisParams order
BALL_BERRY_GROUPS = ("broadleaf", "conifer", "shrub", "c4_grass", "c3_grass")
#: PFT -> Ball-Berry group (python-port/ibis/physics/stomata.py:654)
PFT_SLOPE_GROUP = (0, 0, 0, 1, 0, 1, 0, 1, 2, 2, 3, 4)
#: the 40 zone-type ids in type-index order (python-port/ibis/io/pft_types.py:50-92)
TYPE_IDS = ("T01", "T02", "T03", "T04", "T05", "T06", "T07", "T08", "T09", "T10", "T11", "T12",
            "T13", "T14", "T15", "T16", "T17", "T18", "T19", "T21", "S01", "S02", "S03", "S04",
            "S05", "S06", "S07", "G01", "G02", "G03", "G04", "G05", "G06", "G07", "G08", "W01",
            "W02", "W03", "C01", "C02")
#: yaml per-type keys the hourly step reads -> SimParams override dict
TYPE_KEYS = {"vmax": "vmax_type_overrides", "coefm": "coefm_type_overrides",
             "acclim_s0": "acclim_s0_type_overrides",
             "acclim_smax": "acclim_smax_type_overrides"}


def _pft_vector(table, defaults):
    out = [float(x) for x in defaults]
    for name, v in (table or {}).items():
        if name not in _PFT_INDEX:
            raise KeyError(f"unknown PFT name {name!r}; use one of {PFT_NAMES} or {PFT_SHORT_NAMES}")
        out[_PFT_INDEX[name]] = float(v)
    return np.array(out)


def sim_params_from_dict(d: dict, bedrock: bool = True) -> SimParams:
    """SimParams from a parsed parameter file of the reference (the sections the hourly step
    reads: stomata, veg_rad, canopy, interception, soil, the hourly fire thresholds and the
    per-type vmax / coefm / acclim_s0 / acclim_smax of 'types'), with the reference's defaults
    for absent keys (python-port/ibis/io/physics_params.py:129-247, :442-493;
    io/pft_types.py:246-258).  Per-type stressfac and other per-zone scalars need the zone map
    and are given per cell by the caller (``stomata_params.stressfac`` as an (npoi,) array)."""
    d = d or {}
    s = d.get("stomata", {}) or {}
    sp = StomataParams()
    for k in ("tau15", "kc15", "ko15", "cimax", "o2conc", "alpha3", "theta3", "beta3",
              "alpha4", "theta4", "beta4", "stressfac", "acclim_tau", "acclim_s0",
              "acclim_smax", "cap_kn", "frost_t0", "frost_t1"):
        if k in s:
            setattr(sp, k, float(s[k]))
    for k in SWITCH_NAMES + ("gpp_split_by_capacity",):
        setattr(sp, k, bool(s.get(k, False)))
    if "acclim_pfts" in s:
        sp.acclim_pfts = tuple(int(k) for k in s["acclim_pfts"])
    bb = s.get("ball_berry", {}) or {}
    for name in ("gamma", "coefm", "coefb", "gsmin"):
        base = getattr(sp, name)
        setattr(sp, name, np.array([float((bb.get(g, {}) or {}).get(name, base[i]))
                                    for i, g in enumerate(BALL_BERRY_GROUPS)]))
    vm = s.get("vmax", {}) or {}
    sp.vmax_pft = np.array([float(vm.get(n, DEFAULT_VMAX[i])) for i, n in enumerate(PFT_NAMES)])
    sp.coefm_pft = _pft_vector(s.get("coefm_pft"),
                               [sp.coefm[PFT_SLOPE_GROUP[i]] for i in range(12)])
    sp.acclim_s0_pft = _pft_vector(s.get("acclim_s0_pft"), [sp.acclim_s0] * 12)
    sp.acclim_smax_pft = _pft_vector(s.get("acclim_smax_pft"), [sp.acclim_smax] * 12)

    sim = default_sim_params(bedrock)
    sim.stomata_params = sp
    vr = d.get("veg_rad", {}) or {}
    for k in ("rhoveg", "tauveg", "oriev", "orieh"):
        if k in vr:
            setattr(sim.veg_rad, k, np.asarray(vr[k], dtype=np.float32).astype(np.float64))
    cp = d.get("canopy", {}) or {}
    for k, v in cp.items():
        if hasattr(sim.canopy, k):
            setattr(sim.canopy, k, tuple(float(x) for x in v) if isinstance(v, list) else float(v))
    ip = d.get("interception", {}) or {}
    for k, v in ip.items():
        if hasattr(sim.interception, k):
            setattr(sim.interception, k, float(v))
    so = d.get("soil", {}) or {}
    for k in ("wpudmax", "zwpmax", "bperm", "rhosoi_val", "csoi_mineral_val"):
        if k in so:
            setattr(sim, k, float(so[k]))
    sim.h2o_consistent_weights = bool(so.get("h2o_consistent_weights", False))
    cr = d.get("crop", {}) or {}
    for k, v in cr.items():
        if not hasattr(sim.crop, k):
            raise KeyError(f"unknown crop parameter {k!r}")
        setattr(sim.crop, k, float(v))
    fp = d.get("fire", {}) or {}
    for k in ("d_crit", "vpd_ref", "d_crit_vpd"):
        if k in fp:
            setattr(sim.fire, k, float(fp[k]))
    for tid, kv in (d.get("types", {}) or {}).items():
        if tid not in TYPE_IDS:
            raise KeyError(f"unknown plant type {tid!r}")
        for key, attr in TYPE_KEYS.items():
            if key in (kv or {}):
                getattr(sim, attr)[TYPE_IDS.index(tid)] = float(kv[key])
    return sim


def load_params_yaml(path: str, bedrock: bool = True) -> SimParams:
    """SimParams from one of the reference's parameter files (``configs/params/*.yaml``)."""
    import yaml
    with open(path) as fh:
        return sim_params_from_dict(yaml.safe_load(fh) or {}, bedrock=bedrock)


def release_switches(sim: SimParams) -> SimParams:
    """Set the switches of the reference's release parameter file on ``sim`` (in place):
    sunlit/shaded canopies, composition-weighted Vmax and slopes, acclimation with constants
    per plant type (types 3, 5, 8), the capacity-weighted GPP split and the consistent soil
    water weights (python-port/configs/params/params_31s_ch4h.yaml:55-90, :299).  The
    calibrated values themselves come from the file (``load_params_yaml``)."""
    sp = sim.stomata_params
    sp.acclim_on = sp.acclim_per_pft = True
    sp.acclim_pfts = (3, 5, 8)
    sp.sunshade_u = sp.sunshade_l = True
    sp.cap_kn = 0.15
    sp.coefm_mix = sp.vmax_mix = True
    sp.gpp_split_by_capacity = True
    sp.acclim_light = sp.frost_on = False
    sim.h2o_consistent_weights = True
    return sim
