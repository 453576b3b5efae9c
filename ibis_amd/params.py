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
    stressfac: float = -5.0


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
    #: per-type overrides of Vmax {type index 0..39: value}, IBIS 3.0 zone types
    #: (python-port/ibis/io/pft_types.py:170-189)
    vmax_type_overrides: dict = field(default_factory=dict)


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
