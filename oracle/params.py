"""Reference default constants and parameters, restated for the oracle.

Values follow the reference's code defaults (the F77 path: every optional switch off):

* ``PhysicalConstants``        -- ``python-port/ibis/constants.py:16-79``
* ``StomataParams``            -- ``python-port/ibis/physics/stomata.py:33-161``
* ``CanopyParams``             -- ``python-port/ibis/physics/canopy_exchange.py:68-84``
* ``InterceptionParams``       -- ``python-port/ibis/physics/interception.py:31-45``
* ``VegRadParams``             -- ``python-port/ibis/physics/radiation.py:32-56``
* ``BedrockParams``            -- ``python-port/ibis/physics/bedrock.py:35-66``
* soil constants of SimParams  -- ``python-port/ibis/timestepping/loops.py:86-105``
  with ``RHOSOI_VAL``/``CSOI_VAL`` of ``python-port/ibis/io/soil_init.py`` (2636.5, 880.5)
* fire thresholds              -- ``python-port/ibis/physics/fire.py:203-270``
* snow constants               -- ``python-port/ibis/state/snow.py:34-40``
* layer thicknesses            -- ``python-port/ibis/io/soil_init.py:17`` and root profile
  ``python-port/ibis/io/veg_init.py:78-103``

Attribute names are the reference's, so the product's parameter objects can be handed to
the oracle unchanged and the product defaults can be compared against these.
"""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np


def physical_constants():
    hvap, hfus = 2.5104e6, 0.3336e6
    return SimpleNamespace(
        stef=5.67e-8, vonk=0.4, grav=9.80616, tmelt=273.16, hvap=hvap, hfus=hfus,
        hsub=hvap + hfus, ch2o=4.218e3, cice=2.106e3, cair=1.00464e3, cvap=1.81e3,
        rair=287.04, rvap=461.0, cappa=287.04 / 1.00464e3, rhow=1.0e3, epsilon=1.0e-7,
        pi=3.1415927,
    )


def stomata_params():
    return SimpleNamespace(
        tau15=4500.0, kc15=1.5e-4, ko15=0.25, cimax=2000.0e-6, o2conc=0.209,
        alpha3=0.080, theta3=0.950, beta3=0.990,
        alpha4=0.050, theta4=0.970, beta4=0.800,
        # Ball-Berry groups: broadleaf, conifer, shrub, C4 grass, C3 grass
        gamma=np.array([0.015, 0.015, 0.015, 0.030, 0.015]),
        coefm=np.array([10.0, 6.0, 9.0, 4.0, 9.0]),
        coefb=np.array([0.010, 0.010, 0.010, 0.040, 0.010]),
        gsmin=np.array([1e-5, 1e-5, 1e-5, 1e-5, 1e-5]),
        vmax_pft=np.array([65e-6, 65e-6, 40e-6, 30e-6, 30e-6, 25e-6,
                           30e-6, 30e-6, 27.5e-6, 27.5e-6, 15e-6, 25e-6]),
        stressfac=-5.0,
    )


def canopy_params():
    return SimpleNamespace(
        cleaf=0.01, cstem=0.01, cgrass=0.01, dleaf=(0.10, 0.10), dstem=(0.10, 0.10),
        alaiml=8.0, alaimu=8.0, chu=8436.0, chs=2.109e5, chl=8436.0, soil_beta_power=1.0,
    )


def interception_params():
    return SimpleNamespace(
        wliqumax=0.20, wliqsmax=0.40, wliqlmax=0.20, wsnoumax=2.00, wsnosmax=4.00,
        wsnolmax=2.00, tdripu=7200.0, tdrips=7200.0, tdripl=7200.0, tblowu=43200.0,
        tblows=43200.0, tblowl=43200.0, fmax=0.25,
    )


def veg_rad_params():
    # rhoveg/tauveg indexed [band, story] with story 0 = lower, 1 = upper
    return SimpleNamespace(
        rhoveg=np.array([[0.10, 0.10], [0.60, 0.40]]),
        tauveg=np.array([[0.07, 0.05], [0.25, 0.20]]),
        oriev=np.array([0.5, 0.0]),
        orieh=np.array([0.0, 0.0]),
    )


def bedrock_params():
    return SimpleNamespace(hsoi=(2.0, 2.0, 4.0), consoi=3.0, rhosoi=2700.0,
                           csoi=740.74074074074076)


def fire_params():
    return SimpleNamespace(d_crit=0.5, vpd_ref=3.0, d_crit_vpd=0.33)


def snow_constants():
    return SimpleNamespace(z0sno=0.0005, rhos=250.0, consno=0.20, hsnotop=0.05,
                           hsnomin=0.15, fimin=0.00004, fimax=1.00)


HSOI_LAYERS = np.array([0.10, 0.15, 0.25, 0.50, 1.00, 2.00])


def root_fractions(hsoi=HSOI_LAYERS, beta_lower=0.950, beta_upper=0.975):
    """froot (nsoilay, 2): column 0 lower canopy, 1 upper (io/veg_init.py:78-103)."""
    depth = np.cumsum(hsoi) * 100.0
    out = np.zeros((len(hsoi), 2))
    for col, beta in ((0, beta_lower), (1, beta_upper)):
        cum = 1.0 - beta ** depth
        prev = np.concatenate([[0.0], cum[:-1]])
        out[:, col] = (cum - prev) / (1.0 - beta ** depth[-1])
    return out


def sim_params():
    return SimpleNamespace(
        veg_rad=veg_rad_params(), canopy=canopy_params(), interception=interception_params(),
        stomata_params=stomata_params(), fire=fire_params(), bedrock=bedrock_params(),
        wpudmax=10.0, zwpmax=0.5, bperm=0.04, h2o_consistent_weights=False,
        rhosoi_val=2636.5, csoi_mineral_val=880.5,
    )
