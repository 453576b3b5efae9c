"""CPU FP64 restatement of IBIS 3.0's hourly land-surface step -- TEST INFRASTRUCTURE ONLY.

This package is the checker for the HIP step.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it, and only to check or to time a
baseline: the product path never calls it and never falls back to it.

It restates ``timestep_body`` (``python-port/ibis/timestepping/loops.py:234-1284``) and the
hourly physics it calls, vectorised over cells with NumPy in float64, one module per
reference module:

========================  ==========================================================
``oracle.thermo``         ``physics/thermo.py`` (Lowe 1977 SVP, qsat, latent heats)
``oracle.solvers``        ``physics/solvers.py`` (Thomas) and ``canopy_exchange.py:26-61``
``oracle.radiation``      ``physics/radiation.py`` (two-stream, albedos, IR)
``oracle.interception``   ``physics/interception.py`` (fwetcal, cascade, cascad2, noveg)
``oracle.canopy``         ``physics/canopy_exchange.py`` (canini, turcof, turvap)
``oracle.stomata``        ``physics/stomata.py`` (Farquhar / Ball-Berry, drystress)
``oracle.snow``           ``physics/snow_physics.py``
``oracle.soil``           ``physics/soil_physics.py``, ``physics/bedrock.py``
``oracle.fire``           per-step accumulators of ``physics/fire.py``
``oracle.step``           ``timestep_body`` itself
``oracle.params``         reference default constants and parameters
========================  ==========================================================

Pinning (what checks each function independently of this restatement) is listed in each
module's header; functions with no independent check say "parity unpinned".
"""
