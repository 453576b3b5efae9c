"""ctypes mirror of the C ABI (``include/ibis_step.h``) and arena packing.

``IbisParams`` is the by-value parameter block; ``params_from`` fills it from parameter
objects carrying the reference's attribute names (``PhysicalConstants``,
``SimParams.{stomata_params, canopy, interception, veg_rad, bedrock, fire}`` and the
soil constants, ``python-port/ibis/timestepping/loops.py:72-134``).  ``pack``/``unpack``
convert between reference-shaped arrays ``(npoi,)``/``(npoi, k)`` and the plane-major
arena of ``ibis_amd.layout``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import layout as L

PREC_MIXED, PREC_F32, PREC_F64 = 0, 1, 2
FLAG_BEDROCK = 0x10
FLAG_COPY_IN = 0x20
NTYPE = 40
TYPE_ROWS = NTYPE + L.NPFT

# optional StomataParams switches (include/ibis_step.h, python-port/ibis/physics/stomata.py:93-161)
SW_VMAX_MIX, SW_COEFM_MIX, SW_SUNSHADE_U, SW_SUNSHADE_L = 0x001, 0x002, 0x004, 0x008
SW_CAP_SCALE_U, SW_CAP_SCALE_L, SW_SCALE_LEAF_ONLY = 0x010, 0x020, 0x040
SW_ACCLIM, SW_ACCLIM_PER_PFT, SW_ACCLIM_LIGHT, SW_FROST = 0x080, 0x100, 0x200, 0x400
SW_CROP = 0x800
SW_ALL = 0xfff
#: StomataParams switch name -> bit (acclim_on maps to SW_ACCLIM)
SWITCHES = {"vmax_mix": SW_VMAX_MIX, "coefm_mix": SW_COEFM_MIX, "sunshade_u": SW_SUNSHADE_U,
            "sunshade_l": SW_SUNSHADE_L, "cap_scale_u": SW_CAP_SCALE_U,
            "cap_scale_l": SW_CAP_SCALE_L, "scale_leaf_only": SW_SCALE_LEAF_ONLY,
            "acclim_on": SW_ACCLIM, "acclim_per_pft": SW_ACCLIM_PER_PFT,
            "acclim_light": SW_ACCLIM_LIGHT, "frost_on": SW_FROST}
#: per-slot tables of ibis_step_ext, in order
NTABLE = 4

_D = C.c_double


class IbisArena(C.Structure):
    _fields_ = [("planes", C.c_void_p), ("pitch", C.c_int64)]


class IbisParams(C.Structure):
    _fields_ = (
        [(n, _D) for n in ("stef", "vonk", "grav", "tmelt", "hvap", "hfus", "hsub", "ch2o",
                           "cice", "cair", "cvap", "rair", "rvap", "cappa", "rhow", "epsilon")]
        + [(n, _D) for n in ("tau15", "kc15", "ko15", "cimax", "o2conc", "alpha3", "theta3",
                             "beta3", "alpha4", "theta4", "beta4")]
        + [("gamma", _D * 5), ("coefm", _D * 5), ("coefb", _D * 5), ("gsmin", _D * 5),
           ("stressfac", _D)]
        + [("cleaf", _D), ("cstem", _D), ("cgrass", _D), ("dleaf", _D * 2), ("dstem", _D * 2),
           ("alaiml", _D), ("alaimu", _D), ("chu", _D), ("chs", _D), ("chl", _D),
           ("soil_beta_power", _D)]
        + [(n, _D) for n in ("wliqumax", "wliqsmax", "wliqlmax", "wsnoumax", "wsnosmax",
                             "wsnolmax", "tdripu", "tdrips", "tdripl", "tblowu", "tblows",
                             "tblowl", "fmax")]
        + [("rhoveg", (_D * 2) * 2), ("tauveg", (_D * 2) * 2), ("oriev", _D * 2),
           ("orieh", _D * 2)]
        + [(n, _D) for n in ("wpudmax", "zwpmax", "bperm", "rhosoi_val", "csoi_mineral_val")]
        + [("h2o_consistent_weights", C.c_int32), ("pad0", C.c_int32)]
        + [("bed_hsoi", _D * L.NBEDROCK), ("bed_consoi", _D), ("bed_rhosoi", _D),
           ("bed_csoi", _D)]
        + [("d_crit", _D), ("vpd_ref", _D), ("d_crit_vpd", _D)]
        + [(n, _D) for n in ("z0sno", "rhos", "consno", "hsnotop", "hsnomin", "fimin", "fimax")]
        + [("flux_max", _D), ("t_max_dev", _D)]
        + [("hsoi", _D * (L.NSOILAY + 1)), ("froot", (_D * 2) * L.NSOILAY), ("co2conc", _D)]
    )


class IbisStomataExt(C.Structure):
    _fields_ = [("switches", C.c_int32), ("acclim_mask", C.c_int32), ("cap_kn", _D),
                ("acclim_s0", _D), ("acclim_smax", _D), ("frost_t0", _D), ("frost_t1", _D),
                ("type_tables", C.c_void_p), ("acclim_x", C.c_void_p), ("td", C.c_void_p),
                ("stressfac", C.c_void_p), ("cropfrac", C.c_void_p), ("crop_c4_share", _D),
                ("crop_vmax_scale_c3", _D), ("crop_vmax_scale_c4", _D)]


def switch_bits(sp) -> int:
    """IBIS_SW_* bits of a StomataParams-like object (absent attributes are off)."""
    return sum(bit for name, bit in SWITCHES.items() if bool(getattr(sp, name, False)))


def stomata_ext(sp, type_tables=0, acclim_x=0, td=0, stressfac=0, cropfrac=0,
                crop=None) -> IbisStomataExt:
    """IbisStomataExt from a StomataParams-like object, device pointers and, with a cropland
    share, the CropParams-like ``crop``."""
    x = IbisStomataExt()
    x.switches = switch_bits(sp)
    x.acclim_mask = sum(1 << int(j) for j in getattr(sp, "acclim_pfts", ()))
    x.cap_kn = float(getattr(sp, "cap_kn", 0.15))
    x.acclim_s0 = float(getattr(sp, "acclim_s0", -3.063))
    x.acclim_smax = float(getattr(sp, "acclim_smax", 17.72))
    x.frost_t0 = float(getattr(sp, "frost_t0", -5.0))
    x.frost_t1 = float(getattr(sp, "frost_t1", 5.0))
    x.type_tables, x.acclim_x, x.td, x.stressfac = type_tables, acclim_x, td, stressfac
    x.crop_c4_share, x.crop_vmax_scale_c3, x.crop_vmax_scale_c4 = -1.0, 1.0, 1.0
    if cropfrac:
        x.switches |= SW_CROP
        x.cropfrac = cropfrac
        if crop is not None:
            x.crop_c4_share = float(crop.c4_share)
            x.crop_vmax_scale_c3 = float(crop.vmax_scale_c3)
            x.crop_vmax_scale_c4 = float(crop.vmax_scale_c4)
    return x


def params_from(consts, sim, snow, hsoi, froot, co2conc, flux_max=2000.0, t_max_dev=40.0):
    """Fill IbisParams from reference-named parameter objects.

    ``consts``: PhysicalConstants; ``sim``: SimParams (stomata_params, canopy, interception,
    veg_rad, bedrock, fire, wpudmax, zwpmax, bperm, rhosoi_val, csoi_mineral_val,
    h2o_consistent_weights); ``snow``: the SnowState constants; ``hsoi`` (7,); ``froot`` (6, 2).
    """
    p = IbisParams()
    for n in ("stef", "vonk", "grav", "tmelt", "hvap", "hfus", "hsub", "ch2o", "cice", "cair",
              "cvap", "rair", "rvap", "cappa", "rhow", "epsilon"):
        setattr(p, n, float(getattr(consts, n)))
    sp = sim.stomata_params
    for n in ("tau15", "kc15", "ko15", "cimax", "o2conc", "alpha3", "theta3", "beta3", "alpha4",
              "theta4", "beta4"):
        setattr(p, n, float(getattr(sp, n)))
    for n in ("gamma", "coefm", "coefb", "gsmin"):
        v = np.asarray(getattr(sp, n), dtype=np.float64)
        getattr(p, n)[:] = list(v)
    sf = np.asarray(sp.stressfac, dtype=np.float64)
    # a per-cell stressfac travels as its own plane (IbisStomataExt.stressfac); the scalar
    # slot is then never read and holds NaN
    p.stressfac = float(sf) if sf.ndim == 0 else float("nan")
    cp = sim.canopy
    for n in ("cleaf", "cstem", "cgrass", "alaiml", "alaimu", "chu", "chs", "chl",
              "soil_beta_power"):
        setattr(p, n, float(getattr(cp, n)))
    p.dleaf[:] = [float(x) for x in cp.dleaf]
    p.dstem[:] = [float(x) for x in cp.dstem]
    ip = sim.interception
    for n in ("wliqumax", "wliqsmax", "wliqlmax", "wsnoumax", "wsnosmax", "wsnolmax", "tdripu",
              "tdrips", "tdripl", "tblowu", "tblows", "tblowl", "fmax"):
        setattr(p, n, float(getattr(ip, n)))
    vr = sim.veg_rad
    rho = np.asarray(vr.rhoveg, dtype=np.float64)
    tau = np.asarray(vr.tauveg, dtype=np.float64)
    for b in range(2):
        for s in range(2):
            p.rhoveg[b][s] = float(rho[b, s])
            p.tauveg[b][s] = float(tau[b, s])
    p.oriev[:] = [float(x) for x in np.asarray(vr.oriev)]
    p.orieh[:] = [float(x) for x in np.asarray(vr.orieh)]
    for n in ("wpudmax", "zwpmax", "bperm", "rhosoi_val", "csoi_mineral_val"):
        setattr(p, n, float(getattr(sim, n)))
    p.h2o_consistent_weights = 1 if bool(getattr(sim, "h2o_consistent_weights", False)) else 0
    bp = sim.bedrock
    if bp is not None:
        if len(bp.hsoi) != L.NBEDROCK:
            raise ValueError(f"bedrock must have {L.NBEDROCK} layers, got {len(bp.hsoi)}")
        p.bed_hsoi[:] = [float(x) for x in bp.hsoi]
        p.bed_consoi, p.bed_rhosoi, p.bed_csoi = float(bp.consoi), float(bp.rhosoi), float(bp.csoi)
    fp = sim.fire
    p.d_crit, p.vpd_ref, p.d_crit_vpd = float(fp.d_crit), float(fp.vpd_ref), float(fp.d_crit_vpd)
    for n in ("z0sno", "rhos", "consno", "hsnotop", "hsnomin", "fimin", "fimax"):
        setattr(p, n, float(getattr(snow, n)))
    p.flux_max, p.t_max_dev = float(flux_max), float(t_max_dev)
    h = np.asarray(hsoi, dtype=np.float64)
    p.hsoi[:] = [float(x) for x in h[:L.NSOILAY + 1]]
    fr = np.asarray(froot, dtype=np.float64)
    for k in range(L.NSOILAY):
        p.froot[k][0], p.froot[k][1] = float(fr[k, 0]), float(fr[k, 1])
    p.co2conc = float(co2conc)
    return p


def params_bytes(p: "IbisParams"):
    """The IbisParams block as a CPU uint8 tensor (the ``params`` argument of torch.ops.ibis.step)."""
    import torch
    return torch.frombuffer(bytearray(bytes(p)), dtype=torch.uint8)


def pack(fields: dict, n: int, dtype=np.float32, pitch: int | None = None, fill=0.0):
    """Reference-shaped arrays -> plane-major arena (NPLANE, pitch).  Missing fields stay
    at ``fill``."""
    pitch = n if pitch is None else pitch
    a = np.full((L.NPLANE, pitch), fill, dtype=dtype)
    for f in L.FIELDS:
        if f.key not in fields:
            continue
        v = np.asarray(fields[f.key], dtype=np.float64)
        off = L.PLANE_OFFSET[f.key]
        if f.width == 1:
            a[off, :n] = v.reshape(n)
        else:
            a[off:off + f.width, :n] = v.reshape(n, f.width).T
    return a


def unpack(arena, n: int, keys=None):
    """Plane-major arena -> dict of reference-shaped float64 arrays."""
    out = {}
    a = np.asarray(arena)
    for f in L.FIELDS:
        if keys is not None and f.key not in keys:
            continue
        off = L.PLANE_OFFSET[f.key]
        if f.width == 1:
            out[f.key] = a[off, :n].astype(np.float64)
        else:
            out[f.key] = a[off:off + f.width, :n].T.astype(np.float64)
    return out


def pack_slots(slot_type, n: int, pitch: int | None = None):
    """(N, 12) int8 slot ids -> (12, pitch) planes."""
    pitch = n if pitch is None else pitch
    s = np.full((L.NPFT, pitch), -1, dtype=np.int8)
    s[:, :n] = np.asarray(slot_type, dtype=np.int8).T
    return s
