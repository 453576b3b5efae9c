"""Write ``site_transient_1996_step.npz``: the leaves of a real IBIS state that the hourly
step reads, taken from the reference's saved state
``verify_out_stage1/site_transient_1996.npz`` (171 site cells, float32, production run).

The saved file is a flattened pytree: ``leaf_i`` in the field order of the reference's
state dataclasses (``python-port/ibis/state/soil.py``, ``snow.py``, ``canopy.py``,
``atmosphere.py``, ``vegetation.py``), soil first.  The order below is that field order; it
is checked against the leaf shapes (and the leaf count of each sub-state against the
file's ``treedef_repr``) before anything is written.  Only ``numpy.load`` with
``allow_pickle=False`` touches the file.

Run from the repository root:  python tests/golden/make_site_fixture.py
"""
import os
import re

import numpy as np

SRC = "/root/reference/verify_out_stage1/site_transient_1996.npz"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "site_transient_1996_step.npz")

ORDER = {
    "soil": "tsoi wsoi wisoi tg wpud wipud consoi csoi rhosoi hflo qglif heatg hvasug z0soi "
            "stresstl stresstu stressl stressu upsoil upsoiu porosflo poros hydraul suction "
            "bex sfield swilt sand clay hsoi albsav albsan tsoi_deep",
    "snow": "fi tsno hsno ti tlsub heati hvasui z0sno rhos consno hsnotop hsnomin fimin fimax",
    "canopy": "tu ts tl t12 t34 q12 q34 wliqu wliqs wliql wsnou wsnos wsnol ci_upper gs_upper "
              "cs_upper an_upper ag_upper ci_lower gs_lower cs_lower an_lower ag_lower fwetu "
              "fwets fwetl rliqu rliqs rliql solu sols soll solg soli firu firs firl firg firi "
              "firb asurd asuri topparu topparl scalcoefu scalcoefl fsenu fsens fsenl fseng "
              "fseni fvapg fvapi raing snowg a10scalparamu a10daylightu a10scalparaml "
              "a10daylightl rhoa",
    "atm": "ta qa ua psurf coszen solad solai fira raina snowa td tmax tmin qd ud precip cloud "
           "rh xstore iwet co2conc o2conc",
    "vegetation": "cbiol cbior cbiow plai frac biomass exist tnpp tgpp lai sai ztop zbot fu fl "
                  "froot tempu templ dropu dropls dropl3 dropl4 agddu agddl tc tw gdd0 gdd5 "
                  "tcmin tnpptot tgpptot tco2mic tco2root tneetot sapfrac firefac nytimes "
                  "nfirestep sdry nfirestep_vpd sdry_vpd totlaiu totlail totbiou totbiol_agg "
                  "disturbf disturbo vegtype0 tnmin gdd0this gdd5this tcthis twthis specla "
                  "aleaf aroot awood",
}
#: step inputs kept, with the expected width per cell (0 = global leaf)
KEEP = {
    "soil": dict(tsoi=6, wsoi=6, wisoi=6, tg=1, wpud=1, wipud=1, csoi=6, rhosoi=6, z0soi=1,
                 poros=6, hydraul=6, suction=6, bex=6, sfield=6, swilt=6, sand=6, clay=6,
                 albsav=1, albsan=1, tsoi_deep=3, hsoi=0),
    "snow": dict(fi=1, tsno=3, hsno=3, ti=1, tlsub=1, z0sno=0, rhos=0, consno=0, hsnotop=0,
                 hsnomin=0, fimin=0, fimax=0),
    "canopy": dict(tu=1, ts=1, tl=1, t12=1, t34=1, q12=1, q34=1, wliqu=1, wliqs=1, wliql=1,
                   wsnou=1, wsnos=1, wsnol=1, ci_upper=2, gs_upper=2, cs_upper=2, ci_lower=3,
                   gs_lower=3, cs_lower=3, a10scalparamu=1, a10daylightu=1, a10scalparaml=1,
                   a10daylightl=1),
    "atm": dict(ta=1, qa=1, ua=1, psurf=1, coszen=1, solad=2, solai=2, fira=1, raina=1,
                snowa=1, co2conc=0),
    "vegetation": dict(frac=12, exist=12, lai=2, sai=2, ztop=2, zbot=2, fu=1, fl=1, froot=0,
                       firefac=1, nfirestep=1, sdry=1, nfirestep_vpd=1, sdry_vpd=1, nytimes=0),
}
NODE = {"soil": "SoilState", "snow": "SnowState", "canopy": "CanopyState",
        "atm": "AtmState", "vegetation": "VegetationState"}


def main():
    z = np.load(SRC, allow_pickle=False)
    tree = str(z["treedef_repr"])
    i = 0
    out = {}
    ncell = z["leaf_0"].shape[0]
    for group, names in ORDER.items():
        names = names.split()
        m = re.search(NODE[group] + r"\[\(Missing,\)\], \[([*, ]*)\]", tree)
        assert m and m.group(1).count("*") == len(names), (group, m and m.group(1).count("*"))
        for name in names:
            a = z[f"leaf_{i}"]
            i += 1
            w = KEEP[group].get(name)
            if w is None:
                continue
            if w == 0:
                out[f"glob.{group}.{name}"] = a
                continue
            assert a.shape[0] == ncell and (a.ndim == 1) == (w == 1), (group, name, a.shape)
            assert w == 1 or a.shape[1] == w, (group, name, a.shape)
            out[f"{group}.{name}"] = a
    assert sum(len(v) for v in KEEP.values()) == len(out)
    np.savez_compressed(DST, **out)
    print(f"{DST}: {ncell} cells, {len(out)} leaves")


if __name__ == "__main__":
    main()
