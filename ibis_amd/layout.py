"""Field registry of the hourly step's device state: the one source of truth for the SoA layout.

The reference threads an Equinox pytree ``IBISState`` through ``timestep_body``
(``python-port/ibis/state/model_state.py:28-63``).  Each leaf the step touches is listed
here with its reference group, name and trailing width ``k`` (the reference shape is
``(npoi,)`` for ``k == 1`` and ``(npoi, k)`` otherwise) and its role in the step:

* ``in``  -- read, never written (static soil texture, forcing, vegetation structure);
* ``io``  -- read and written (prognostic state);
* ``out`` -- written only (diagnostics the step recomputes every hour).

On the device every field is stored plane-major: ``k`` contiguous planes of ``pitch``
elements, plane ``j`` holding column ``j`` of the reference array for every cell.  A
wavefront of 64 consecutive cells therefore reads 64 consecutive words of each plane,
which is the coalesced pattern HBM3E wants.  The reference ``(npoi, k)`` array is the
transposed view ``planes.T`` of that storage, so the Python mirror hands out exactly the
reference shapes without copying.

Counting the planes reproduces the reference's I/O contract of the step: 179 words read
and 180 written per cell (the leaves of ``state/soil.py:15-67``, ``state/snow.py:13-43``,
``state/canopy.py:13-103``, ``state/atmosphere.py:13-49``, ``state/vegetation.py:13-97``,
``state/hydrology.py:13-34`` that ``loops.py:234-1284`` reads and repacks).  The ``diag``
group adds the two total-to-air fluxes after ``cascad2`` that the reference computes and
drops (``loops.py:616-627``).

``render_header()`` writes ``include/ibis_layout.h``; a test checks the committed header is
current.
"""
from __future__ import annotations

from dataclasses import dataclass

NSOILAY = 6
NSNOLAY = 3
NBEDROCK = 3
NPFT = 12
NBAND = 2


@dataclass(frozen=True)
class Field:
    group: str
    name: str
    width: int
    role: str  # 'in' | 'io' | 'out'

    @property
    def key(self) -> str:
        return f"{self.group}.{self.name}"

    @property
    def cname(self) -> str:
        return f"IBIS_F_{self.group.upper()}_{self.name.upper()}"


def _f(group, spec):
    out = []
    for item in spec:
        name, width, role = item
        out.append(Field(group, name, width, role))
    return out


# Order inside each group follows the reference dataclass field order.
FIELDS: tuple[Field, ...] = tuple(
    _f("soil", [
        ("tsoi", NSOILAY, "io"), ("wsoi", NSOILAY, "io"), ("wisoi", NSOILAY, "io"),
        ("tg", 1, "io"), ("wpud", 1, "io"), ("wipud", 1, "io"),
        ("consoi", NSOILAY, "out"), ("csoi", NSOILAY, "in"), ("rhosoi", NSOILAY, "in"),
        ("hflo", NSOILAY + 1, "out"), ("qglif", 4, "out"), ("heatg", 1, "out"),
        ("hvasug", 1, "out"), ("z0soi", 1, "in"), ("stresstl", 1, "out"), ("stresstu", 1, "out"),
        ("stressl", NSOILAY, "out"), ("stressu", NSOILAY, "out"), ("porosflo", NSOILAY, "out"),
        ("poros", NSOILAY, "in"), ("hydraul", NSOILAY, "in"), ("suction", NSOILAY, "in"),
        ("bex", NSOILAY, "in"), ("sfield", NSOILAY, "in"), ("swilt", NSOILAY, "in"),
        ("sand", NSOILAY, "in"), ("clay", NSOILAY, "in"),
        ("albsav", 1, "in"), ("albsan", 1, "in"), ("tsoi_deep", NBEDROCK, "io"),
    ])
    + _f("snow", [
        ("fi", 1, "io"), ("tsno", NSNOLAY, "io"), ("hsno", NSNOLAY, "io"), ("ti", 1, "io"),
        ("tlsub", 1, "io"), ("heati", 1, "out"), ("hvasui", 1, "out"),
    ])
    + _f("canopy", [
        ("tu", 1, "io"), ("ts", 1, "io"), ("tl", 1, "io"), ("t12", 1, "io"), ("t34", 1, "io"),
        ("q12", 1, "io"), ("q34", 1, "io"),
        ("wliqu", 1, "io"), ("wliqs", 1, "io"), ("wliql", 1, "io"),
        ("wsnou", 1, "io"), ("wsnos", 1, "io"), ("wsnol", 1, "io"),
        ("ci_upper", 2, "io"), ("gs_upper", 2, "io"), ("cs_upper", 2, "io"),
        ("an_upper", 2, "out"), ("ag_upper", 2, "out"),
        ("ci_lower", 3, "io"), ("gs_lower", 3, "io"), ("cs_lower", 3, "io"),
        ("an_lower", 3, "out"), ("ag_lower", 3, "out"),
        ("fwetu", 1, "out"), ("fwets", 1, "out"), ("fwetl", 1, "out"),
        ("rliqu", 1, "out"), ("rliqs", 1, "out"), ("rliql", 1, "out"),
        ("solu", 1, "out"), ("sols", 1, "out"), ("soll", 1, "out"), ("solg", 1, "out"),
        ("soli", 1, "out"),
        ("firu", 1, "out"), ("firs", 1, "out"), ("firl", 1, "out"), ("firg", 1, "out"),
        ("firi", 1, "out"), ("firb", 1, "out"),
        ("asurd", NBAND, "out"), ("asuri", NBAND, "out"),
        ("topparu", 1, "out"), ("topparl", 1, "out"),
        ("scalcoefu", 4, "out"), ("scalcoefl", 4, "out"),
        ("fsenu", 1, "out"), ("fsens", 1, "out"), ("fsenl", 1, "out"), ("fseng", 1, "out"),
        ("fseni", 1, "out"), ("fvapg", 1, "out"), ("fvapi", 1, "out"),
        ("raing", 1, "out"), ("snowg", 1, "out"),
        ("a10scalparamu", 1, "io"), ("a10daylightu", 1, "io"),
        ("a10scalparaml", 1, "io"), ("a10daylightl", 1, "io"),
        ("rhoa", 1, "out"),
    ])
    + _f("atm", [
        ("ta", 1, "in"), ("qa", 1, "in"), ("ua", 1, "in"), ("psurf", 1, "in"),
        ("coszen", 1, "in"), ("solad", NBAND, "in"), ("solai", NBAND, "in"),
        ("fira", 1, "in"), ("raina", 1, "in"), ("snowa", 1, "in"),
    ])
    + _f("vegetation", [
        ("frac", NPFT, "in"), ("exist", NPFT, "in"),
        ("lai", 2, "in"), ("sai", 2, "in"), ("ztop", 2, "in"), ("zbot", 2, "in"),
        ("fu", 1, "in"), ("fl", 1, "in"),
        ("firefac", 1, "io"), ("nfirestep", 1, "io"), ("sdry", 1, "io"),
        ("nfirestep_vpd", 1, "io"), ("sdry_vpd", 1, "io"),
    ])
    + _f("hydrology", [
        ("ginvap", 1, "out"), ("gsuvap", 1, "out"), ("gtrans", 1, "out"),
        ("gtransu", 1, "out"), ("gtransl", 1, "out"), ("grunof", 1, "out"),
        ("gdrain", 1, "out"), ("gadjust", 1, "out"), ("wtot", 1, "out"),
        ("raing_avg", 1, "out"), ("traing_avg", 1, "out"), ("fhtop_soil", 1, "out"),
        ("rainl", 1, "out"), ("snowl", 1, "out"), ("trainl", 1, "out"), ("tsnowl", 1, "out"),
        ("sflo_bot", 1, "out"), ("fhtop_snow", 1, "out"),
    ])
    + _f("diag", [
        ("fsena", 1, "out"), ("fvapa", 1, "out"),
    ])
)

FIELD_INDEX = {f.key: i for i, f in enumerate(FIELDS)}
NFIELD = len(FIELDS)

# Plane offsets of every field inside one packed state buffer (all fields back to back).
PLANE_OFFSET: dict[str, int] = {}
_off = 0
for _fld in FIELDS:
    PLANE_OFFSET[_fld.key] = _off
    _off += _fld.width
NPLANE = _off
del _off, _fld


def words_read() -> int:
    return sum(f.width for f in FIELDS if f.role in ("in", "io"))


def words_written(include_diag: bool = False) -> int:
    return sum(f.width for f in FIELDS if f.role in ("io", "out")
               and (include_diag or f.group != "diag"))


# Per-slot type ids: one int8 plane per slot, [NPFT][pitch].
SLOT_BYTES_PER_CELL = NPFT

# Bytes the step must move per cell-hour if every touched leaf is read once and written
# once: 179 fp32 words in, 180 out (the reference state only), plus 12 int8 slot ids.
ALGORITHMIC_BYTES_PER_CELL_HOUR = 4 * (words_read() + words_written()) + SLOT_BYTES_PER_CELL


def fields_of(group: str):
    return [f for f in FIELDS if f.group == group]


def render_header() -> str:
    lines = [
        "/* Generated by ibis_amd/layout.py (render_header); do not edit by hand.",
        " *",
        " * Plane-major layout of the hourly step's per-cell state: field F of width k occupies",
        " * planes [IBIS_PLANE_OFF(F), IBIS_PLANE_OFF(F)+k), each plane holding one word per cell.",
        " * Roles: IN = read only, IO = read and written, OUT = written only.",
        " */",
        "#ifndef IBIS_LAYOUT_H",
        "#define IBIS_LAYOUT_H",
        "",
        f"#define IBIS_NSOILAY {NSOILAY}",
        f"#define IBIS_NSNOLAY {NSNOLAY}",
        f"#define IBIS_NBEDROCK {NBEDROCK}",
        f"#define IBIS_NPFT {NPFT}",
        f"#define IBIS_NBAND {NBAND}",
        f"#define IBIS_NFIELD {NFIELD}",
        f"#define IBIS_NPLANE {NPLANE}",
        f"#define IBIS_WORDS_READ {words_read()}",
        f"#define IBIS_WORDS_WRITTEN {words_written()}",
        "",
        "enum IbisRole { IBIS_ROLE_IN = 0, IBIS_ROLE_IO = 1, IBIS_ROLE_OUT = 2 };",
        "",
        "enum IbisField {",
    ]
    for i, f in enumerate(FIELDS):
        lines.append(f"  {f.cname} = {i},")
    lines.append("};")
    lines.append("")
    lines.append("/* first plane of each field inside a packed arena */")
    lines.append("enum IbisPlane {")
    for f in FIELDS:
        lines.append(f"  {f.cname.replace('IBIS_F_', 'IBIS_P_')} = {PLANE_OFFSET[f.key]},")
    lines.append("};")
    lines.append("")
    lines.append("/* width, plane offset and role of each field, indexed by IbisField */")
    lines.append("#define IBIS_FIELD_TABLE \\")
    for f in FIELDS:
        role = {"in": "IBIS_ROLE_IN", "io": "IBIS_ROLE_IO", "out": "IBIS_ROLE_OUT"}[f.role]
        lines.append(f"  X({f.cname}, {f.width}, {PLANE_OFFSET[f.key]}, {role}) \\")
    lines.append("")
    lines.append("#endif /* IBIS_LAYOUT_H */")
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    import sys
    sys.stdout.write(render_header())
