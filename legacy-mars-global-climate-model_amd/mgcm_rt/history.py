"""The GCM history file, for the radiation's diagnostics.

A history file ("fort.11") is a Fortran sequential unformatted file: each record
is framed by its byte length as a 4-byte little-endian integer before and after
the payload, and every REAL is REAL*4.  MHISTH writes three header records once
(code/mhisth.f:124-128) and MHISTV 21 records per output time (code/mhistv.f:
182-204); the reference's analysis tools read them back with
tutorial/historymod.f90 (readheader :64-73, readvariables :74-102,
skiprecords :103-138).

The radiation supplies records 17-19 of a time step (fuptopv, fdntopv, fupsurfv,
fdnsurfv / fuptopir, fupsurfir, fdnsurfir / surfalb), which
mgcm_rt.Radiation.history_fields computes on the device, and the radiative part
of DHEAT (hr_nlte, K/sol).  The other fields belong to the rest of the GCM: the
writer takes whatever the caller has and writes zeros for what it leaves out, so
that a history file with the radiation's diagnostics can be opened by the
reference's tools.

Arrays are given in the history's index order, (L_J, L_I[, L][, N]) with J
(latitude) fastest, i.e. numpy arrays of that shape in Fortran order or any
array of that shape (they are written with order="F").  A grid's columns in
the order col = (I-1)*L_J + (J-1) become such an array with
``field.reshape((L_J, L_I), order="F")`` (``grid_field``).
"""
from __future__ import annotations

import numpy as np

VERSION = "V24-001"  # code/modules.f90:16-20 (version_h)

# header records (code/mhisth.f:124-128); shapes name the dimensions
HEADER_RECORDS = (
    (("runnum", "f4", ()), ("jm", "i4", ()), ("im", "i4", ()), ("layers", "i4", ()), ("nl", "i4", ()),
     ("ntrace", "i4", ()), ("version", "S7", ())),
    (("dsig", "f4", ("layers",)), ("dxyp", "f4", ("jm",)), ("grav", "f4", ()), ("rgas", "f4", ()),
     ("cp", "f4", ()), ("stbo", "f4", ()), ("xlhtc", "f4", ()), ("kapa", "f4", ()), ("cmk", "f4", ()),
     ("decmax", "f4", ()), ("eccn", "f4", ()), ("orbinc", "f4", ()), ("vinc", "f4", ()),
     ("sdepth", "f4", ("nl2",)), ("alicen", "f4", ()), ("alices", "f4", ()), ("egoco2n", "f4", ()),
     ("egoco2s", "f4", ()), ("npcwikg", "f4", ()), ("gidn", "f4", ()), ("gids", "f4", ())),
    (("topog", "f4", ("jm", "im")), ("alsp", "f4", ("jm", "im")), ("zin", "f4", ("jm", "im", "nl")),
     ("npcflag", "l4", ("jm", "im"))),
)

_JI = ("jm", "im")
_JIL = ("jm", "im", "layers")
# one output time (code/mhistv.f:182-204)
STEP_RECORDS = (
    tuple((k, "f4", ()) for k in ("tau", "vpout", "rsdist", "tofday", "psf", "ptrop", "tautot", "rptau", "sind",
                                  "gasp")),
    (("nc3", "i4", ()), ("ncycle", "i4", ())),
    (("p", "f4", _JI),), (("t", "f4", _JIL),), (("u", "f4", _JIL),), (("v", "f4", _JIL),),
    (("gt", "f4", _JI),), (("co2ice", "f4", _JI),), (("stressx", "f4", _JI),), (("stressy", "f4", _JI),),
    (("tstrat", "f4", _JI),), (("tausurf", "f4", _JI),), (("ssun", "f4", _JI),),
    (("qtrace", "f4", ("jm", "im", "layers", "ntrace")),), (("qcond", "f4", ("jm", "im", "ntrace")),),
    (("stemp", "f4", ("jm", "im", "nl")),),
    (("fuptopv", "f4", _JI), ("fdntopv", "f4", _JI), ("fupsurfv", "f4", _JI), ("fdnsurfv", "f4", _JI)),
    (("fuptopir", "f4", _JI), ("fupsurfir", "f4", _JI), ("fdnsurfir", "f4", _JI)),
    (("surfalb", "f4", _JI),),
    (("dheat", "f4", _JIL),),
    (("geop", "f4", _JIL),),
)

RADIATION_FIELDS = ("fuptopv", "fdntopv", "fupsurfv", "fdnsurfv", "fuptopir", "fupsurfir", "fdnsurfir", "surfalb")


def grid_field(field, jm: int, im: int) -> np.ndarray:
    """A per-column field [jm*im] (J fastest) as the history's (L_J, L_I) array."""
    return np.asarray(field).reshape((jm, im), order="F")


def _dims(d: dict) -> dict:
    return {"jm": int(d["jm"]), "im": int(d["im"]), "layers": int(d["layers"]), "nl": int(d["nl"]),
            "ntrace": int(d["ntrace"]), "nl2": 2 * int(d["nl"]) + 1}


def _item_bytes(name: str, typ: str, shape: tuple, value, dims: dict) -> bytes:
    shp = tuple(dims[s] for s in shape)
    if typ == "S7":
        v = (value if value is not None else VERSION)
        v = v.encode() if isinstance(v, str) else bytes(v)
        return v[:7].ljust(7)
    dt = {"f4": "<f4", "i4": "<i4", "l4": "<i4"}[typ]
    if value is None:
        a = np.zeros(shp, dtype=dt)
    else:
        a = np.asarray(value)
        if typ == "l4":
            a = a.astype(bool).astype(dt)  # LOGICAL(4): .TRUE. = 1
        else:
            a = a.astype(dt)  # REAL*8 -> REAL*4 assignment: round to nearest
        if a.shape != shp:
            if a.size == int(np.prod(shp)) and a.ndim <= 1:
                a = a.reshape(shp, order="F")
            else:
                raise ValueError(f"history field {name}: shape {a.shape}, expected {shp}")
    return a.tobytes(order="F")


def write_record(f, payload: bytes) -> None:
    """One sequential unformatted record: 4-byte LE length, payload, length."""
    n = np.array([len(payload)], dtype="<u4").tobytes()
    f.write(n + payload + n)


def read_record(f) -> bytes | None:
    """The payload of the next record, or None at the end of the file."""
    head = f.read(4)
    if len(head) == 0:
        return None
    if len(head) != 4:
        raise ValueError("history file: truncated record marker")
    n = int(np.frombuffer(head, "<u4")[0])
    payload = f.read(n)
    tail = f.read(4)
    if len(payload) != n or len(tail) != 4 or int(np.frombuffer(tail, "<u4")[0]) != n:
        raise ValueError("history file: truncated or unframed record")
    return payload


def _write_records(f, specs, values: dict, dims: dict) -> None:
    for rec in specs:
        write_record(f, b"".join(_item_bytes(n, t, s, values.get(n), dims) for n, t, s in rec))


def _parse_record(payload: bytes, rec, dims: dict) -> dict:
    out, off = {}, 0
    for name, typ, shape in rec:
        shp = tuple(dims[s] for s in shape)
        if typ == "S7":
            out[name] = payload[off:off + 7].decode(errors="replace")
            off += 7
            continue
        n = int(np.prod(shp))
        a = np.frombuffer(payload, dtype={"f4": "<f4", "i4": "<i4", "l4": "<i4"}[typ], count=n, offset=off)
        off += 4 * n
        a = a.reshape(shp, order="F") if shp else a[0]
        out[name] = (np.asarray(a) != 0) if typ == "l4" else (a.copy() if shp else a.item())
    if off > len(payload):
        raise ValueError("history file: record shorter than its fields")
    return out


def write_header(f, **values) -> None:
    """MHISTH's three records (code/mhisth.f:124-128).  Needs jm, im, layers, nl,
    ntrace; every other value defaults to zero (version to VERSION)."""
    _write_records(f, HEADER_RECORDS, values, _dims(values))


def write_step(f, dims: dict, **values) -> None:
    """MHISTV's 21 records of one output time (code/mhistv.f:182-204).  dims holds
    jm, im, layers, nl, ntrace; absent fields are written as zeros."""
    _write_records(f, STEP_RECORDS, values, _dims(dims))


def read_history(path: str) -> tuple:
    """(header, [step, ...]) as historymod's readheader/readvariables read them."""
    with open(path, "rb") as f:
        first = read_record(f)
        if first is None:
            raise ValueError("history file: empty")
        head = _parse_record(first, HEADER_RECORDS[0], {})
        dims = _dims(head)
        for rec in HEADER_RECORDS[1:]:
            head.update(_parse_record(read_record(f), rec, dims))
        steps = []
        while True:
            payload = read_record(f)
            if payload is None:
                break
            step = _parse_record(payload, STEP_RECORDS[0], dims)
            for rec in STEP_RECORDS[1:]:
                p = read_record(f)
                if p is None:
                    raise ValueError("history file: time step cut short")
                step.update(_parse_record(p, rec, dims))
            steps.append(step)
    return head, steps
