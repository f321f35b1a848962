"""Runs the GCM builds of oracle/build_model.sh (test helper): a fresh run directory
with the model's data files, one run of the one-hour tutorial namelist, its
history file and radiation-call dump parsed."""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")
RUN = os.path.join(REF, "gcm_run")
NLAYRAD, NV, NG = 25, 7, 17


def built(*names) -> bool:
    return all(os.path.exists(os.path.join(REF, n)) for n in names) and os.path.exists(os.path.join(RUN, "mars_1h"))


def run_model(name: str, workdir: str, dump: bool = True, timeout: int = 300) -> dict:
    """Runs oracle/_ref/<name> in workdir; returns its history and dump records."""
    os.makedirs(workdir, exist_ok=True)
    link = os.path.join(workdir, "data")
    if not os.path.exists(link):
        os.symlink(os.path.join(RUN, "data"), link)
    env = dict(os.environ)
    dpath = os.path.join(workdir, "radiation.bin")
    if dump:
        env["MGCM_COMP3_DUMP"] = dpath
    print(f"running {name} in {workdir}", flush=True)
    with open(os.path.join(RUN, "mars_1h")) as fin, open(os.path.join(workdir, "m.out"), "w") as fout:
        r = subprocess.run([os.path.join(REF, name)], stdin=fin, stdout=fout, stderr=subprocess.STDOUT, cwd=workdir,
                           env=env, timeout=timeout)
    out = open(os.path.join(workdir, "m.out")).read()
    assert r.returncode == 0 and "HAS STOPPED" in out, out[-2000:]
    from mgcm_rt import history
    head, steps = history.read_history(os.path.join(workdir, "fort.11"))
    return {"head": head, "steps": steps, "dump": read_dump(dpath) if dump else None}


def read_dump(path: str) -> list:
    """Radiation-call records: (j, i, part, values)."""
    raw = open(path, "rb").read()
    recs, o = [], 0
    while o < len(raw):
        j, i, part, n = struct.unpack("<4i", raw[o:o + 16])
        recs.append((j, i, part, np.frombuffer(raw[o + 16:o + 16 + 8 * n], dtype="<f8").copy()))
        o += 16 + 8 * n
    return recs


def split_passes(recs: list) -> list:
    """The records of each COMP3 radiation pass (a pass restarts at its first key)."""
    passes, first = [], None
    for r in recs:
        key = r[:3]
        if first is None:
            first = key
        if key == first:
            passes.append([])
        passes[-1].append(r)
    return passes


def groups(part: int, v: np.ndarray) -> dict:
    """Named quantities of a record, and the net-flux differences QRAD is built from
    (SUNTOT/IRTOTAL(2L+1) = FMNET(L) - FMNET(L-1), (3) = FMNET(1) - NFLUXTOP;
    code/comp3.f:677-681,735-738)."""
    n = NLAYRAD
    fmnet, up, dn, top = v[0:n], v[n:2 * n], v[2 * n:3 * n], v[3 * n]
    d = {"fmnet": fmnet, "fluxup": up, "fluxdn": dn, "nfluxtop": np.array([top]),
         "netdiff": np.concatenate([[fmnet[0] - top], np.diff(fmnet)]),
         "dnflux_surface": np.array([dn[-1]])}  # DNVFLUX / DNIRFLUX (code/comp3.f:697,740)
    if part == 1:
        d["diffvt"] = np.array([v[3 * n + 1]])  # DNDIFFV
        d["cumtauv"] = v[3 * n + 2:3 * n + 2 + NV * NG]
        d["detau"] = v[3 * n + 2 + NV * NG:]
    return d


def history_diff(a: dict, b: dict) -> dict:
    """Per history output time and field: max |a-b| / max |a| over latitude rows
    J = 1..JM-1 (row JM is never computed; it holds uninitialised memory that
    differs between builds)."""
    jm = int(a["head"]["jm"])
    out = []
    for sa, sb in zip(a["steps"], b["steps"]):
        row = {}
        for k in sa:
            x, y = np.atleast_1d(np.asarray(sa[k], dtype=np.float64)), np.atleast_1d(np.asarray(sb[k], dtype=np.float64))
            if x.ndim >= 2 and x.shape[0] == jm:
                x, y = x[:jm - 1], y[:jm - 1]
            s = float(np.max(np.abs(x))) if x.size else 0.0
            row[k] = float(np.max(np.abs(x - y))) / s if s > 0 else float(np.max(np.abs(x - y)))
        out.append(row)
    return out
