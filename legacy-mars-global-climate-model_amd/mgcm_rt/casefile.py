"""Case and result files of the Fortran host program fortran/mgcm_rt_driver.

A case holds the radiation tables and a batch of columns in the reference's
Fortran layouts, written as a native-endian stream (int32 header, FP64 arrays);
the result file holds the status and the outputs.  The order of the arrays is
the one documented in fortran/mgcm_rt_driver.f90.
"""
from __future__ import annotations

import numpy as np

NV, NI, NG = 7, 5, 17
TABLE_KEYS = ("co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi", "fzerov", "fzeroi", "tauray",
              "dwni", "planckir")
AEROSOL_KEYS = ("qxvdst", "qsvdst", "gvdst", "qxidst", "qsidst", "gidst",
                "qxvcld", "qsvcld", "gvcld", "qxicld", "qsicld", "gicld")
COLUMN_KEYS = ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz")
RESULT_KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
               "diffvt", "detau", "cumtauv", "hr")


def write_case(path: str, tables: dict, cols: dict, precision: int = 0) -> None:
    ncol, nlay = int(cols["ncol"]), int(cols["nlay"])
    with open(path, "wb") as f:
        np.array([ncol, nlay, precision], dtype=np.int32).tofile(f)
        for k in TABLE_KEYS:
            np.ascontiguousarray(tables[k], dtype=np.float64).tofile(f)
        for k in AEROSOL_KEYS:  # level-constant: one value per band
            np.ascontiguousarray(np.asarray(tables[k])[:, 0], dtype=np.float64).tofile(f)
        np.array([np.asarray(tables["qextrefdst"])[0], np.asarray(tables["qextrefcld"])[0]]).tofile(f)
        np.ascontiguousarray(tables["sol"], dtype=np.float64).tofile(f)
        for k in COLUMN_KEYS:
            np.ascontiguousarray(cols[k], dtype=np.float64).tofile(f)


def read_result(path: str, ncol: int, nlay: int) -> tuple:
    nl = nlay + 1
    shapes = {k: (ncol, nl) for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "hr")}
    shapes.update({k: (ncol,) for k in ("nfluxtopv", "nfluxtopi", "diffvt")})
    shapes.update({k: (ncol, NG, NV) for k in ("detau", "cumtauv")})
    with open(path, "rb") as f:
        status = int(np.fromfile(f, dtype=np.int32, count=1)[0])
        out = {}
        for k in RESULT_KEYS:
            n = int(np.prod(shapes[k]))
            out[k] = np.fromfile(f, dtype=np.float64, count=n).reshape(shapes[k])
    return status, out
