"""Prescribed (Conrath) dust: DUSTPROFILE's vertical profile (code/dustprofile.f) and
FILLTAUCUM's TAUREF on the GCM's sub-levels (code/filltaucum.f), on the device
(mgcm_rt_dust_tauref).

    nu = dust.conrath_nu(vpout, j, l_j, vary=True)      # per latitude row, as DUSTPROFILE
    prdst = dust.pressure_grid(psf, ptrop)               # PRDST(100)
    dust.tauref(prdst, rptau, ptrop, tautot, nu, p, sigma, nlay)   # device: TAUREF, TAUDST
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

NPDST = 100


def _f32(x) -> float:
    return float(np.float32(x))


def conrath_nu(vpout: float, j: int, l_j: int, vary: bool = True, conrnu: float = 0.03) -> float:
    """DUSTPROFILE's Conrath parameter of latitude row j (1-based; code/dustprofile.f:41-54):
    0.04 - (0.04 - conrtmp) * cos(lat)**0.75 with conrtmp from the season VPOUT (L_s), or
    CONRNU when VARY_CONR is off.  REAL(4) literals and the REAL(4) latitude arithmetic kept;
    X**0.75 is SQRT(X)*SQRT(SQRT(X)), as the reference's build (amdflang/LLVM) evaluates it,
    X**1.5 the C library's pow."""
    if not vary:
        return conrnu
    deltals = (vpout - _f32(240.0)) / _f32(180.0) * math.pi
    conrtmp = _f32(0.023) * abs(math.sin(deltals / 2.0)) ** 1.5 + _f32(0.007)
    lat = np.float32(-90.0) + np.float32(180.0) / np.float32(l_j) * np.float32(j)   # REAL(4)
    tetalat = float(lat / np.float32(180.0)) * math.pi
    c = math.cos(tetalat)
    return _f32(0.04) - (_f32(0.04) - conrtmp) * (math.sqrt(c) * math.sqrt(math.sqrt(c)))


def pressure_grid(psf: float, ptrop: float, npdst: int = NPDST) -> np.ndarray:
    """PRDST (code/dustprofile.f:30-37): 1/(FLOAT(NPDST)-1) in REAL(4)."""
    refpr = (5.0 * psf / ptrop) ** float(np.float32(1.0) / (np.float32(npdst) - np.float32(1.0)))
    prdst = np.empty(npdst)
    prdst[0] = ptrop
    for n in range(1, npdst):
        prdst[n] = refpr * prdst[n - 1]
    return prdst


class DustColumns(C.Structure):
    _fields_ = [("ncol", C.c_int), ("nlay", C.c_int)] + [(n, C.c_void_p) for n in (
        "tautot", "conr", "p", "sigma", "prdst")] + [("npdst", C.c_int), ("rptau", C.c_double), ("ptrop", C.c_double)]


def tauref(prdst, rptau: float, ptrop: float, tautot, conr, p, sigma, nlay: int) -> dict:
    """mgcm_rt_dust_tauref_host on host arrays (the library copies in and out and
    synchronises): TAUREF [ncol][nlev+1] and TAUDST [ncol][npdst]."""
    from . import _check, load_library
    lib = load_library()
    lib.mgcm_rt_dust_tauref_host.argtypes = [C.POINTER(DustColumns), C.c_void_p, C.c_void_p]
    a = {k: np.ascontiguousarray(v, dtype=np.float64)
         for k, v in (("tautot", tautot), ("conr", conr), ("p", p), ("sigma", sigma), ("prdst", prdst))}
    ncol, npdst = a["tautot"].size, a["prdst"].size
    tr = np.empty((ncol, 2 * nlay + 4))
    td = np.empty((ncol, npdst))
    s = DustColumns(ncol, nlay, *[a[k].ctypes.data for k in ("tautot", "conr", "p", "sigma", "prdst")], npdst,
                    float(rptau), float(ptrop))
    _check(lib.mgcm_rt_dust_tauref_host(C.byref(s), tr.ctypes.data, td.ctypes.data), "mgcm_rt_dust_tauref_host")
    return {"tauref": tr, "taudst": td}


def tauref_device(d: dict, rptau: float, ptrop: float, nlay: int, tr, td=None, stream: int = 0) -> None:
    """mgcm_rt_dust_tauref on device tensors d[tautot, conr, p, sigma, prdst] into
    tr [ncol][nlev+1] (and td [ncol][npdst]); asynchronous on stream."""
    from . import _check, load_library
    lib = load_library()
    lib.mgcm_rt_dust_tauref.argtypes = [C.POINTER(DustColumns), C.c_void_p, C.c_void_p, C.c_void_p]
    s = DustColumns(int(d["tautot"].numel()), nlay, *[d[k].data_ptr() for k in ("tautot", "conr", "p", "sigma",
                                                                                 "prdst")],
                    int(d["prdst"].numel()), float(rptau), float(ptrop))
    _check(lib.mgcm_rt_dust_tauref(C.byref(s), C.c_void_p(tr.data_ptr()),
                                   C.c_void_p(None if td is None else td.data_ptr()), C.c_void_p(stream)),
           "mgcm_rt_dust_tauref")
