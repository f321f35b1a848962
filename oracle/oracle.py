"""ctypes bindings for the CPU checkers (TEST INFRASTRUCTURE ONLY).

* ``Oracle``   -- oracle/libmgcm_oracle.so, the C FP64 restatement (OpenMP batch).
* ``Reference`` -- oracle/_ref/libmgcm_ref.so, the reference Fortran routines
  compiled from their own sources by oracle/build_ref.sh (only where that
  library was built).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this
module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libmgcm_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmgcm_ref.so")

NV, NI, NG = 7, 5, 17
DP = C.POINTER(C.c_double)


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(DP)


class _Tables(C.Structure):
    _fields_ = [("nlay", C.c_int)] + [(n, DP) for n in (
        "co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi", "fzerov", "fzeroi",
        "tauray", "dwni", "planckir",
        "qxvdst", "qsvdst", "gvdst", "qxvcld", "qsvcld", "gvcld",
        "qxidst", "qsidst", "gidst", "qxicld", "qsicld", "gicld",
        "qextrefdst", "qextrefcld")]


# per-column aerosol optics: column-dict key (the library's name) -> oracle field
AER_FIELDS = (("qxv_dst", "qxvdst"), ("qsv_dst", "qsvdst"), ("gv_dst", "gvdst"),
              ("qxv_cld", "qxvcld"), ("qsv_cld", "qsvcld"), ("gv_cld", "gvcld"),
              ("qxi_dst", "qxidst"), ("qsi_dst", "qsidst"), ("gi_dst", "gidst"),
              ("qxi_cld", "qxicld"), ("qsi_cld", "qsicld"), ("gi_cld", "gicld"),
              ("qextref_dst", "qextrefdst"), ("qextref_cld", "qextrefcld"))


class _BatchIn(C.Structure):
    _fields_ = [("ncol", C.c_int)] + [(n, DP) for n in (
        "plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz")] + [
        (f, DP) for _, f in AER_FIELDS]


class _BatchOut(C.Structure):
    _fields_ = [(n, DP) for n in (
        "fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni",
        "nfluxtopv", "nfluxtopi", "diffvt", "detau", "cumtauv", "hr")]


class _OpCount(C.Structure):
    _fields_ = [("flops", C.c_double), ("exps", C.c_double), ("sqrts", C.c_double), ("logs", C.c_double)]


OUT_KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni",
            "nfluxtopv", "nfluxtopi", "diffvt", "detau", "cumtauv", "hr")


def alloc_outputs(ncol: int, nlay: int, fill: float = 0.0) -> dict:
    nl = nlay + 1
    o = {}
    for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni"):
        o[k] = np.full((ncol, nl), fill)
    for k in ("nfluxtopv", "nfluxtopi", "diffvt"):
        o[k] = np.full(ncol, fill)
    o["detau"] = np.full((ncol, NG, NV), fill)    # Fortran (7,17) per column
    o["cumtauv"] = np.full((ncol, NG, NV), fill)
    o["hr"] = np.full((ncol, nlay + 1), fill)
    return o


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _AerTables(C.Structure):
    _fields_ = [("radb_rt", DP), ("q_v", DP), ("q_i", DP)]


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run oracle/build_oracle.sh)")
        self.lib = C.CDLL(path)
        self.lib.or_batch_run.argtypes = [C.POINTER(_Tables), DP, C.POINTER(_BatchIn), C.POINTER(_BatchOut), C.c_int]
        self.lib.or_batch_count.argtypes = [C.POINTER(_Tables), DP, C.POINTER(_BatchIn), C.POINTER(_OpCount)]
        self.lib.or_tpindex.argtypes = [C.c_double, C.c_double, C.c_double, DP, DP, DP, DP,
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), DP]
        self.lib.or_dtridgl.argtypes = [C.c_int, DP, DP, DP, DP, DP]
        self.lib.or_dsolver.argtypes = [C.c_int] + [DP] * 9 + [C.c_double] * 3 + [DP, DP]
        self.lib.or_lagrange.argtypes = [C.c_double, DP, DP, DP]
        self.lib.or_laginterp.argtypes = [DP] * 6
        self.lib.or_planck_table.argtypes = [DP, DP, DP]
        self.lib.or_setspv.argtypes = [DP] * 4
        self.lib.or_aer_grid.argtypes = [DP, DP]
        self.lib.or_dust_grid.argtypes = [C.c_double, C.c_double, C.c_int, DP]
        self.lib.or_dsolflux.argtypes = [DP, C.c_double, DP, DP, DP]
        self.lib.or_dsolflux.restype = C.c_double
        self.lib.or_dust_columns.argtypes = [C.c_int, C.c_int, C.c_int, DP, C.c_double, C.c_double] + [DP] * 7
        self.lib.or_opt_aer.argtypes = [C.c_int, C.POINTER(_AerTables), C.c_int, C.c_int] + [DP] * 11
        self.lib.or_libm.argtypes = [C.c_int, DP, DP, DP, DP]

    @staticmethod
    def _tables(t: dict, keep: list) -> _Tables:
        s = _Tables()
        s.nlay = int(t["nlay"])
        for name, _ in _Tables._fields_[1:]:
            a = _c(t[name])
            keep.append(a)
            setattr(s, name, _p(a))
        return s

    @staticmethod
    def _inputs(c: dict, keep: list) -> _BatchIn:
        s = _BatchIn()
        s.ncol = int(c["ncol"])
        for name in ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz"):
            a = _c(c[name])
            keep.append(a)
            setattr(s, name, _p(a))
        for key, field in AER_FIELDS:
            if key in c:
                a = _c(c[key])
                keep.append(a)
                setattr(s, field, _p(a))
        return s

    @staticmethod
    def _check_layers(t: dict, c: dict) -> None:
        if int(c["nlay"]) != int(t["nlay"]):
            raise ValueError(f"tables are for {int(t['nlay'])} layers, columns have {int(c['nlay'])}")

    def run(self, t: dict, c: dict, nthreads: int = 0, out: dict | None = None) -> dict:
        self._check_layers(t, c)
        keep: list = []
        ts = self._tables(t, keep)
        ins = self._inputs(c, keep)
        o = out if out is not None else alloc_outputs(int(c["ncol"]), int(t["nlay"]))
        os_ = _BatchOut(*[_p(o[k]) for k in OUT_KEYS])
        sol = _c(t["sol"])
        self.lib.or_batch_run(C.byref(ts), _p(sol), C.byref(ins), C.byref(os_), int(nthreads))
        return o

    def aer_grid(self) -> tuple:
        """INITCLD's radiative size grid: rad_rt(20), radb_rt(21)."""
        rad, radb = np.zeros(20), np.zeros(21)
        self.lib.or_aer_grid(_p(rad), _p(radb))
        return rad, radb

    def opt_aer(self, cloud: bool, radb: np.ndarray, q_v: np.ndarray, q_i: np.ndarray, qtr: np.ndarray,
                pl: np.ndarray) -> dict:
        """OPT_DST / OPT_CLD over columns: qtr [ncol][6][nlay], pl [ncol][nlev]; tables as
        Reference.initcld returns them (dust (3,nb,20), cloud (3,nb,20,15))."""
        qtr, pl = _c(qtr), _c(pl)
        ncol, nlay = qtr.shape[0], qtr.shape[2]
        ld = 2 * nlay + 4
        keep = [_c(radb), _c(q_v), _c(q_i)]
        tb = _AerTables(_p(keep[0]), _p(keep[1]), _p(keep[2]))
        o = {k: np.zeros((ncol, NV if k[-1] == "v" else NI, ld)) for k in ("qxv", "qsv", "gv", "qxi", "qsi", "gi")}
        o.update(qextref=np.zeros((ncol, ld)), tauref=np.zeros((ncol, ld)), tau2=np.zeros((ncol, 2)))
        self.lib.or_opt_aer(int(cloud), C.byref(tb), ncol, nlay, _p(qtr), _p(pl),
                            *[_p(o[k]) for k in ("qxv", "qsv", "gv", "qxi", "qsi", "gi", "qextref", "tauref", "tau2")])
        return o

    def dsolflux(self, t: dict, acosz: np.ndarray, detau: np.ndarray) -> np.ndarray:
        """DSOLFLUX per column (0 where ACOSZ < REAL(4) 1e-5, as COMP3 sets it)."""
        sol, gw, fz = _c(t["sol"]), _c(t["gwv"]), _c(t["fzerov"])
        out = np.zeros(len(acosz))
        for i, u in enumerate(acosz):
            if u >= float(np.float32(1e-5)):
                out[i] = self.lib.or_dsolflux(_p(sol), float(u), _p(gw), _p(fz), _p(_c(detau[i])))
        return out

    def dust_grid(self, psf: float, ptrop: float, npdst: int = 100) -> np.ndarray:
        prdst = np.zeros(npdst)
        self.lib.or_dust_grid(psf, ptrop, npdst, _p(prdst))
        return prdst

    def dust_columns(self, prdst, rptau, ptrop, tautot, conr, p, sigma, nlay: int) -> dict:
        """DUSTPROFILE's column profile and FILLTAUCUM: TAUDST, TAUREF, TAUCUM per column."""
        tautot, conr, p, sigma, prdst = _c(tautot), _c(conr), _c(p), _c(sigma), _c(prdst)
        ncol, nlev = tautot.size, 2 * nlay + 3
        o = {"taudst": np.zeros((ncol, prdst.size)), "tauref": np.zeros((ncol, nlev + 1)),
             "taucum": np.zeros((ncol, nlev))}
        self.lib.or_dust_columns(ncol, nlay, prdst.size, _p(prdst), rptau, ptrop, _p(tautot), _p(conr), _p(p),
                                 _p(sigma), _p(o["taudst"]), _p(o["tauref"]), _p(o["taucum"]))
        return o

    def count(self, t: dict, c: dict) -> dict:
        self._check_layers(t, c)
        keep: list = []
        ts = self._tables(t, keep)
        ins = self._inputs(c, keep)
        n = _OpCount()
        sol = _c(t["sol"])
        self.lib.or_batch_count(C.byref(ts), _p(sol), C.byref(ins), C.byref(n))
        return {"flops": n.flops, "exps": n.exps, "sqrts": n.sqrts, "logs": n.logs}

    def tpindex(self, pw, tw, q, pref, tref, wref):
        lc = np.zeros(4)
        mt, mp, nh = C.c_int(), C.c_int(), C.c_int()
        wr = np.zeros(1)
        self.lib.or_tpindex(pw, tw, q, _p(_c(pref)), _p(_c(tref)), _p(_c(wref)), _p(lc),
                            C.byref(mt), C.byref(mp), C.byref(nh), _p(wr))
        return lc, mt.value, mp.value, nh.value, wr[0]

    def dtridgl(self, af, bf, cf, df):
        n = len(af)
        pad = lambda a: _c(np.concatenate([[0.0], a]))  # noqa: E731  (1-based)
        xk = np.zeros(n + 1)
        self.lib.or_dtridgl(n, _p(pad(af)), _p(pad(bf)), _p(pad(cf)), _p(pad(df)), _p(xk))
        return xk[1:]

    def dsolver(self, gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4, btop, bsurf, rsf):
        nl = len(gama)
        pad = lambda a: _c(np.concatenate([[0.0], a]))  # noqa: E731
        xk1, xk2 = np.zeros(nl + 1), np.zeros(nl + 1)
        arrs = [pad(a) for a in (gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4)]
        self.lib.or_dsolver(nl, *[_p(a) for a in arrs], btop, bsurf, rsf, _p(xk1), _p(xk2))
        return xk1[1:], xk2[1:]

    def lagrange(self, x, xi, yi):
        ans = np.zeros(1)
        self.lib.or_lagrange(x, _p(_c(xi)), _p(_c(yi)), _p(ans))
        return ans[0]

    def laginterp(self, pgref11, co2v8, co2i8):
        pint = np.zeros(51)
        co2v = np.zeros((NG, NV, 10, 51, 7))
        co2i = np.zeros((NG, NI, 10, 51, 7))
        self.lib.or_laginterp(_p(_c(pgref11)), _p(_c(co2v8)), _p(_c(co2i8)), _p(pint), _p(co2v), _p(co2i))
        return pint, co2v, co2i

    def planck_table(self):
        p = np.zeros((8501, NI))
        dwni, wnoi = np.zeros(NI), np.zeros(NI)
        self.lib.or_planck_table(_p(p), _p(dwni), _p(wnoi))
        return p, dwni, wnoi

    def setspv(self):
        a = [np.zeros(NV) for _ in range(4)]
        self.lib.or_setspv(*[_p(x) for x in a])
        return tuple(a)

    def libm(self, x: np.ndarray) -> tuple:
        """The host C library's exp, log10 and log of x (logs 0 where x <= 0)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        e, l10, ln = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
        self.lib.or_libm(int(x.size), _p(x), _p(e), _p(l10), _p(ln))
        return e, l10, ln


class Reference:
    """The reference Fortran (amdflang build), one column per call; 24 layers."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run oracle/build_ref.sh)")
        self.lib = C.CDLL(path)
        n = C.c_int()
        self.lib.ref_nlay(C.byref(n))
        self.nlay = n.value
        self.lib.ref_column.argtypes = [DP] * 33 + [C.c_double] * 3 + [DP] * 11
        self.lib.ref_grid.argtypes = [C.c_int] + [DP] * 47
        self.lib.ref_set_planck.argtypes = [DP]
        self.lib.ref_get_planck.argtypes = [DP]
        self.lib.ref_setspv.argtypes = [DP] * 5
        self.lib.ref_setspi.argtypes = [DP] * 3
        self.lib.ref_setrad.argtypes = [DP] * 14
        self.lib.ref_dsolver.argtypes = [C.c_int] + [DP] * 9 + [C.c_double] * 3 + [DP, DP]
        self.lib.ref_tpindex.argtypes = [C.c_double] * 3 + [DP] * 4 + [C.POINTER(C.c_int)] * 3 + [DP]
        self.lib.ref_aer_tables.argtypes = [DP] * 7
        self.lib.ref_opt_aer.argtypes = [C.c_int] + [DP] * 11
        self.lib.ref_dust_setup.argtypes = [C.c_double] * 4 + [C.c_int, C.c_double, DP]
        self.lib.ref_filltaucum.argtypes = [C.c_int, C.c_int, C.c_double] + [DP] * 5
        self.lib.ref_dsolflux.argtypes = [DP, C.c_double, DP, DP, DP, DP]
        self.lib.ref_fillpt.argtypes = [DP, C.c_double, C.c_double, C.c_double] + [DP] * 5

    def run(self, t: dict, c: dict, cols=None) -> dict:
        assert int(t["nlay"]) == self.nlay, "reference build is compiled for %d layers" % self.nlay
        ncol = int(c["ncol"])
        cols = range(ncol) if cols is None else cols
        o = alloc_outputs(ncol, self.nlay)
        self.lib.ref_set_planck(_p(_c(t["planckir"])))
        tabs = [_c(t[k]) for k in ("co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi",
                                   "fzerov", "fzeroi", "tauray", "dwni", "sol",
                                   "qxvdst", "qsvdst", "gvdst", "qxvcld", "qsvcld", "gvcld",
                                   "qxidst", "qsidst", "gidst", "qxicld", "qsicld", "gicld",
                                   "qextrefdst", "qextrefcld")]
        tab_names = ("co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi", "fzerov", "fzeroi", "tauray",
                     "dwni", "sol", "qxvdst", "qsvdst", "gvdst", "qxvcld", "qsvcld", "gvcld",
                     "qxidst", "qsidst", "gidst", "qxicld", "qsicld", "gicld", "qextrefdst", "qextrefcld")
        per_col = {tab_names.index(f): key for key, f in AER_FIELDS if key in c}
        for i in cols:
            prof = [_c(c[k][i]) for k in ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld")]
            outs = [o[k][i] for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni")]
            sc = [o[k][i:i + 1] for k in ("nfluxtopv", "nfluxtopi", "diffvt")]
            # the column's own aerosol optics (OPT_DST/OPT_CLD outputs) where given
            tabs_i = [(_c(c[per_col[j]][i]) if j in per_col else a) for j, a in enumerate(tabs)]
            self.lib.ref_column(*[_p(a) for a in tabs_i], *[_p(a) for a in prof],
                                float(c["als"][i]), float(c["albi"][i]), float(c["acosz"][i]),
                                *[_p(a) for a in outs], *[_p(a) for a in sc],
                                _p(o["detau"][i]), _p(o["cumtauv"][i]))
        _heating(t, c, o)
        return o

    def run_grid(self, t: dict, c: dict) -> dict:
        """All columns in one call of the harness's column loop (no per-column Python)."""
        assert int(t["nlay"]) == self.nlay, "reference build is compiled for %d layers" % self.nlay
        ncol = int(c["ncol"])
        o = alloc_outputs(ncol, self.nlay)
        self.lib.ref_set_planck(_p(_c(t["planckir"])))
        tabs = [_c(t[k]) for k in ("co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi",
                                   "fzerov", "fzeroi", "tauray", "dwni", "sol",
                                   "qxvdst", "qsvdst", "gvdst", "qxvcld", "qsvcld", "gvcld",
                                   "qxidst", "qsidst", "gidst", "qxicld", "qsicld", "gicld",
                                   "qextrefdst", "qextrefcld")]
        assert not any(k in c for k, _ in AER_FIELDS), "run_grid takes the shared aerosol tables; use run"
        prof = [_c(c[k][:ncol]) for k in ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld",
                                          "als", "albi", "acosz")]
        outs = [o[k] for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni",
                               "nfluxtopv", "nfluxtopi", "diffvt", "detau", "cumtauv")]
        self.lib.ref_grid(ncol, *[_p(a) for a in tabs], *[_p(a) for a in prof], *[_p(a) for a in outs])
        _heating(t, c, o)
        return o

    def initcld(self, rundir: str) -> dict:
        """INITCLD: the aerosol size grid and the optical tables from rundir/data."""
        cwd = os.getcwd()
        os.chdir(rundir)
        try:
            self.lib.ref_initcld()
        finally:
            os.chdir(cwd)
        rad, radb = np.zeros(20), np.zeros(21)
        qv_d, qi_d = np.zeros((3, NV, 20)), np.zeros((3, NI, 20))
        qv_c, qi_c = np.zeros((3, NV, 20, 15)), np.zeros((3, NI, 20, 15))
        consts = np.zeros(6)
        self.lib.ref_aer_tables(_p(rad), _p(radb), _p(qv_d), _p(qi_d), _p(qv_c), _p(qi_c), _p(consts))
        return dict(rad_rt=rad, radb_rt=radb, qv_dst=qv_d, qi_dst=qi_d, qv_cld=qv_c, qi_cld=qi_c,
                    athird=consts[0], dev_dt=consts[1], dev_ice=consts[2], dpden_dt=consts[3], dpden_ice=consts[4])

    def opt_aer(self, cloud: bool, qtr: np.ndarray, pl: np.ndarray) -> dict:
        """OPT_DST (cloud False) or OPT_CLD on one column: qtr (NTRACE=6, nlay), pl (nlev)."""
        nlev = 2 * self.nlay + 3
        o = {k: np.zeros((NV if k[-1] == "v" else NI, nlev + 1)) for k in ("qxv", "qsv", "gv", "qxi", "qsi", "gi")}
        qextref, tauref, tau2 = np.zeros(nlev + 1), np.zeros(nlev + 1), np.zeros(2)
        self.lib.ref_opt_aer(int(cloud), _p(_c(qtr)), _p(_c(pl)), _p(o["qxv"]), _p(o["qxi"]), _p(o["qsv"]),
                             _p(o["qsi"]), _p(o["gv"]), _p(o["gi"]), _p(qextref), _p(tauref), _p(tau2))
        o.update(qextref=qextref, tauref=tauref, tau2=tau2)
        return o

    def fillpt(self, pl, ptrop: float, tg, tstrat, tl) -> dict:
        """FILLPT of each column: pl, tl [ncol][L_LEVELS], tg, tstrat [ncol]."""
        ncol, nlev = pl.shape
        assert nlev == 2 * self.nlay + 3
        out = {k: np.zeros((ncol, nlev)) for k in ("plev", "tlev", "pmid", "tmid")}
        for i in range(ncol):
            row = {k: np.zeros(nlev) for k in out}
            self.lib.ref_fillpt(_p(np.ascontiguousarray(pl[i])), ptrop, float(tg[i]), float(tstrat[i]),
                                _p(np.ascontiguousarray(tl[i])), _p(row["plev"]), _p(row["tlev"]),
                                _p(row["pmid"]), _p(row["tmid"]))
            for k in out:
                out[k][i] = row[k]
        return out

    def dsolflux(self, sol, ubar0: float, gweight, fzerov, detau_col) -> float:
        out = np.zeros(1)
        self.lib.ref_dsolflux(_p(_c(sol)), float(ubar0), _p(_c(gweight)), _p(_c(fzerov)), _p(_c(detau_col)), _p(out))
        return float(out[0])

    def dust(self, psf, ptrop, rptau, conrnu, vary, vpout, tautot, p, sigma, cols) -> dict:
        """DUSTPROFILE over the grid (tautot (L_J, L_I) C-shaped (L_I, L_J)), then FILLTAUCUM
        for each (j, i) in cols (1-based) with P(j, i) = p[(j, i)]."""
        self.lib.ref_dust_setup(psf, ptrop, rptau, conrnu, int(vary), vpout, _p(_c(tautot)))
        nlev = 2 * self.nlay + 3
        out = {"prdst": np.zeros(100), "taudst": [], "tauref": [], "taucum": []}
        for (j, i) in cols:
            td, tr, tc = np.zeros(100), np.zeros(nlev + 1), np.zeros(nlev)
            self.lib.ref_filltaucum(j, i, float(p[(j, i)]), _p(_c(sigma)), _p(out["prdst"]), _p(td), _p(tr), _p(tc))
            out["taudst"].append(td)
            out["tauref"].append(tr)
            out["taucum"].append(tc)
        return {k: np.array(v) for k, v in out.items()}

    def tables(self, rundir: str):
        """SETSPV/SETSPI/SETRAD (+LAGINTERP) as the reference computes them at start-up."""
        wnov, dwnv, wavev, solarf, tauray = (np.zeros(NV) for _ in range(5))
        self.lib.ref_setspv(_p(wnov), _p(dwnv), _p(wavev), _p(solarf), _p(tauray))
        wnoi, dwni, wavei = (np.zeros(NI) for _ in range(3))
        self.lib.ref_setspi(_p(wnoi), _p(dwni), _p(wavei))
        planck = np.zeros((8501, NI))
        self.lib.ref_get_planck(_p(planck))
        tg, pf = np.zeros(7), np.zeros(51)
        co2v, co2i = np.zeros((NG, NV, 10, 51, 7)), np.zeros((NG, NI, 10, 51, 7))
        misc = [np.zeros(NV) for _ in range(4)] + [np.zeros(NI) for _ in range(4)] + [np.zeros(NI), np.zeros(NV)]
        cwd = os.getcwd()
        os.chdir(rundir)
        try:
            self.lib.ref_setrad(_p(tg), _p(pf), _p(co2v), _p(co2i), *[_p(a) for a in misc])
        finally:
            os.chdir(cwd)
        return dict(wnov=wnov, dwnv=dwnv, solarf=solarf, tauray=tauray, wnoi=wnoi, dwni=dwni,
                    planckir=planck, tgasref=tg, pfgasref=pf, co2v=co2v, co2i=co2i,
                    fzeroi=misc[8], fzerov=misc[9])

    def dsolver(self, gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4, btop, bsurf, rsf):
        nl = len(gama)
        xk1, xk2 = np.zeros(nl), np.zeros(nl)
        arrs = [_c(a) for a in (gama, cp, cm, cpm1, cmm1, e1, e2, e3, e4)]
        self.lib.ref_dsolver(nl, *[_p(a) for a in arrs], btop, bsurf, rsf, _p(xk1), _p(xk2))
        return xk1, xk2

    def tpindex(self, pw, tw, q, pref, tref, wref):
        lc = np.zeros(4)
        mt, mp, nh = C.c_int(), C.c_int(), C.c_int()
        wr = np.zeros(1)
        self.lib.ref_tpindex(pw, tw, q, _p(_c(pref)), _p(_c(tref)), _p(_c(wref)), _p(lc),
                             C.byref(mt), C.byref(mp), C.byref(nh), _p(wr))
        return lc, mt.value, mp.value, nh.value, wr[0]


def _heating(t: dict, c: dict, o: dict) -> None:
    """Layer heating (K/sol) from net fluxes, same statement order as the oracle's."""
    nlay = int(t["nlay"])
    grav, cpair, scalep, sps = 3.72, 7.3594e2, 1.0e2, 88775.0
    acosz = np.asarray(c["acosz"])
    up = acosz >= np.float64(np.float32(1.0e-5))
    fv, fi = o["fmnetv"], o["fmneti"]
    sun = np.zeros_like(fv)
    sun[:, 0] = fv[:, 0] - o["nfluxtopv"]
    sun[:, 1:] = fv[:, 1:] - fv[:, :-1]
    sun[~up] = 0.0
    ir = np.empty_like(fi)
    ir[:, 0] = fi[:, 0] - o["nfluxtopi"]
    ir[:, 1:] = fi[:, 1:] - fi[:, :-1]
    plev = np.asarray(c["plev"])
    for l in range(nlay + 1):
        dp = (plev[:, 0] + plev[:, 1]) if l == 0 else (plev[:, 2 * l + 2] - plev[:, 2 * l])
        df = sun[:, l] + ir[:, l]
        o["hr"][:, l] = ((df * grav) / ((cpair * scalep) * dp)) * sps


def history_fields(c: dict, o: dict, p=None, sigma=None, ptrop: float = 0.0, surfalb=None) -> dict:
    """COMP3's history diagnostics of one grid from its radiation outputs.

    REAL*4 fields (code/comp3.f:654,703-706,745-747; REAL*4 declarations
    code/modules.f90:513,984-986): the TOA/surface flux components rounded to
    nearest, VIS ones zero where the sun is down (code/comp3.f:690-695), surfalb
    = ALS only where the sun is up (the previous value, `surfalb`, elsewhere).
    With p, sigma: hr_nlte, the heating rate in K/sol with COMP3's non-LTE factor
    on SUNTOT (code/comp3.f:765-771), PNLTE = SIGMA(K)*(P-PTROP)+PTROP, K = 2l+2.
    """
    up = np.asarray(c["acosz"]) >= np.float64(np.float32(1.0e-5))
    f32 = lambda a: np.asarray(a, dtype=np.float64).astype(np.float32)  # noqa: E731
    z = np.float32(0.0)
    h = {"fuptopv": np.where(up, f32(o["fluxupv"][:, 0]), z),
         "fdntopv": np.where(up, f32(o["fluxdnv"][:, 0]), z),
         "fupsurfv": np.where(up, f32(o["fluxupv"][:, -1]), z),
         "fdnsurfv": np.where(up, f32(o["fluxdnv"][:, -1]), z),
         "fuptopir": f32(o["fluxupi"][:, 0]),
         "fupsurfir": f32(o["fluxupi"][:, -1]),
         "fdnsurfir": f32(o["fluxdni"][:, -1])}
    prev = np.zeros(up.size, np.float32) if surfalb is None else np.asarray(surfalb, np.float32)
    h["surfalb"] = np.where(up, f32(c["als"]), prev)
    if p is not None:
        nlay = o["fmnetv"].shape[1] - 1
        grav, cpair, scalep, sps = 3.72, 7.3594e2, 1.0e2, 88775.0
        k22, one = float(np.float32(2.2e4)), float(np.float32(1.0))
        fv, fi = o["fmnetv"], o["fmneti"]
        sun = np.empty_like(fv)
        sun[:, 0] = fv[:, 0] - o["nfluxtopv"]
        sun[:, 1:] = fv[:, 1:] - fv[:, :-1]
        sun[~up] = 0.0
        ir = np.empty_like(fi)
        ir[:, 0] = fi[:, 0] - o["nfluxtopi"]
        ir[:, 1:] = fi[:, 1:] - fi[:, :-1]
        plev, p, sigma = np.asarray(c["plev"]), np.asarray(p, np.float64), np.asarray(sigma, np.float64)
        hr = np.empty_like(fv)
        for l in range(nlay + 1):
            pn = sigma[2 * l + 1] * (p - ptrop) + ptrop
            sunlte = ((sun[:, l] * k22) * pn) / (one + k22 * pn)
            dp = (plev[:, 0] + plev[:, 1]) if l == 0 else (plev[:, 2 * l + 2] - plev[:, 2 * l])
            hr[:, l] = (((sunlte + ir[:, l]) * grav) / ((cpair * scalep) * dp)) * sps
        h["hr_nlte"] = hr
    return h
