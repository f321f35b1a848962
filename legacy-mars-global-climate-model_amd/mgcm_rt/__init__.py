"""Python host binding of the MI355X column radiative transfer (include/mgcm_rt.h).

The library libmgcm_rt.so is built in-tree by ../build.py.  Loading fails loudly
when it is missing: there is no CPU path behind this module.

    from mgcm_rt import Radiation, synth
    tables, cols = synth.make_workload(seed=1, ncol=2160, nlay=24)
    rt = Radiation(tables)
    out = rt.run_host(cols)              # numpy in, numpy out (copies + sync)
    rt.run_device(dev_cols, dev_out)      # device pointers (torch tensors), async

Argument meaning follows the reference's OPTCV/SFLUXV/OPTCI/SFLUXI as called by
COMP3 (code/comp3.f:637-738): PLEV/TLEV/PMID/TMID/QH2O per level, TAUREF and
TAUREFCLD per sub-level, surface albedo ALS, IR albedo ALBI = 1 - emissivity,
cos(zenith) ACOSZ, SOL = SOLARF/RSDIST.  Outputs are FMNETV/FLUXUPV/FLUXDNV,
FMNETI/FLUXUPI/FLUXDNI per radiation layer, NFLUXTOPV/I, DIFFVT, DETAU and
CUMTAUV per (band, Gauss point), and the layer heating rates in K/sol.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import history, synth  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmgcm_rt.so")

NV, NI, NG = 7, 5, 17
FP64, MIXED, STUDY = 0, 1, 2
# precision-study component bits, OR-ed with STUDY (grids of 25..36 layers): k-tables
# in FP32, k-interpolation arithmetic in FP32, sub-level optical depths in FP32,
# layer transmissions in FP32, FP64 division by a Newton-refined reciprocal
P_TAB32, P_INTERP32, P_TAU32, P_TRANS32, P_FASTDIV = 0x100, 0x200, 0x400, 0x800, 0x1000
SKIP_VIS, SKIP_IR = 0x10, 0x20  # OR-ed into the precision argument: one part only
DP = C.POINTER(C.c_double)

STATUS = {0: "ok", -1: "invalid argument", -2: "layer count outside 1..99", -3: "HIP runtime error",
          -4: "out of device memory", -5: "temperature outside the Planck table or a NaN input",
          -6: "unknown precision", -7: "data file missing, unreadable or not in the expected record layout",
          -8: "grid larger than the no-allocation entry point's scratch"}


class RadiationError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {STATUS.get(status, status)} ({status})")
        self.status = status


class Tables(C.Structure):
    _fields_ = [(n, DP) for n in (
        "co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi", "fzerov", "fzeroi",
        "tauray", "dwni", "planckir",
        "qxv_dst", "qsv_dst", "gv_dst", "qxi_dst", "qsi_dst", "gi_dst",
        "qxv_cld", "qsv_cld", "gv_cld", "qxi_cld", "qsi_cld", "gi_cld")] + [
        ("qextref_dst", C.c_double), ("qextref_cld", C.c_double)]


# per-level aerosol optics of each column (mgcm_rt_aerosol): optional column-dict keys,
# [ncol][nb][nlev+1] for the band quantities, [ncol][nlev+1] for the two references
AER_KEYS = ("qxv_dst", "qsv_dst", "gv_dst", "qxi_dst", "qsi_dst", "gi_dst",
            "qxv_cld", "qsv_cld", "gv_cld", "qxi_cld", "qsi_cld", "gi_cld", "qextref_dst", "qextref_cld")


class Aerosol(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in AER_KEYS]


class Columns(C.Structure):
    _fields_ = [("ncol", C.c_int), ("nlay", C.c_int)] + [(n, C.c_void_p) for n in (
        "plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz")] + [
        ("sol", C.c_double * NV), ("aerosol", C.POINTER(Aerosol))]


class Tracers(C.Structure):
    _fields_ = [("ncol", C.c_int), ("nlay", C.c_int), ("pl", C.c_void_p), ("qtrace", C.c_void_p)]


AER_OUT_KEYS = ("qxv", "qsv", "gv", "qxi", "qsi", "gi", "qextref", "tauref", "tau2")
AER_TABLE_DOUBLES = 11552


class AerosolOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in AER_OUT_KEYS]


HISTORY_KEYS = ("fuptopv", "fdntopv", "fupsurfv", "fdnsurfv", "fuptopir", "fupsurfir", "fdnsurfir", "surfalb")


class History(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in HISTORY_KEYS + ("hr_nlte", "p", "sigma")] + [("ptrop", C.c_double)]


class Fluxes(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni",
        "nfluxtopv", "nfluxtopi", "diffvt", "detau", "cumtauv", "hr", "directsol", "dnvflux")]


NCAND = 10  # MGCM_RT_NCAND


class BatchClass(C.Structure):
    _fields_ = [("max_layers", C.c_int), ("items", C.c_int), ("tile", C.c_int), ("waves_per_simd", C.c_int),
                ("blocks", C.c_int), ("tuned_items", C.c_int), ("cand_tile", C.c_int * NCAND),
                ("cand_waves", C.c_int * NCAND), ("cand_mult", C.c_int * NCAND), ("cand_ms", C.c_float * NCAND)]


IN_KEYS = ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz")
OUT_KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni",
            "nfluxtopv", "nfluxtopi", "diffvt", "detau", "cumtauv", "hr")

_lib = None


def _one_hip_runtime() -> None:
    """A process must map one HIP runtime.  PyTorch ships its own libamdhip64 and
    HSA runtime; loaded first, they satisfy the library's HIP dependency by
    soname.  Loaded after the library, they would be a second runtime next to
    /opt/rocm's and torch would find no device.  So torch, when installed, is
    imported before the library is loaded."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libmgcm_rt.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: build it with build.py (no CPU fallback exists)")
        _one_hip_runtime()
        lib = C.CDLL(path)
        lib.mgcm_rt_init.argtypes = [C.POINTER(Tables), C.POINTER(C.c_void_p)]
        lib.mgcm_rt_destroy.argtypes = [C.c_void_p]
        lib.mgcm_rt_run.argtypes = [C.c_void_p, C.POINTER(Columns), C.POINTER(Fluxes), C.c_int, C.c_void_p]
        lib.mgcm_rt_run_host.argtypes = [C.c_void_p, C.POINTER(Columns), C.POINTER(Fluxes), C.c_int]
        lib.mgcm_rt_run_noalloc.argtypes = [C.c_void_p, C.POINTER(Columns), C.POINTER(Fluxes), C.c_int,
                                            C.c_void_p]
        lib.mgcm_rt_run_batched.argtypes = [C.c_void_p, C.c_int, C.POINTER(Columns), C.POINTER(Fluxes),
                                            C.c_int, C.c_void_p]
        lib.mgcm_rt_check.argtypes = [C.c_void_p]
        lib.mgcm_rt_check_stream.argtypes = [C.c_void_p, C.c_void_p]
        lib.mgcm_rt_batched_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(BatchClass)]
        lib.mgcm_rt_set_tile.argtypes = [C.c_void_p, C.c_int]
        lib.mgcm_rt_tile_for.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.mgcm_rt_planck_table.argtypes = [DP, DP, DP]
        lib.mgcm_rt_laginterp.argtypes = [DP, DP, C.c_int, DP, DP]
        lib.mgcm_rt_read_ktable.argtypes = [C.c_char_p, C.c_int, DP, DP]
        lib.mgcm_rt_read_aerosol_tables.argtypes = [C.c_char_p, DP]
        lib.mgcm_rt_dsolflux.argtypes = [C.c_void_p, C.c_int, DP] + [C.c_void_p] * 6
        lib.mgcm_rt_dsolflux_host.argtypes = [C.c_void_p, C.c_int, DP, DP, DP, DP, DP, DP]
        lib.mgcm_rt_history_fields.argtypes = [C.POINTER(Columns), C.POINTER(Fluxes), C.POINTER(History),
                                               C.c_void_p]
        lib.mgcm_rt_set_aerosol_tables.argtypes = [C.c_void_p, DP]
        lib.mgcm_rt_opt_dst.argtypes = [C.c_void_p, C.POINTER(Tracers), C.POINTER(AerosolOut), C.c_void_p]
        lib.mgcm_rt_opt_cld.argtypes = [C.c_void_p, C.POINTER(Tracers), C.POINTER(AerosolOut), C.c_void_p]
        lib.mgcm_rt_opt_dst_host.argtypes = [C.c_void_p, C.POINTER(Tracers), C.POINTER(AerosolOut)]
        lib.mgcm_rt_opt_cld_host.argtypes = [C.c_void_p, C.POINTER(Tracers), C.POINTER(AerosolOut)]
        lib.mgcm_rt_libm_check.argtypes = [DP, C.c_int, DP, DP, DP, DP]
        lib.mgcm_rt_strerror.restype = C.c_char_p
        lib.mgcm_rt_strerror.argtypes = [C.c_int]
        _lib = lib
    return _lib


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(DP)


def _check(status: int, what: str) -> None:
    if status != 0:
        raise RadiationError(status, what)


def alloc_outputs(ncol: int, nlay: int) -> dict:
    nl = nlay + 1
    o = {k: np.zeros((ncol, nl)) for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni")}
    for k in ("nfluxtopv", "nfluxtopi", "diffvt"):
        o[k] = np.zeros(ncol)
    o["detau"] = np.zeros((ncol, NG, NV))
    o["cumtauv"] = np.zeros((ncol, NG, NV))
    o["hr"] = np.zeros((ncol, nlay + 1))
    return o


def column_struct(cols: dict, ptr=lambda a: a.ctypes.data, sol=None) -> Columns:
    """The mgcm_rt_columns of a column dict; its aerosol block (if the dict has any of
    AER_KEYS) is kept alive on the returned struct."""
    s = Columns()
    s.ncol = int(cols["ncol"])
    s.nlay = int(cols["nlay"])
    for k in IN_KEYS:
        setattr(s, k, ptr(cols[k]))
    for i in range(NV):
        s.sol[i] = float(sol[i])
    if any(k in cols for k in AER_KEYS):
        a = Aerosol()
        for k in AER_KEYS:
            if k in cols:
                setattr(a, k, ptr(cols[k]))
        s._aerosol = a
        s.aerosol = C.pointer(a)
    return s


SURFACE_KEYS = ("directsol", "dnvflux")  # optional outputs: DSOLFLUX and COMP3's DNVFLUX, [ncol]


def flux_struct(out: dict, ptr=lambda a: a.ctypes.data) -> Fluxes:
    s = Fluxes()
    for k in OUT_KEYS + SURFACE_KEYS:
        v = out.get(k)
        setattr(s, k, None if v is None else ptr(v))
    return s


class Radiation:
    """One radiation context (tables resident on the current HIP device)."""

    def __init__(self, tables: dict):
        self.lib = load_library()
        self.tables = tables
        self.sol = np.asarray(tables["sol"], dtype=np.float64)
        self._keep = {}
        t = Tables()
        for name in ("co2v", "co2i", "pfgasref", "tgasref", "wrefh2o", "gwv", "gwi", "fzerov", "fzeroi",
                     "tauray", "dwni", "planckir"):
            a = _f64(tables[name])
            self._keep[name] = a
            setattr(t, name, _dp(a))
        # per-band aerosol constants (the tables' per-level arrays are level-constant)
        for src, dst in (("qxvdst", "qxv_dst"), ("qsvdst", "qsv_dst"), ("gvdst", "gv_dst"),
                         ("qxidst", "qxi_dst"), ("qsidst", "qsi_dst"), ("gidst", "gi_dst"),
                         ("qxvcld", "qxv_cld"), ("qsvcld", "qsv_cld"), ("gvcld", "gv_cld"),
                         ("qxicld", "qxi_cld"), ("qsicld", "qsi_cld"), ("gicld", "gi_cld")):
            a = _f64(np.asarray(tables[src])[:, 0])
            self._keep[dst] = a
            setattr(t, dst, _dp(a))
        t.qextref_dst = float(np.asarray(tables["qextrefdst"])[0])
        t.qextref_cld = float(np.asarray(tables["qextrefcld"])[0])
        h = C.c_void_p()
        _check(self.lib.mgcm_rt_init(C.byref(t), C.byref(h)), "mgcm_rt_init")
        self.ctx = h

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.mgcm_rt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_host(self, cols: dict, precision: int = FP64, out: dict | None = None) -> dict:
        """Host arrays in, host arrays out (copies, runs, synchronises)."""
        c = {k: _f64(cols[k]) for k in IN_KEYS + AER_KEYS if k in cols}
        c["ncol"], c["nlay"] = int(cols["ncol"]), int(cols["nlay"])
        o = out if out is not None else alloc_outputs(c["ncol"], c["nlay"])
        cs = column_struct(c, sol=self.sol)
        fs = flux_struct(o)
        _check(self.lib.mgcm_rt_run_host(self.ctx, C.byref(cs), C.byref(fs), int(precision)), "mgcm_rt_run_host")
        return o

    def run_device(self, cols: dict, out: dict, precision: int = FP64, stream: int = 0) -> None:
        """Device tensors (torch, float64, contiguous) in and out; asynchronous."""
        ptr = lambda t: t.data_ptr()  # noqa: E731
        cs = column_struct(cols, ptr=ptr, sol=self.sol)
        fs = flux_struct(out, ptr=ptr)
        _check(self.lib.mgcm_rt_run(self.ctx, C.byref(cs), C.byref(fs), int(precision), C.c_void_p(stream)),
               "mgcm_rt_run")

    def run_noalloc(self, cols: dict, out: dict, precision: int = FP64, stream: int = 0) -> None:
        """mgcm_rt_run_noalloc on device tensors: no device allocation; grids beyond the
        static scratch raise RadiationError (MGCM_RT_ERR_SCRATCH)."""
        ptr = lambda t: t.data_ptr()  # noqa: E731
        cs = column_struct(cols, ptr=ptr, sol=self.sol)
        fs = flux_struct(out, ptr=ptr)
        _check(self.lib.mgcm_rt_run_noalloc(self.ctx, C.byref(cs), C.byref(fs), int(precision),
                                            C.c_void_p(stream)), "mgcm_rt_run_noalloc")

    def run_batched(self, grids: list, outs: list, precision: int = FP64, stream: int = 0) -> None:
        ptr = lambda t: t.data_ptr()  # noqa: E731
        n = len(grids)
        structs = [column_struct(g, ptr=ptr, sol=self.sol) for g in grids]  # keeps the aerosol blocks alive
        cs = (Columns * n)(*structs)
        fs = (Fluxes * n)(*[flux_struct(o, ptr=ptr) for o in outs])
        _check(self.lib.mgcm_rt_run_batched(self.ctx, n, cs, fs, int(precision), C.c_void_p(stream)),
               "mgcm_rt_run_batched")
        del structs

    def batched_tuning(self) -> list:
        """The batched tuner's choice per layer class of the last planned batch."""
        info = (BatchClass * 8)()
        n = self.lib.mgcm_rt_batched_tuning(self.ctx, 8, info)
        _check(min(n, 0), "mgcm_rt_batched_tuning")
        out = []
        for c in info[:n]:
            cands = [{"tile": c.cand_tile[j], "waves_per_simd": c.cand_waves[j], "mult": c.cand_mult[j],
                      "ms": float(c.cand_ms[j])} for j in range(NCAND) if c.cand_ms[j] > 0]
            out.append({"max_layers": c.max_layers, "items": c.items, "tile": c.tile,
                        "waves_per_simd": c.waves_per_simd, "blocks": c.blocks, "tuned_items": c.tuned_items,
                        "candidates": cands})
        return out

    def set_tile(self, band_slots: int) -> None:
        """Column kernel tile of single-grid calls: 1, 2 or 4 band slots per
        column, 12 = one band per wave, 0 = chosen per grid by the library
        (mgcm_rt_set_tile)."""
        _check(self.lib.mgcm_rt_set_tile(self.ctx, int(band_slots)), "mgcm_rt_set_tile")

    def tile_for(self, ncol: int, nlay: int) -> int:
        """The tile a single-grid call of ncol x nlay runs with."""
        t = self.lib.mgcm_rt_tile_for(self.ctx, int(ncol), int(nlay))
        _check(min(t, 0), "mgcm_rt_tile_for")
        return t

    def check(self) -> None:
        _check(self.lib.mgcm_rt_check(self.ctx), "mgcm_rt_check")

    def check_stream(self, stream: int = 0) -> None:
        _check(self.lib.mgcm_rt_check_stream(self.ctx, C.c_void_p(stream)), "mgcm_rt_check_stream")

    def dsolflux(self, cols: dict, out: dict, stream: int = 0) -> dict:
        """DSOLFLUX and COMP3's DNVFLUX from a run's outputs (device tensors): returns
        device tensors directsol[ncol] and dnvflux[ncol]."""
        import torch
        n = int(cols["ncol"])
        ds = torch.empty(n, dtype=torch.float64, device=out["detau"].device)
        dn = torch.empty_like(ds)
        sol = _f64(self.sol)
        _check(self.lib.mgcm_rt_dsolflux(self.ctx, n, _dp(sol), cols["acosz"].data_ptr(), out["detau"].data_ptr(),
                                         out["diffvt"].data_ptr(), ds.data_ptr(), dn.data_ptr(), C.c_void_p(stream)),
               "mgcm_rt_dsolflux")
        return {"directsol": ds, "dnvflux": dn}

    def dsolflux_host(self, acosz, detau, diffvt) -> dict:
        """mgcm_rt_dsolflux_host: DSOLFLUX and DNVFLUX from host arrays (numpy):
        acosz [ncol], detau [ncol][17][7], diffvt [ncol]."""
        acosz, detau, diffvt = _f64(acosz), _f64(detau), _f64(diffvt)
        n = acosz.shape[0]
        ds, dn = np.zeros(n), np.zeros(n)
        sol = _f64(self.sol)
        _check(self.lib.mgcm_rt_dsolflux_host(self.ctx, n, _dp(sol), _dp(acosz), _dp(detau), _dp(diffvt), _dp(ds),
                                              _dp(dn)), "mgcm_rt_dsolflux_host")
        return {"directsol": ds, "dnvflux": dn}

    def history_fields(self, cols: dict, out: dict, hist: dict | None = None, p=None, sigma=None,
                       ptrop: float = 0.0, stream: int = 0) -> dict:
        """COMP3's history diagnostics of a run (device tensors): the REAL*4 fields
        HISTORY_KEYS [ncol] and, given p [ncol] and sigma [nlev], hr_nlte [ncol][nlay+1]
        (K/sol with the non-LTE factor).  hist may hold preallocated tensors (its
        surfalb keeps its entries for sun-down columns); the filled dict is returned."""
        import torch
        n, nl = int(cols["ncol"]), int(cols["nlay"]) + 1
        dev = out["fluxupv"].device
        h = dict(hist or {})
        for k in HISTORY_KEYS:
            if h.get(k) is None:
                h[k] = torch.zeros(n, dtype=torch.float32, device=dev)
        if p is not None and h.get("hr_nlte") is None:
            h["hr_nlte"] = torch.empty((n, nl), dtype=torch.float64, device=dev)
        hs = History(*[h[k].data_ptr() for k in HISTORY_KEYS],
                     h["hr_nlte"].data_ptr() if p is not None else None,
                     p.data_ptr() if p is not None else None,
                     sigma.data_ptr() if sigma is not None else None, float(ptrop))
        cs = column_struct(cols, ptr=lambda t: t.data_ptr(), sol=self.sol)
        fs = flux_struct(out, ptr=lambda t: t.data_ptr())
        _check(self.lib.mgcm_rt_history_fields(C.byref(cs), C.byref(fs), C.byref(hs), C.c_void_p(stream)),
               "mgcm_rt_history_fields")
        return h

    def set_aerosol_tables(self, tables: np.ndarray) -> None:
        """Upload INITCLD's size grid and optical tables (read_aerosol_tables' block)."""
        t = _f64(tables)
        assert t.size == AER_TABLE_DOUBLES
        _check(self.lib.mgcm_rt_set_aerosol_tables(self.ctx, _dp(t)), "mgcm_rt_set_aerosol_tables")

    def opt_aerosol(self, cloud: bool, pl, qtrace, out: dict, stream: int = 0) -> None:
        """OPT_DST (cloud False) / OPT_CLD on device tensors: pl [ncol][nlev],
        qtrace [ncol][6][nlay]; out holds AER_OUT_KEYS tensors (tau2 optional)."""
        ncol, nlay = int(qtrace.shape[0]), int(qtrace.shape[2])
        tr = Tracers(ncol, nlay, pl.data_ptr(), qtrace.data_ptr())
        ao = AerosolOut(*[(out[k].data_ptr() if out.get(k) is not None else None) for k in AER_OUT_KEYS])
        fn = self.lib.mgcm_rt_opt_cld if cloud else self.lib.mgcm_rt_opt_dst
        _check(fn(self.ctx, C.byref(tr), C.byref(ao), C.c_void_p(stream)), "mgcm_rt_opt_cld" if cloud else "mgcm_rt_opt_dst")

    def opt_aerosol_host(self, cloud: bool, pl: np.ndarray, qtrace: np.ndarray) -> dict:
        """OPT_DST / OPT_CLD on host arrays through mgcm_rt_opt_dst_host / _cld_host
        (the library copies in and out and synchronises)."""
        pl, qtrace = _f64(pl), _f64(qtrace)
        ncol, nlay = qtrace.shape[0], qtrace.shape[2]
        ld = 2 * nlay + 4
        shapes = {"qxv": (ncol, NV, ld), "qsv": (ncol, NV, ld), "gv": (ncol, NV, ld), "qxi": (ncol, NI, ld),
                  "qsi": (ncol, NI, ld), "gi": (ncol, NI, ld), "qextref": (ncol, ld), "tauref": (ncol, ld),
                  "tau2": (ncol, 2)}
        out = {k: np.empty(v) for k, v in shapes.items()}
        tr = Tracers(ncol, nlay, pl.ctypes.data, qtrace.ctypes.data)
        ao = AerosolOut(*[out[k].ctypes.data for k in AER_OUT_KEYS])
        fn = self.lib.mgcm_rt_opt_cld_host if cloud else self.lib.mgcm_rt_opt_dst_host
        _check(fn(self.ctx, C.byref(tr), C.byref(ao)), "mgcm_rt_opt_cld_host" if cloud else "mgcm_rt_opt_dst_host")
        return out


def planck_table():
    """PLANCKIR(5,8501) as C (8501, 5), DWNI, WNOI -- built on the device."""
    lib = load_library()
    p, dwni, wnoi = np.zeros((8501, NI)), np.zeros(NI), np.zeros(NI)
    _check(lib.mgcm_rt_planck_table(_dp(p), _dp(dwni), _dp(wnoi)), "mgcm_rt_planck_table")
    return p, dwni, wnoi


def laginterp(pgref11, raw) -> tuple:
    """LAGINTERP on the device: raw (17, nb, 10, 11, 7) -> (pint[51], (17, nb, 10, 51, 7))."""
    lib = load_library()
    raw = _f64(raw)
    nb = raw.shape[1]
    out = np.zeros((NG, nb, 10, 51, 7))
    pint = np.zeros(51)
    _check(lib.mgcm_rt_laginterp(_dp(_f64(pgref11)), _dp(raw), nb, _dp(pint), _dp(out)), "mgcm_rt_laginterp")
    return pint, out


def read_aerosol_tables(datadir: str) -> np.ndarray:
    """INITCLD's size grid and the four aerosol optical-property files of datadir (host)."""
    lib = load_library()
    t = np.zeros(AER_TABLE_DOUBLES)
    _check(lib.mgcm_rt_read_aerosol_tables(os.fsencode(datadir), _dp(t)), "mgcm_rt_read_aerosol_tables")
    return t


def aerosol_table_views(t: np.ndarray) -> dict:
    """Named views of a read_aerosol_tables block (C shapes of the Fortran arrays)."""
    return {"radb_rt": t[0:21], "q_v_dst": t[32:452].reshape(3, NV, 20), "q_i_dst": t[452:752].reshape(3, NI, 20),
            "q_v_cld": t[752:7052].reshape(3, NV, 20, 15), "q_i_cld": t[7052:11552].reshape(3, NI, 20, 15)}


def read_ktable(path: str, nbands: int) -> tuple:
    """A CO2H2O_*_INTEL k-coefficient file (LAGINTERP's reads): raw (17, nb, 10, 11, 7), FZERO[nb]."""
    lib = load_library()
    raw, fz = np.zeros((NG, nbands, 10, 11, 7)), np.zeros(nbands)
    _check(lib.mgcm_rt_read_ktable(os.fsencode(path), int(nbands), _dp(raw), _dp(fz)), "mgcm_rt_read_ktable")
    return raw, fz


def libm_check(x) -> tuple:
    """Device exp/log10/sqrt/log of x (diagnostic entry of the library)."""
    lib = load_library()
    x = _f64(x)
    e, l, q, ln = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    _check(lib.mgcm_rt_libm_check(_dp(x), x.size, _dp(e), _dp(l), _dp(q), _dp(ln)), "mgcm_rt_libm_check")
    return e, l, q, ln


def to_device(d: dict, device="cuda:0") -> dict:
    """Copy a dict of numpy arrays to contiguous float64 torch tensors (plumbing only)."""
    import torch
    r = {}
    for k, v in d.items():
        if isinstance(v, np.ndarray):
            r[k] = torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(device)
        else:
            r[k] = v
    return r
