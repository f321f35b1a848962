"""Seeded synthetic radiation workloads (numpy PCG64; identical bytes on every host).

Builds the inputs of the column radiation call of the Legacy Mars GCM:
k-tables in the reference's layout CO2V(NT,NP,NH2O,band,g), per-band Gauss weights,
aerosol optical constants per sub-level, and per-column profiles PLEV/TLEV/PMID/TMID,
QH2O, TAUREF, TAUREFCLD, albedos and cos(zenith).

Generator decisions (the workload statement leaves them open):
  * k-table values: log10(Cmk*k) ~ U(-8, 2), i.e. k = 10**u / Cmk, so a sub-layer of
    dp mbar has gas optical depth dp*10**u (thin, thick and TLIMIT-skipped items all
    occur).  g = 17 entries are zero, as in the reference's data files.
  * p grid: 51 log10 points from -6 to 5 (mbar); T grid 50..350 K; H2O grid 10
    log-spaced vmr from 1e-7 to 1e-2.
  * Gauss weights: per band Dirichlet(1,...,1) over g = 1..16, g = 17 weight 0,
    FZERO = 0.  weights="reference" gives every band the reference's Gauss-split 0.95
    set (code/modules.f90:337-345).
  * Sigma levels: the reference's DSIG (code/modules.f90:731-737) for 24 layers; for
    other layer counts the 24-layer cumulative sigma is resampled linearly in layer
    index.  PTROP = 0.0008 mbar (run/mars).  PL(K) = PTROP + SIGMA(K)*ps
    (code/potemp1.f:95-105); PLEV/TLEV/PMID/TMID follow FILLPT (code/fillpt.f90).
  * T: layer mid-point temperature linear in log p from 140 K (top) to 250 K (bottom)
    plus N(0, 5 K); TSTRAT = 140 K + N(0, 5 K); ground TG = bottom layer + U(-5, 15) K.
  * Dust: Conrath profile with nu = 0.03, tau_vis ~ U(0.1, 1) at 6.1 mbar, built as
    DUSTPROFILE (code/dustprofile.f) and FILLTAUCUM (code/filltaucum.f:44-53) do.
    Dust Q/omega/g are the reference's constants (code/setrad.f90:73-154).
  * H2O: per column vmr log-U(1e-7, 1e-3) times a per-layer factor 10**U(-0.3, 0.3);
    QH2O(1..3) = 0 as the reference leaves them (zero-initialised locals).
  * Cloud ice (clouds=True): tau log-U(1e-3, 2) in 0-3 random layers, split over the
    layer's two sub-levels; ice Q/omega/g per band are fixed constants listed below.
  * RSDIST = 2.4277 (sun-Mars distance squared, documentation/verification.txt:15).
"""
from __future__ import annotations

import functools
import math

import numpy as np

NT, NP, NH, NG, NV, NI = 7, 51, 10, 17, 7, 5
NPLANCK = 8501
CMK = 3.51e22
PTROP = 0.0008
RSDIST = 2.4277
GRAV = 3.72
SCALEP = 100.0

# code/modules.f90:731-737
DSIG24 = np.array([
    0.0001237, 0.0002186, 0.0003877, 0.0006877, 0.0012199, 0.0021622, 0.0038572,
    0.0057287, 0.0094287, 0.0155443, 0.0256262, 0.0422609, 0.0696418, 0.1148966,
    0.1530633, 0.1572269, 0.1352191, 0.1212412, 0.0718975, 0.0435678, 0.014, 0.008,
    0.003, 0.001])

# code/modules.f90:337-345 (Gauss-split 0.95)
GWEIGHT_REF = np.array([
    4.8083554740e-02, 1.0563099137e-01, 1.4901065679e-01, 1.7227479710e-01,
    1.7227479710e-01, 1.4901065679e-01, 1.0563099137e-01, 4.8083554740e-02,
    2.5307134073e-03, 5.5595258613e-03, 7.8426661469e-03, 9.0670945845e-03,
    9.0670945845e-03, 7.8426661469e-03, 5.5595258613e-03, 2.5307134073e-03, 0.0])

# dust constants, code/setrad.f90:73-105
QEV1 = np.array([1.834, 2.296, 2.672, 2.829, 2.698, 2.452, 2.261])
QSV1 = np.array([1.695, 2.031, 2.583, 2.744, 2.626, 2.225, 1.525])
GV1 = np.array([0.551, 0.640, 0.661, 0.678, 0.690, 0.743, 0.868])
QEI1 = np.array([0.008, 0.262, 0.491, 1.017, 0.444])
QSI1 = np.array([0.001, 0.037, 0.122, 0.351, 0.336])
GI1 = np.array([0.004, 0.030, 0.095, 0.214, 0.316])
# water-ice cloud constants (generator decision; 2-4 um effective radius class)
QEV_CLD = np.array([2.20, 2.18, 2.15, 2.12, 2.10, 2.08, 2.06])
QSV_CLD = np.array([1.60, 2.02, 2.14, 2.11, 2.09, 2.07, 2.05])
GV_CLD = np.array([0.86, 0.83, 0.81, 0.80, 0.79, 0.78, 0.77])
QEI_CLD = np.array([0.35, 0.95, 0.60, 0.40, 1.10])
QSI_CLD = np.array([0.05, 0.45, 0.25, 0.15, 0.55])
GI_CLD = np.array([0.20, 0.45, 0.40, 0.35, 0.60])

L_NREFV = 6  # 1-based reference visible band (code/modules.f90:238)


def _f32(x) -> np.ndarray:
    """A REAL(4) literal promoted to REAL(8)."""
    return np.asarray(np.float32(x), dtype=np.float64)


def setspv():
    """Visible band centres, solar fluxes and Rayleigh tau (code/setspv.f90:67-105)."""
    bwnv = np.array([2222.22, 3087.37, 4030.63, 5370.57, 7651.11, 12500.00, 25000.00, 41666.67])
    solar = np.array([12.7, 24.2, 54.6, 145.9, 354.9, 657.5, 106.3], dtype=np.float32).astype(np.float64)
    wnov = 0.5 * (bwnv[1:] + bwnv[:-1])
    dwnv = bwnv[1:] - bwnv[:-1]
    wl = _f32(1.0e4) / wnov
    p0 = 9.423e6
    # WL**4 as the reference build evaluates it: ((WL*WL)*WL)*WL
    tauray = ((_f32(8.7) / GRAV) * (_f32(1.527) * (1.0 + _f32(0.013) / (wl * wl)) / (((wl * wl) * wl) * wl))
              * SCALEP / p0)
    return wnov, dwnv, solar, tauray


@functools.lru_cache(maxsize=1)
def _planck_table() -> tuple:
    """SETSPI's band edges and PLANCKIR(5,8501) as the reference computes them
    (code/setspi.f90:48-105): the C library's exp (math.exp) and Y**5 as left-to-right
    products, so every entry equals the reference build's bit for bit."""
    bwni = np.array([10.000, 166.667, 416.667, 833.333, 1250.000, 2222.222])
    x = (-0.981560634246719, -0.904117256370475, -0.769902674194305, -0.587317954286617,
         -0.367831498998180, -0.125233408511469, 0.125233408511469, 0.367831498998180,
         0.587317954286617, 0.769902674194305, 0.904117256370475, 0.981560634246719)
    w = (0.047175336386512, 0.106939325995318, 0.160078328543346, 0.203167426723066,
         0.233492536538355, 0.249147045813403, 0.249147045813403, 0.233492536538355,
         0.203167426723066, 0.160078328543346, 0.106939325995318, 0.047175336386512)
    c1, c2 = 3.741832e-16, 1.438786e-2
    wnoi = 0.5 * (bwni[1:] + bwni[:-1])
    dwni = bwni[1:] - bwni[:-1]
    planck = np.empty((NPLANCK, NI))
    exp = math.exp
    for nw in range(NI):
        a = 1.0e-2 / float(bwni[nw + 1])
        b = 1.0e-2 / float(bwni[nw])
        bpa, bma = (b + a) / 2.0, (b - a) / 2.0
        ys = [bma * xm + bpa for xm in x]
        y5s = [y * y * y * y * y for y in ys]
        den = math.pi * float(dwni[nw])
        col = planck[:, nw]
        for i in range(NPLANCK):
            t = (i + 500) / 1.0e1
            ans = 0.0
            for m in range(12):
                ans = ans + w[m] * c1 / (y5s[m] * (exp(c2 / (ys[m] * t)) - 1.0))
            col[i] = ans * bma / den
    return wnoi, dwni, planck


def setspi():
    """IR band centres/widths and the Planck table PLANCKIR(5,8501) (code/setspi.f90:48-105).

    Returns planckir in Fortran layout as a C array of shape (8501, 5)."""
    wnoi, dwni, planck = _planck_table()
    return wnoi.copy(), dwni.copy(), planck.copy()


def sigma_levels(nlay: int) -> np.ndarray:
    """SIGMA(1..nlev) 0-based (code/emiss.f:103-115); SIGMA(1), SIGMA(2) = 0."""
    sb24 = np.concatenate([[0.0], np.cumsum(DSIG24)])
    if nlay == 24:
        sb = sb24
    else:
        xi = np.arange(nlay + 1) * (24.0 / nlay)
        sb = np.interp(xi, np.arange(25), sb24)
    nlev = 2 * nlay + 3
    sig = np.zeros(nlev)
    # boundaries at K = 2l+3 (1-based) => 0-based index 2l+2, l = 0..nlay
    for l in range(nlay + 1):
        sig[2 * l + 2] = sb[l]
    for l in range(nlay):
        sig[2 * l + 3] = 0.5 * (sig[2 * l + 4] + sig[2 * l + 2])
    return sig


def make_tables(seed: int, nlay: int, weights: str = "dirichlet") -> dict:
    """k-tables, grids, Gauss weights and aerosol constants for one workload."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ks = {}
    for name, nb in (("co2v", NV), ("co2i", NI)):
        u = rng.uniform(-8.0, 2.0, size=(NG, nb, NH, NP, NT))
        k = np.power(10.0, u) / CMK
        k[NG - 1] = 0.0
        ks[name] = np.ascontiguousarray(k)
    gws = {}
    for name, nb in (("gwv", NV), ("gwi", NI)):
        gw = np.zeros((nb, NG))
        if weights == "reference":
            gw[:] = GWEIGHT_REF
        else:
            gw[:, :16] = rng.dirichlet(np.ones(16), size=nb)
        gws[name] = gw
    t = base_tables(nlay, gws["gwv"], gws["gwi"])
    t.update(ks)
    t["tgasref"] = np.arange(50.0, 351.0, 50.0)
    t["pfgasref"] = np.linspace(-6.0, 5.0, NP)
    t["wrefh2o"] = np.logspace(-7.0, -2.0, NH)
    t["fzerov"] = np.zeros(NV)
    t["fzeroi"] = np.zeros(NI)
    return t


def base_tables(nlay: int, gwv: np.ndarray, gwi: np.ndarray) -> dict:
    """The table entries that do not come from a k-distribution: Gauss weights,
    SETSPV's solar fluxes and Rayleigh tau, SETSPI's DWNI and Planck table, and
    the per-level aerosol constants (dust: code/setrad.f90:73-154)."""
    t = {"nlay": nlay, "gwv": np.asarray(gwv, dtype=np.float64), "gwi": np.asarray(gwi, dtype=np.float64)}
    _, _, solarf, tauray = setspv()
    t["solarf"] = solarf
    t["tauray"] = tauray
    _, dwni, planck = setspi()
    t["dwni"] = dwni
    t["planckir"] = planck
    nlev = 2 * nlay + 3
    nlp1 = nlev + 1

    def per_level(vals):
        return np.ascontiguousarray(np.repeat(np.asarray(vals, dtype=np.float64)[:, None], nlp1, axis=1))

    qsv = np.where(QSV1 >= QEV1, _f32(0.99999) * QEV1, QSV1)  # code/setrad.f90:137-139
    qsi = np.where(QSI1 >= QEI1, _f32(0.99999) * QEI1, QSI1)
    t["qxvdst"], t["qsvdst"], t["gvdst"] = per_level(QEV1), per_level(qsv), per_level(GV1)
    t["qxidst"], t["qsidst"], t["gidst"] = per_level(QEI1), per_level(qsi), per_level(GI1)
    t["qxvcld"], t["qsvcld"], t["gvcld"] = per_level(QEV_CLD), per_level(QSV_CLD), per_level(GV_CLD)
    t["qxicld"], t["qsicld"], t["gicld"] = per_level(QEI_CLD), per_level(QSI_CLD), per_level(GI_CLD)
    t["qextrefdst"] = np.full(nlp1, QEV1[L_NREFV - 1])  # code/ini_optdst.f
    t["qextrefcld"] = np.full(nlp1, QEV_CLD[L_NREFV - 1])
    t["sol"] = solarf / RSDIST  # code/comp3.f:300-304
    return t


def _conrath_taudst(tautot, conr=0.03, rptau=6.1, psf=7.01, npdst=100):
    """DUSTPROFILE's cumulative dust tau on its pressure grid (code/dustprofile.f)."""
    refpr = (5.0 * psf / PTROP) ** (1.0 / (npdst - 1.0))
    prdst = PTROP * refpr ** np.arange(npdst)
    s = 0.0
    n = 1
    while n < npdst:
        if prdst[n] < rptau:
            pave = 0.5 * (prdst[n] + prdst[n - 1])
            s += np.exp(conr * (1.0 - rptau / pave)) * (prdst[n] - prdst[n - 1])
        if prdst[n] >= rptau:
            break
        n += 1
    pave = 0.5 * (rptau + prdst[n - 1])
    s += np.exp(conr * (1.0 - rptau / pave)) * (rptau - prdst[n - 1])
    q0 = tautot / s
    q = np.empty(npdst - 1)
    for i in range(npdst - 1):
        if rptau > prdst[i + 1]:
            pave = 0.5 * (prdst[i + 1] + prdst[i])
            q[i] = q0 * np.exp(conr * (1.0 - rptau / pave))
        if prdst[i] <= rptau <= prdst[i + 1]:
            pave = 0.5 * (prdst[i] + rptau)
            q[i] = q0 * (np.exp(conr * (1.0 - rptau / pave)) * (rptau - prdst[i]) + (prdst[i + 1] - rptau)) / (
                prdst[i + 1] - prdst[i])
        if rptau < prdst[i]:
            q[i] = q0
    taudst = np.concatenate([[0.0], np.cumsum(q * np.diff(prdst))])
    return prdst, taudst


def _clib_log(x: np.ndarray) -> np.ndarray:
    """Natural log by the C library (math.log calls it for positive finite x), element by
    element: the reference's DLOG.  NumPy's log is its own vector code, whose last bit
    differs from the C library's for some arguments."""
    flat = np.ascontiguousarray(x, dtype=np.float64).ravel()
    return np.fromiter(map(math.log, flat.tolist()), dtype=np.float64, count=flat.size).reshape(np.shape(x))


def fillpt(pl: np.ndarray, ptrop: float, tg: np.ndarray, tstrat: np.ndarray, tl: np.ndarray) -> dict:
    """FILLPT (code/fillpt.f90:30-72) for ncol columns: the GCM's level pressures PL and
    temperatures TL ([ncol][L_LEVELS], TL at the even levels K = 4..L_LEVELS-1), ground
    temperature TG and stratosphere temperature TSTRAT ([ncol]) -> PLEV, TLEV, PMID, TMID
    on the radiation grid.  The odd levels' log-p interpolation is evaluated with the same
    operations in the same order, DLOG being the C library's log."""
    ncol, nlev = pl.shape
    nlay = (nlev - 3) // 2
    # level-major work arrays (row K-1 holds level K of every column)
    plev = np.array(np.asarray(pl, dtype=np.float64).T, copy=True)
    plev[0] = ptrop / 2.0
    plev[1] = ptrop / 2.0
    tlev = np.empty((nlev, ncol))
    tlev[0:3] = np.asarray(tstrat)[None, :]
    tlev[3:nlev - 1:2] = np.asarray(tl, dtype=np.float64).T[3:nlev - 1:2]  # K = 4, 6, ..., L_LEVELS-1
    if nlev - 1 > 5:  # K = 5, 7, ..., L_LEVELS-2 (each depends only on even levels)
        ks = np.arange(4, nlev - 2, 2)
        num = _clib_log(plev[ks] / plev[ks + 1])
        den = _clib_log(plev[ks - 1] / plev[ks + 1])
        tlev[ks] = tlev[ks + 1] + (tlev[ks - 1] - tlev[ks + 1]) * num / den
    tlev[nlev - 1] = tg
    tmid = np.empty_like(tlev)
    pmid = np.empty_like(plev)
    tmid[0] = tlev[1]
    tmid[1] = tlev[1]
    pmid[0] = plev[0]
    pmid[1] = plev[1]
    tmid[2:2 * nlay + 2:2] = tlev[2:2 * nlay + 2:2]        # TMID(2L+1) = TLEV(2L+1)
    tmid[3:2 * nlay + 3:2] = tlev[2:2 * nlay + 2:2]        # TMID(2L+2) = TLEV(2L+1)
    pmid[2:2 * nlay + 2:2] = plev[2:2 * nlay + 2:2]
    pmid[3:2 * nlay + 3:2] = plev[2:2 * nlay + 2:2]
    tmid[nlev - 1] = tlev[nlev - 1]
    pmid[nlev - 1] = plev[nlev - 1]
    plev, tlev, pmid, tmid = (np.ascontiguousarray(a.T) for a in (plev, tlev, pmid, tmid))
    return {"plev": plev, "tlev": tlev, "pmid": pmid, "tmid": tmid}


def gcm_levels(rng: np.random.Generator, ncol: int, nlay: int) -> dict:
    """The GCM-side fields FILLPT takes, drawn from rng: surface pressure ps (mbar), level
    pressures PL = PTROP + sigma*ps ([ncol][L_LEVELS]; PL(1..2) unused), layer
    temperatures TL at the even levels K = 2l+2 (a 140..250 K profile in log p plus
    N(0, 5 K)), TSTRAT and TG."""
    nlev = 2 * nlay + 3
    sig = sigma_levels(nlay)
    ps = rng.uniform(4.0, 10.0, ncol)                         # mbar
    pl = np.empty((ncol, nlev))
    pl[:, 1] = PTROP / 2.0
    pl[:, 2:nlev - 1] = PTROP + sig[None, 2:nlev - 1] * ps[:, None]
    pl[:, nlev - 1] = PTROP + ps
    pl[:, 0] = PTROP / 2.0
    kmid = np.arange(1, nlay + 1) * 2 + 1                     # 0-based index of K = 2l+2
    lp = np.log(pl[:, kmid])
    den = lp[:, -1:] - lp[:, :1]                              # 0 for a single layer
    frac = np.divide(lp - lp[:, :1], den, out=np.full_like(lp, 0.5), where=den > 0)
    tlay = 140.0 + 110.0 * frac + rng.normal(0.0, 5.0, (ncol, nlay))
    tstrat = 140.0 + rng.normal(0.0, 5.0, ncol)
    tg = tlay[:, -1] + rng.uniform(-5.0, 15.0, ncol)
    tl = np.zeros((ncol, nlev))
    tl[:, kmid] = tlay
    return {"ps": ps, "pl": pl, "tl": tl, "tstrat": tstrat, "tg": tg, "sigma": sig}


def make_columns(seed: int, ncol: int, nlay: int, clouds: bool = False) -> dict:
    """Per-column inputs, per-column contiguous (Fortran order within a column)."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    nlev = 2 * nlay + 3
    g = gcm_levels(rng, ncol, nlay)
    ps, sig = g["ps"], g["sigma"]
    f = fillpt(g["pl"], PTROP, g["tg"], g["tstrat"], g["tl"])
    plev, tlev, pmid, tmid = f["plev"], f["tlev"], f["pmid"], f["tmid"]
    # water vapour (code/comp3.f:576-589)
    q0 = 10.0 ** rng.uniform(-7.0, -3.0, ncol)
    qf = 10.0 ** rng.uniform(-0.3, 0.3, (ncol, nlay))
    qh2o = np.zeros((ncol, nlev))
    for l in range(nlay):
        qh2o[:, 2 * l + 3] = q0 * qf[:, l]                    # K = 2l+4
        qh2o[:, 2 * l + 4] = q0 * qf[:, l]                    # K = 2l+5
    # Conrath dust (code/filltaucum.f:44-53)
    tautot = rng.uniform(0.1, 1.0, ncol)
    # the Conrath profile is linear in tau_vis: one unit profile serves every column
    prdst, taudst1 = _conrath_taudst(1.0)
    pstar = ps[:, None] * sig[None, 3:nlev] + PTROP           # K = 4..L_LEVELS
    pstar = np.maximum(pstar, prdst[0])
    taucum = np.interp(pstar, prdst, taudst1) * tautot[:, None]
    tauref = np.zeros((ncol, nlev + 1))
    tauref[:, 3:nlev] = np.diff(np.concatenate([np.zeros((ncol, 1)), taucum], axis=1), axis=1)
    taurefcld = np.zeros((ncol, nlev + 1))
    if clouds:
        ncl = rng.integers(0, 4, ncol)
        tau_c = 10.0 ** rng.uniform(-3.0, np.log10(2.0), (ncol, 3))
        lay = rng.integers(1, nlay + 1, (ncol, 3))
        for i in range(ncol):
            for j in range(ncl[i]):
                l = lay[i, j]
                taurefcld[i, 2 * l + 1] += 0.5 * tau_c[i, j]   # K = 2l+2
                taurefcld[i, 2 * l + 2] += 0.5 * tau_c[i, j]   # K = 2l+3
    c = {
        "nlay": nlay,
        "ncol": ncol,
        "plev": np.ascontiguousarray(plev),
        "tlev": np.ascontiguousarray(tlev),
        "pmid": np.ascontiguousarray(pmid),
        "tmid": np.ascontiguousarray(tmid),
        "qh2o": np.ascontiguousarray(qh2o),
        "tauref": np.ascontiguousarray(tauref),
        "taurefcld": np.ascontiguousarray(taurefcld),
        "als": rng.uniform(0.1, 0.4, ncol),
        "albi": 1.0 - rng.uniform(0.9, 1.0, ncol),
        "acosz": rng.uniform(0.0, 1.0, ncol),
    }
    return c


def make_aerosol(seed: int, ncol: int, nlay: int, dust: bool = True, cloud: bool = True) -> dict:
    """Per-column, per-level aerosol optics in the shape OPT_DST/OPT_CLD hand to the
    radiation (code/comp3.f:595-631), as column-dict keys of mgcm_rt.AER_KEYS.

    Generator decision: each (column, band, level) value is the reference constant times
    U(0.6, 1.4) (g times U(0.8, 1.2), capped at 0.95); Qscat is capped at 0.99999 Qext as
    OPT_DST does (code/opt_dst.f:136,148); the references follow OPT_DST: QEXTREFDST(K) =
    QXVDST(K, L_NREFV) (code/opt_dst.f:151), and likewise for the cloud."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    ld = 2 * nlay + 4
    out = {}

    def group(qe, qsc, gg, nb, suffix):
        qx = qe[None, :, None] * rng.uniform(0.6, 1.4, (ncol, nb, ld))
        qs = np.minimum(qsc[None, :, None] * rng.uniform(0.6, 1.4, (ncol, nb, ld)), _f32(0.99999) * qx)
        g = np.minimum(gg[None, :, None] * rng.uniform(0.8, 1.2, (ncol, nb, ld)), 0.95)
        return {f"qx{suffix}": qx, f"qs{suffix}": qs, f"g{suffix}": g}

    if dust:
        out.update(group(QEV1, QSV1, GV1, NV, "v_dst"))
        out.update(group(QEI1, QSI1, GI1, NI, "i_dst"))
        out["qextref_dst"] = np.ascontiguousarray(out["qxv_dst"][:, L_NREFV - 1, :])
    if cloud:
        out.update(group(QEV_CLD, QSV_CLD, GV_CLD, NV, "v_cld"))
        out.update(group(QEI_CLD, QSI_CLD, GI_CLD, NI, "i_cld"))
        out["qextref_cld"] = np.ascontiguousarray(out["qxv_cld"][:, L_NREFV - 1, :])
    return {k: np.ascontiguousarray(v) for k, v in out.items()}


def make_tracers(seed: int, ncol: int, nlay: int) -> np.ndarray:
    """QTRACE(L_LAYERS, NTRACE) per column, [ncol][6][nlay]: dust mass and number,
    ice mass and number, dust-core mass, vapour (code/initcld.f:37-48).

    Generator decision: log-uniform mixing ratios spanning both sides of OPT_DST's and
    OPT_CLD's thresholds (mass 1e-8 / 1e-7, number 1), dust mass 1e-10..1e-4 kg/kg and
    number 1e-1..1e10 /kg, ice mass 1e-10..1e-4 and number 1e-1..1e8, core mass
    1e-12..1e-6, vapour 1e-7..1e-3; one layer in eight of each species is empty."""
    rng = np.random.Generator(np.random.PCG64(seed + 15485863))
    q = np.empty((ncol, 6, nlay))
    for m, (lo, hi) in enumerate(((-10, -4), (-1, 10), (-10, -4), (-1, 8), (-12, -6), (-7, -3))):
        q[:, m] = 10.0 ** rng.uniform(lo, hi, (ncol, nlay))
        q[:, m][rng.random((ncol, nlay)) < 0.125] = 0.0
    return q


def make_workload(seed: int, ncol: int, nlay: int, clouds: bool = False, weights: str = "dirichlet"):
    return make_tables(seed, nlay, weights), make_columns(seed, ncol, nlay, clouds)


def ragged_shapes(seed: int, ngrids: int = 512, cols=(2000, 20000), lays=(16, 64)) -> list:
    """(ncol, nlay) of each grid of a ragged batch: U{cols} columns, U{lays} layers."""
    rng = np.random.Generator(np.random.PCG64(seed + 104729))
    nc = rng.integers(cols[0], cols[1], size=ngrids, endpoint=True)
    nl = rng.integers(lays[0], lays[1], size=ngrids, endpoint=True)
    return [(int(a), int(b)) for a, b in zip(nc, nl)]


def ragged_pools(seed: int, shapes: list, clouds: bool = False) -> dict:
    """One pool of columns per layer count, as long as the largest grid using it.

    Grid i reads the first ncol_i columns of pool[nlay_i]: inputs are shared
    read-only between grids of equal layer count, which keeps a batch of
    millions of columns cheap to generate; every grid writes its own outputs.
    """
    need = {}
    for n, nl in shapes:
        need[nl] = max(need.get(nl, 0), n)
    return {nl: make_columns(seed * 1000 + nl, n, nl, clouds) for nl, n in sorted(need.items())}


CONFIGS = {
    # name: (ncol, nlay, clouds)
    "c1_cpu_2160x24": (2160, 24, False),
    "c2_gpu_64800x24": (64800, 24, False),
    "c3_mixed_64800x36": (64800, 36, True),
    # c4: ragged_shapes(seed) -- 512 grids, 2000..20000 columns, 16..64 layers
}
