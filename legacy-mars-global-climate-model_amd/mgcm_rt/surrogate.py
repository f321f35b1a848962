"""Learned IR heating-rate surrogate (include/mgcm_sg.h) with exact fallback.

The network maps a 36-layer column (T, log10 p, log10 dust optical depth,
log10 H2O per layer; surface temperature, emissivity, surface pressure) to the IR heating
of its 36 GCM layers.  Its matrix products run in the library's BF16
matrix-core kernels; training uses the library's own backward kernels and
Adam.  Targets come from the plain FP64 path run IR-only on the device.

Targets are learned as asinh(h / S) (S = 10 K/sol), standardised per layer:
heating rates here span 1e-3..1e5 K/sol and the transform keeps both ends in
the loss.  A column is trusted only if (1) its features lie inside the
training range and (2) two independently trained networks agree within a
threshold; every other column is re-solved exactly (IR-only plain path), so
the returned heating is the surrogate's where trusted and the exact one
elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import SKIP_IR, SKIP_VIS, FP64, Radiation, load_library, _check  # noqa: F401

NLAY = 36
NIN, NOUT, NPARAM = 147, 36, 160 * 512 + 512 + 2 * (512 * 512 + 512) + 512 * 36 + 36
SCALE = 10.0  # K/sol in the asinh target transform
GRAV, CPAIR, SCALEP, SOL_S = 3.72, 735.94, 100.0, 88775.0
FP = C.POINTER(C.c_float)


def _bind(lib):
    if getattr(lib, "_sg_bound", False):
        return lib
    lib.mgcm_sg_create.argtypes = [C.c_int, C.c_ulonglong, C.POINTER(C.c_void_p)]
    lib.mgcm_sg_destroy.argtypes = [C.c_void_p]
    lib.mgcm_sg_get_params.argtypes = [C.c_void_p, FP]
    lib.mgcm_sg_set_params.argtypes = [C.c_void_p, FP]
    lib.mgcm_sg_forward.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mgcm_sg_train_step.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p,
                                       C.c_void_p]
    lib.mgcm_sg_gradients.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, FP]
    lib.mgcm_sg_features.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 10 + [C.c_void_p]
    lib.mgcm_sg_gemm_check.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mgcm_sg_check.argtypes = [C.c_int] + [C.c_void_p] * 7 + [C.c_double] + [C.c_void_p] * 4
    lib._sg_bound = True
    return lib


def raw_features(cols: dict) -> np.ndarray:
    """The features of mgcm_sg_features before standardisation (numpy, FP64)."""
    nlay = int(cols["nlay"])
    assert nlay == NLAY
    L = np.arange(1, NLAY + 1) + 1
    km = 2 * L - 1
    t = cols["tmid"][:, km]
    lp = np.log10(cols["pmid"][:, km])
    dust = np.log10(np.maximum(cols["tauref"][:, 2 * L] + cols["tauref"][:, 2 * L + 1], 1e-12))
    lq = np.log10(np.maximum(cols["qh2o"][:, km], 1e-30))
    nlev = 2 * nlay + 3
    s = np.stack([cols["tlev"][:, nlev - 1], 1.0 - cols["albi"], cols["plev"][:, nlev - 1]], axis=1)
    return np.concatenate([t, lp, dust, lq, s], axis=1)


def ir_heating(fmneti, plev):
    """IR heating of GCM layers 1..36 (K/sol) from FMNETI: IRTOTAL(2L+1) =
    FMNETI(L) - FMNETI(L-1) over the layer's pressure thickness (code/comp3.f:
    735-738,765-771).  Works on numpy arrays and torch tensors alike."""
    l = np.arange(1, NLAY + 1)
    dp = plev[:, 2 * l + 2] - plev[:, 2 * l]
    return (((fmneti[:, l] - fmneti[:, l - 1]) * GRAV) / ((CPAIR * SCALEP) * dp)) * SOL_S


class Net:
    """One MLP instance on the device (parameters, Adam state, activations)."""

    def __init__(self, max_batch: int, seed: int):
        self.lib = _bind(load_library())
        h = C.c_void_p()
        _check(self.lib.mgcm_sg_create(int(max_batch), C.c_ulonglong(seed), C.byref(h)), "mgcm_sg_create")
        self.h = h
        self.max_batch = max_batch

    def close(self):
        if getattr(self, "h", None):
            self.lib.mgcm_sg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(NPARAM, dtype=np.float32)
        _check(self.lib.mgcm_sg_get_params(self.h, p.ctypes.data_as(FP)), "mgcm_sg_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.lib.mgcm_sg_set_params(self.h, p.ctypes.data_as(FP)), "mgcm_sg_set_params")

    def forward(self, x, y, stream: int = 0) -> None:
        """x: device float32 [B][147], y: device float32 [B][36]."""
        _check(self.lib.mgcm_sg_forward(self.h, int(x.shape[0]), x.data_ptr(), y.data_ptr(), C.c_void_p(stream)),
               "mgcm_sg_forward")

    def train_step(self, x, t, lr: float, loss, stream: int = 0) -> None:
        _check(self.lib.mgcm_sg_train_step(self.h, int(x.shape[0]), x.data_ptr(), t.data_ptr(), C.c_float(lr),
                                           loss.data_ptr(), C.c_void_p(stream)), "mgcm_sg_train_step")

    def gradients(self, x, t) -> np.ndarray:
        g = np.zeros(NPARAM, dtype=np.float32)
        _check(self.lib.mgcm_sg_gradients(self.h, int(x.shape[0]), x.data_ptr(), t.data_ptr(),
                                          g.ctypes.data_as(FP)), "mgcm_sg_gradients")
        return g


def features(dc: dict, mean, inv_std, x, stream: int = 0) -> None:
    """Standardised features of device columns dc into x (device float32)."""
    lib = _bind(load_library())
    args = [dc[k].data_ptr() for k in ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "albi")]
    _check(lib.mgcm_sg_features(int(dc["ncol"]), int(dc["nlay"]), *args, mean.data_ptr(), inv_std.data_ptr(),
                                x.data_ptr(), C.c_void_p(stream)), "mgcm_sg_features")


def exact_ir_heating(rt: Radiation, dc: dict, stream: int = 0):
    """IR heating (K/sol, torch float64 [ncol][36]) from the plain path, IR part only."""
    import torch
    ncol, nlay = int(dc["ncol"]), int(dc["nlay"])
    dev = dc["plev"].device
    out = {k: torch.empty((ncol, nlay + 1), dtype=torch.float64, device=dev)
           for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni")}
    for k in ("nfluxtopv", "nfluxtopi", "diffvt"):
        out[k] = torch.empty(ncol, dtype=torch.float64, device=dev)
    out["detau"] = torch.zeros((ncol, 17, 7), dtype=torch.float64, device=dev)
    out["cumtauv"] = torch.zeros((ncol, 17, 7), dtype=torch.float64, device=dev)
    out["hr"] = None
    out["hr"] = torch.empty((ncol, nlay + 1), dtype=torch.float64, device=dev)
    rt.run_device(dc, out, FP64 | SKIP_VIS, stream)
    # with the VIS part skipped the kernel's heating is the IR part alone: the same
    # operations as ir_heating(fmneti, plev) for GCM layers 1..36
    return out["hr"][:, 1:]


class Surrogate:
    """Two trained networks, the feature/target statistics, and the check."""

    def __init__(self, stats: dict, nets: list, threshold: float):
        self.stats = stats
        self.nets = nets
        self.threshold = threshold

    def predict(self, dc: dict, stream: int = 0):
        """(heating K/sol [ncol][36] from net 0, disagreement [ncol], in-range [ncol]) -- torch, device."""
        import torch
        dev = dc["plev"].device
        ncol = int(dc["ncol"])
        x = torch.empty((ncol, NIN), dtype=torch.float32, device=dev)
        features(dc, self.stats["mean"], self.stats["inv_std"], x, stream)
        ys = []
        for net in self.nets:
            y = torch.empty((ncol, NOUT), dtype=torch.float32, device=dev)
            for b0 in range(0, ncol, net.max_batch):
                b1 = min(ncol, b0 + net.max_batch)
                net.forward(x[b0:b1], y[b0:b1], stream)
            ys.append(y)
        # de-standardisation, disagreement and range check in one device kernel
        h = torch.empty((ncol, NOUT), dtype=torch.float64, device=dev)
        dis = torch.empty(ncol, dtype=torch.float64, device=dev)
        inr = torch.empty(ncol, dtype=torch.int32, device=dev)
        lib = _bind(load_library())
        st = self.stats
        _check(lib.mgcm_sg_check(ncol, ys[0].data_ptr(), ys[1].data_ptr() if len(ys) > 1 else None, x.data_ptr(),
                                 st["tmean"].data_ptr(), st["tstd"].data_ptr(), st["xlo"].data_ptr(),
                                 st["xhi"].data_ptr(), SCALE, h.data_ptr(), dis.data_ptr(), inr.data_ptr(),
                                 C.c_void_p(stream)), "mgcm_sg_check")
        return h, dis, inr.bool()

    def heating(self, rt: Radiation, dc: dict, stream: int = 0):
        """IR heating with exact fallback: (heating [ncol][36], fallback mask [ncol])."""
        import torch
        h, dis, inr = self.predict(dc, stream)
        bad = (~inr) | (dis > self.threshold)
        idx = torch.nonzero(bad).flatten()
        if idx.numel():
            sub = {k: (v.index_select(0, idx).contiguous() if hasattr(v, "index_select") else v)
                   for k, v in dc.items()}
            sub["ncol"] = int(idx.numel())
            h = h.clone()
            h[idx] = exact_ir_heating(rt, sub, stream)
        return h, bad


    def total_heating(self, rt: Radiation, dc: dict, out: dict, stream: int = 0, combined_above: float = 0.5):
        """The learned form's radiation call: heating of GCM layers 1..36 (K/sol, VIS
        exact, IR from the network where the check trusts it, exact elsewhere).  out
        receives the plain path's outputs of the run made (mgcm_rt output dict).

        With few fallback columns, a VIS-only run of the grid and an IR-only run of the
        fallback columns (gathered) cost least.  With more than combined_above of the
        columns falling back, one combined plain run of the grid (VIS and IR together,
        one continuum-point launch, no gather) costs less than those two runs; its
        heating is then taken whole for the fallback columns, and for trusted columns
        its IR part (from FMNETI) is replaced by the network's.
        Returns (heating [ncol][36], fallback mask [ncol], combined: bool)."""
        import torch
        ncol = int(dc["ncol"])
        h, dis, inr = self.predict(dc, stream)
        bad = (~inr) | (dis > self.threshold)
        nbad = int(bad.sum())
        if nbad > combined_above * ncol:
            rt.run_device(dc, out, FP64, stream)
            full = out["hr"][:, 1:]
            if nbad == ncol:
                return full, bad, True
            exact_ir = ir_heating(out["fmneti"], dc["plev"])
            return torch.where(bad[:, None], full, (full - exact_ir) + h), bad, True
        rt.run_device(dc, out, FP64 | SKIP_IR, stream)
        if nbad:
            idx = torch.nonzero(bad).flatten()
            sub = {k: (v.index_select(0, idx).contiguous() if hasattr(v, "index_select") else v)
                   for k, v in dc.items()}
            sub["ncol"] = nbad
            h = h.clone()
            h[idx] = exact_ir_heating(rt, sub, stream)
        return out["hr"][:, 1:] + h, bad, False


def calibrate(disagreement, err_max, bar: float) -> float:
    """Disagreement threshold of the check.  It is the smaller of (1) just below
    the smallest disagreement of any validation column whose max error exceeds
    the bar -- so every trusted validation column meets it -- and (2) half the
    bar, so that a column is trusted only when the two networks agree closer
    than the bar itself (validation alone does not bound fresh columns).
    -1 (trust nothing) when a failing column has zero disagreement."""
    bad = err_max > bar
    tau = 0.5 * bar
    if bool(bad.any()):
        d = float(disagreement[bad].min())
        tau = min(tau, d * (1.0 - 1e-6)) if d > 0 else -1.0
    return tau


def make_stats(raw: np.ndarray, targets: np.ndarray, device) -> dict:
    """Feature and target standardisation and the trusted feature range (the
    training range widened by 5% of its span plus half a standard deviation)."""
    import torch
    mean = raw.mean(axis=0)
    std = np.maximum(raw.std(axis=0), 1e-6 * np.maximum(np.abs(mean), 1.0))
    ty = np.arcsinh(targets / SCALE)
    tmean, tstd = ty.mean(axis=0), ty.std(axis=0)
    tstd = np.where(tstd > 0, tstd, 1.0)
    z = (raw - mean) / std
    span = z.max(axis=0) - z.min(axis=0)
    f = lambda a, dt=torch.float32: torch.tensor(np.asarray(a), dtype=dt, device=device)  # noqa: E731
    return {"mean": f(mean), "inv_std": f(1.0 / std), "tmean": f(tmean, torch.float64),
            "tstd": f(tstd, torch.float64), "xlo": f(z.min(axis=0) - 0.05 * span - 0.5),
            "xhi": f(z.max(axis=0) + 0.05 * span + 0.5),
            "np": {"mean": mean, "std": std, "tmean": tmean, "tstd": tstd}}


def train(net: Net, x, ty, epochs: int, batch: int, lr: float, seed: int, stream: int = 0,
          log=None) -> list:
    """Adam on the standardised targets ty (device float32 [N][36]); cosine-decayed
    learning rate; returns the mean loss per epoch."""
    import math

    import torch
    n = int(x.shape[0]) // batch * batch
    g = torch.Generator(device="cpu").manual_seed(seed)
    loss = torch.zeros(1, dtype=torch.float32, device=x.device)
    hist = []
    steps, total = 0, epochs * (n // batch)
    for ep in range(epochs):
        perm = torch.randperm(int(x.shape[0]), generator=g)[:n].to(x.device)
        acc = torch.zeros((), dtype=torch.float32, device=x.device)
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            xb, tb = x.index_select(0, idx), ty.index_select(0, idx)
            rate = lr * 0.5 * (1.0 + math.cos(math.pi * steps / max(1, total)))
            net.train_step(xb, tb, rate, loss, stream)
            acc += loss[0]
            steps += 1
        hist.append(float(acc) / (n // batch))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e}")
    return hist
