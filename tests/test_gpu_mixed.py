"""Mixed-precision path (k-tables stored in FP32, all arithmetic FP64) on the GPU.

Bar: the reference formulation is ill-conditioned in the k-coefficients on these
inputs (thin IR layers: B1 = dB/DTAU enters with TERM; thick layers clip at 35),
so rounding the k-tables to FP32 moves even an all-FP64 evaluation by up to ~1 K/sol.
That floor is measured here with the oracle (FP64 arithmetic, FP32-rounded
tables), and the mixed path must stay within twice it (plus 0.05 K/sol), with an
RMS error below 0.05 K/sol.  DETAU of sunlit columns agrees to FP32 precision.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _floor(oracle, t, c):
    a = oracle.run(t, c)
    t32 = dict(t)
    for k in ("co2v", "co2i"):
        t32[k] = t[k].astype(np.float32).astype(np.float64)
    b = oracle.run(t32, c)
    return a, b


@pytest.mark.parametrize("nlay,clouds,seed", [(36, True, 31), (24, False, 32)])
def test_mixed_within_conditioning_floor(oracle, nlay, clouds, seed):
    from mgcm_rt import MIXED, Radiation, synth
    t, c = synth.make_workload(seed, 1024, nlay, clouds=clouds)
    ref, r32 = _floor(oracle, t, c)
    got = Radiation(t).run_host(c, precision=MIXED)
    floor_hr = float(np.abs(r32["hr"] - ref["hr"]).max())
    err = np.abs(got["hr"] - ref["hr"])
    rms = float(np.sqrt((err ** 2).mean()))
    print(f"nlay {nlay}: max {err.max():.3e} K/sol (floor {floor_hr:.3e}), rms {rms:.3e} K/sol", flush=True)
    assert err.max() <= 2.0 * floor_hr + 0.05
    assert rms <= 0.05
    for k in ("fmnetv", "fmneti", "fluxupv", "fluxdnv", "fluxupi", "fluxdni"):
        floor = float(np.abs(r32[k] - ref[k]).max())
        assert float(np.abs(got[k] - ref[k]).max()) <= 2.0 * floor + 1e-3, k
    sun = c["acosz"] >= np.float32(1e-5)
    assert np.allclose(got["detau"][sun], ref["detau"][sun], rtol=1e-5, atol=1e-12)


def test_precision_study_modes():
    """The precision study's column kernels (MGCM_RT_STUDY | component bits): every
    built combination runs and moves the result (its component is active), FP64 with
    the Newton-refined reciprocal -- a few ulps per division, which the ill-conditioned
    solves amplify to ~1e-5 in some fluxes -- stays within the 0.05 K/sol bar of the
    FP64 heating, and the study is refused outside the 25..36-layer class and on the
    batched path."""
    import torch
    from mgcm_rt import (FP64, P_FASTDIV, P_INTERP32, P_TAB32, P_TAU32, P_TRANS32, STUDY, Radiation,
                         RadiationError, alloc_outputs, synth, to_device)
    t, c = synth.make_workload(35, 600, 36, clouds=True)
    rt = Radiation(t)
    dc = to_device(c)
    s = torch.cuda.current_stream().cuda_stream
    ref = to_device(alloc_outputs(600, 36))
    rt.run_device(dc, ref, FP64, s)
    modes = {"interp": STUDY | P_TAB32 | P_INTERP32, "tau": STUDY | P_TAU32, "trans": STUDY | P_TRANS32,
             "all": STUDY | P_TAB32 | P_INTERP32 | P_TAU32 | P_TRANS32, "fastdiv": STUDY | P_FASTDIV,
             "fastdiv_tab": STUDY | P_FASTDIV | P_TAB32}
    for name, mode in modes.items():
        o = to_device(alloc_outputs(600, 36))
        rt.run_device(dc, o, mode, s)
        torch.cuda.synchronize()
        rt.check()
        e = max(rel_err(o[k].cpu().numpy(), ref[k].cpu().numpy()) for k in ("fmnetv", "fmneti", "fluxupv", "fluxupi"))
        dh = float((o["hr"] - ref["hr"]).abs().max())
        print(name, f"flux rel {e:.1e}, max |dhr| {dh:.3e} K/sol", flush=True)
        assert all(bool(torch.isfinite(v).all()) for v in o.values()), name
        assert e > 0.0, name                       # the component is active
        if name == "fastdiv":
            assert dh <= 0.05, dh
    o = to_device(alloc_outputs(600, 24))
    _, c24 = synth.make_workload(36, 600, 24)
    with pytest.raises(RadiationError):
        rt.run_device(to_device(c24), o, STUDY | P_TRANS32, s)
    with pytest.raises(RadiationError) as err:
        rt.run_batched([dc], [to_device(alloc_outputs(600, 36))], STUDY | P_TRANS32, s)
    assert err.value.status == -6
