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
