"""Host-side logic of the IR surrogate: the fallback threshold calibration and
the target/feature statistics (no device needed)."""
from __future__ import annotations

import numpy as np
import torch

from mgcm_rt.surrogate import calibrate, ir_heating, make_stats, raw_features

def test_calibrated_threshold_meets_the_bar_on_validation():
    dis = torch.tensor([0.5, 0.1, 2.0, 0.05, 3.0])
    err = torch.tensor([0.01, 0.2, 0.01, 0.0, 5.0])
    tau = calibrate(dis, err, 0.05)
    assert tau <= 0.025 and bool((err[dis <= tau] <= 0.05).all())
    assert calibrate(dis, torch.zeros(5), 0.05) == 0.025       # capped at half the bar
    assert calibrate(dis, err, 100.0) == 50.0
    assert abs(calibrate(dis, err, 1.0) - 3.0 * (1 - 1e-6)) > 0 and calibrate(dis, err, 1.0) <= 0.5
    assert calibrate(torch.zeros(5), err, 0.05) == -1.0


def test_stats_standardise_and_floor():
    from mgcm_rt import synth
    c = synth.make_columns(4, 300, 36)
    raw = raw_features(c)
    h = np.random.Generator(np.random.PCG64(1)).normal(0, 100, (300, 36))
    st = make_stats(raw, h, "cpu")
    z = (raw - st["np"]["mean"]) / st["np"]["std"]
    assert np.all(np.isfinite(z)) and np.all(np.abs(z.mean(axis=0)) < 1e-6)
    assert np.all(st["np"]["std"] > 0)
    assert torch.all(st["xlo"] < st["xhi"])


def test_ir_heating_definition():
    """IR heating of GCM layer l from FMNETI over the layer's pressure thickness."""
    from mgcm_rt import synth
    c = synth.make_columns(5, 3, 36)
    f = np.cumsum(np.ones((3, 37)), axis=1)          # FMNETI(L) = L: IRTOTAL = 1 W/m2 in every layer
    h = ir_heating(f, c["plev"])
    l = 5
    dp = c["plev"][:, 2 * l + 2] - c["plev"][:, 2 * l]
    assert np.allclose(h[:, l - 1], 3.72 / (735.94 * 100.0 * dp) * 88775.0)
