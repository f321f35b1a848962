"""The headline workload whole: every column of the 64,800-column x 24-layer grid
(seed 2024, the bench's step) against the oracle at rel 1e-10, through the C ABI with
the inputs resident on the device, as the bench runs it; and the same grid against
the reference's own Fortran (oracle/_ref, built from its sources), with the
reference's Gauss weights, which its SFLUXV/SFLUXI take from the module."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
        "diffvt", "detau", "cumtauv", "hr")


def test_headline_grid_every_column(oracle):
    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    ncol, nlay = 64800, 24
    t, c = synth.make_workload(2024, ncol, nlay)
    rt = Radiation(t)
    dc = to_device(c)
    do = to_device(alloc_outputs(ncol, nlay))
    s = torch.cuda.current_stream()
    rt.run_device(dc, do, FP64, s.cuda_stream)
    torch.cuda.synchronize()
    rt.check()
    want = oracle.run(t, c)
    worst = {}
    for k in KEYS:
        g = do[k].cpu().numpy()
        # per column: |got - want| / max(|want|, the column's max |want|)
        worst[k] = rel_err(g, want[k])
    print({k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    bad = {k: v for k, v in worst.items() if not v <= 1e-10}
    assert not bad, bad


def _reference_built() -> bool:
    import os
    from oracle import REF_SO
    return os.path.exists(REF_SO)


@pytest.mark.skipif(not _reference_built(), reason="oracle/_ref not built (needs the reference's sources)")
def test_headline_grid_every_column_against_the_reference_fortran():
    import torch
    from oracle import Reference
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    ncol, nlay = 64800, 24
    t = synth.make_tables(2024, nlay, weights="reference")
    c = synth.make_columns(2024, ncol, nlay)
    rt = Radiation(t)
    do = to_device(alloc_outputs(ncol, nlay))
    rt.run_device(to_device(c), do, FP64, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rt.check()
    want = Reference().run_grid(t, c)  # about 14 s on one host core
    worst = {k: rel_err(do[k].cpu().numpy(), want[k]) for k in KEYS}
    print("vs the reference Fortran:", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    bad = {k: v for k, v in worst.items() if not v <= 1e-10}
    assert not bad, bad
