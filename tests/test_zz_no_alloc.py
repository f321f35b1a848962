"""mgcm_rt_run_noalloc, the entry point that allocates nothing (its continuum-point
records live in a static device array), through the C ABI against mgcm_rt_run: bit
for bit on two calls in a row, the second on a smaller grid, and grids beyond the
array's 128 MiB rejected with MGCM_RT_ERR_SCRATCH."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
        "diffvt", "detau", "cumtauv", "hr", "directsol", "dnvflux")


def _outputs(ncol, nlay, fill):
    import torch
    from mgcm_rt import alloc_outputs, to_device
    o = to_device(alloc_outputs(ncol, nlay))
    o["directsol"] = torch.zeros(ncol, dtype=torch.float64, device="cuda")
    o["dnvflux"] = torch.zeros(ncol, dtype=torch.float64, device="cuda")
    for v in o.values():
        v.fill_(fill)
    return o


def test_noalloc_equals_run_two_calls_second_smaller():
    import torch
    from mgcm_rt import FP64, MIXED, Radiation, synth, to_device
    rt = Radiation(synth.make_tables(61, 24))
    s = torch.cuda.current_stream().cuda_stream
    # first call: 20,000 columns x 24 layers (100 MB of the 128 MiB of records);
    # second: 3,000 columns x 36 layers, mixed precision; third: back to plain FP64
    for n, nlay, prec, seed in ((20000, 24, FP64, 62), (3000, 36, MIXED, 63), (700, 24, FP64, 64)):
        _, c = synth.make_workload(seed, n, nlay, clouds=(nlay == 36))
        dc = to_device(c)
        want = _outputs(n, nlay, -7.0)   # sun-down columns keep the caller's DETAU/CUMTAUV
        rt.run_device(dc, want, prec, s)
        got = _outputs(n, nlay, -7.0)
        rt.run_noalloc(dc, got, prec, s)
        torch.cuda.synchronize()
        rt.check()
        for k in KEYS:
            a, b = got[k].cpu().numpy(), want[k].cpu().numpy()
            assert np.array_equal(a, b), (n, nlay, k)
        print(n, nlay, prec, "bit-equal to mgcm_rt_run", flush=True)


def test_noalloc_rejects_grids_beyond_its_scratch():
    import torch
    from mgcm_rt import FP64, Radiation, RadiationError, synth, to_device
    rt = Radiation(synth.make_tables(65, 24))
    # 26,887 columns x 24 layers need 26,887 * 12 * 52 * 8 B > 128 MiB
    _, c = synth.make_workload(65, 26887, 24)
    dc = to_device(c)
    o = _outputs(26887, 24, 0.0)
    with pytest.raises(RadiationError) as e:
        rt.run_noalloc(dc, o, FP64, torch.cuda.current_stream().cuda_stream)
    assert e.value.status == -8
    # the largest grid that fits runs and equals mgcm_rt_run
    n = (128 << 20) // (12 * 52 * 8)
    sub = {k: (v[:n] if hasattr(v, "shape") and v.shape[:1] == (26887,) else v) for k, v in dc.items()}
    sub["ncol"] = n
    got, want = _outputs(n, 24, -7.0), _outputs(n, 24, -7.0)
    s = torch.cuda.current_stream().cuda_stream
    rt.run_noalloc(sub, got, FP64, s)
    rt.run_device(sub, want, FP64, s)
    torch.cuda.synchronize()
    for k in KEYS:
        assert torch.equal(got[k], want[k]), k
