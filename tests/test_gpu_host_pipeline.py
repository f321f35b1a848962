"""mgcm_rt_run_host on grids it cuts into chunks (pinned slots, three streams,
copy threads): the host arrays it returns equal a device-resident run of the same
grid bit for bit, for every output.  Covered: sun-down columns keeping the caller's
DETAU/CUMTAUV, a ragged last chunk, aerosol optics (which shrink the chunks to fit
the pinned slots), the optional outputs absent, the other precision modes, 99
layers, and sizes that grow and shrink the staging between calls."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALL = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi", "diffvt",
       "detau", "cumtauv", "hr", "directsol", "dnvflux")


def _outputs(ncol, nlay, hr=True, surface=True):
    from mgcm_rt import alloc_outputs
    o = alloc_outputs(ncol, nlay)
    rng = np.random.default_rng(ncol + nlay)
    # the caller's DETAU/CUMTAUV, which sun-down columns keep
    o["detau"][:] = rng.uniform(-9.0, -1.0, o["detau"].shape)
    o["cumtauv"][:] = rng.uniform(-9.0, -1.0, o["cumtauv"].shape)
    if not hr:
        del o["hr"]
    if surface:
        o["directsol"] = np.zeros(ncol)
        o["dnvflux"] = np.zeros(ncol)
    return o


def _device_run(rt, c, o, precision):
    import torch
    from mgcm_rt import to_device
    dc, do = to_device(c), to_device({k: v.copy() for k, v in o.items()})
    rt.run_device(dc, do, precision=precision)
    torch.cuda.synchronize()
    rt.check()
    return {k: v.cpu().numpy() for k, v in do.items()}


def _grid(seed, ncol, nlay, aerosol=False):
    from mgcm_rt import synth
    t, c = synth.make_workload(seed, ncol, nlay)
    c["acosz"][::9] = 0.0
    if aerosol:
        c = c | synth.make_aerosol(seed, ncol, nlay)
    return t, c


def _same(got, want, keys):
    # the stored bits: with VIS skipped, DSOLFLUX runs on the caller's (here
    # negative) DETAU and both paths give the same NaNs
    for k in keys:
        assert got[k].shape == want[k].shape and np.array_equal(got[k].view(np.uint64), want[k].view(np.uint64)), k


@pytest.mark.parametrize("ncol,nlay,precision,aerosol,hr,surface", [
    (40003, 24, "FP64", False, True, True),      # 8 chunks, the last one ragged
    (40003, 24, "MIXED", False, False, False),   # no hr, no surface outputs
    (12288, 24, "FP64|SKIP_VIS", False, True, True),  # 3 full chunks; VIS skipped keeps DETAU
    (9001, 24, "FP64|SKIP_IR", False, True, False),
    (6000, 24, "FP64", True, True, True),        # aerosol: chunks cut to the pinned slot
    (9000, 99, "FP64", False, True, True),
])
def test_chunked_run_host_equals_device_run(ncol, nlay, precision, aerosol, hr, surface):
    import mgcm_rt
    from mgcm_rt import Radiation
    prec = 0
    for p in precision.split("|"):
        prec |= getattr(mgcm_rt, p)
    print(f"run_host chunked: ncol={ncol} nlay={nlay} precision={precision} aerosol={aerosol}", flush=True)
    t, c = _grid(ncol + nlay, ncol, nlay, aerosol)
    rt = Radiation(t)
    o = _outputs(ncol, nlay, hr, surface)
    want = _device_run(rt, c, o, prec)
    got = rt.run_host(c, precision=prec, out={k: v.copy() for k, v in o.items()})
    _same(got, want, [k for k in ALL if k in o])
    down = c["acosz"] < np.float32(1e-5)
    assert down.any()
    assert np.array_equal(got["detau"][down], o["detau"][down])  # the caller's values, untouched
    if "SKIP_VIS" in precision:
        assert np.array_equal(got["detau"], o["detau"])


def test_run_host_staging_grows_and_shrinks():
    """One context, grids of changing size and layer count in a row: each call's
    output equals the device-resident run of its grid."""
    from mgcm_rt import Radiation
    t, _ = _grid(5, 10, 24)
    rt = Radiation(t)
    for ncol, nlay in ((300, 24), (20000, 24), (5000, 36), (4097, 24), (1, 24), (33000, 12)):
        print(f"run_host sequence: ncol={ncol} nlay={nlay}", flush=True)
        _, c = _grid(ncol * 7 + nlay, ncol, nlay)
        o = _outputs(ncol, nlay)
        want = _device_run(rt, c, o, 0)
        got = rt.run_host(c, out={k: v.copy() for k, v in o.items()})
        _same(got, want, ALL)


def test_run_host_range_flag_on_a_chunked_grid():
    """A column outside the Planck table's range in the last chunk is still
    reported, and the flag is cleared for the next call."""
    from mgcm_rt import Radiation, RadiationError
    t, c = _grid(9, 20000, 24)
    rt = Radiation(t)
    c["tlev"][19990, 10] = 1200.0
    with pytest.raises(RadiationError) as e:
        rt.run_host(c)
    assert e.value.status == -5
    c["tlev"][19990, 10] = 200.0
    rt.run_host(c)
