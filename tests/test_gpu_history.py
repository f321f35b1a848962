"""COMP3's history diagnostics on the GPU (k_history), through the C ABI: the
REAL*4 flux fields, surfalb and the non-LTE heating rate bit for bit against the
oracle's restatement of the same outputs, and a history file written from them."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PTROP = 2.0e-4


@pytest.mark.parametrize("nlay,ncol,seed", [(24, 3000, 81), (36, 777, 82), (1, 65, 83), (99, 70, 84)])
def test_history_fields_match_oracle(oracle, nlay, ncol, seed):
    import torch
    from mgcm_rt import HISTORY_KEYS, Radiation, alloc_outputs, synth, to_device
    from oracle import history_fields
    t, c = synth.make_workload(seed, ncol, nlay)
    c["acosz"][::7] = 0.0
    rt = Radiation(t)
    dc, do = to_device(c), to_device(alloc_outputs(ncol, nlay))
    rt.run_device(dc, do)
    sigma = synth.sigma_levels(nlay)
    p = c["plev"][:, -1] - PTROP
    prev = np.full(ncol, -3.0, np.float32)
    hist = {"surfalb": torch.from_numpy(prev.copy()).cuda()}
    h = rt.history_fields(dc, do, hist, p=torch.from_numpy(p).cuda(), sigma=torch.from_numpy(sigma).cuda(),
                          ptrop=PTROP)
    torch.cuda.synchronize()
    o = {k: v.cpu().numpy() for k, v in do.items()}
    want = history_fields(c, o, p=p, sigma=sigma, ptrop=PTROP, surfalb=prev)
    for k in HISTORY_KEYS + ("hr_nlte",):
        got = h[k].cpu().numpy()
        assert got.dtype == want[k].dtype and np.array_equal(got, want[k]), k
    # the device heating equals the run's hr where the sun is down
    down = c["acosz"] < np.float32(1e-5)
    assert np.array_equal(h["hr_nlte"].cpu().numpy()[down], o["hr"][down])
    # and the run's fluxes are the oracle's, so the fields are the oracle's own
    if ncol <= 800:
        ref = history_fields(c, oracle.run(t, c), surfalb=prev)
        for k in HISTORY_KEYS:
            assert np.array_equal(h[k].cpu().numpy(), ref[k]), k


def test_history_fields_without_optional_outputs_and_errors():
    import ctypes as C

    import torch
    from mgcm_rt import (HISTORY_KEYS, History, Radiation, RadiationError, alloc_outputs, column_struct,
                         flux_struct, synth, to_device)
    t, c = synth.make_workload(84, 100, 24)
    rt = Radiation(t)
    dc, do = to_device(c), to_device(alloc_outputs(100, 24))
    rt.run_device(dc, do)
    h = rt.history_fields(dc, do)  # no hr_nlte
    torch.cuda.synchronize()
    assert "hr_nlte" not in h and float(h["fuptopir"].abs().sum()) > 0
    # hr_nlte without p/sigma is an argument error, reported before any launch
    bad = {k: torch.zeros(100, dtype=torch.float32, device="cuda") for k in HISTORY_KEYS}
    hr = torch.zeros((100, 25), dtype=torch.float64, device="cuda")
    hs = History(*[bad[k].data_ptr() for k in HISTORY_KEYS], hr.data_ptr(), None, None, 0.0)
    cs = column_struct(dc, ptr=lambda x: x.data_ptr(), sol=rt.sol)
    fs = flux_struct(do, ptr=lambda x: x.data_ptr())
    assert rt.lib.mgcm_rt_history_fields(C.byref(cs), C.byref(fs), C.byref(hs), None) == -1
    cs.nlay = 0
    assert rt.lib.mgcm_rt_history_fields(C.byref(cs), C.byref(fs), C.byref(hs), None) == -2
    with pytest.raises(RadiationError):
        rt.history_fields(dc, do, p=torch.zeros(100, dtype=torch.float64, device="cuda"))  # sigma missing


def test_history_file_from_device_fields(tmp_path):
    """A grid run, its diagnostics on the device, written as a history time step
    and read back: the stored fields are the device's."""
    import torch
    from mgcm_rt import HISTORY_KEYS, Radiation, alloc_outputs, synth, to_device
    from mgcm_rt import history as H
    jm, im, nlay = 6, 8, 24
    t, c = synth.make_workload(85, jm * im, nlay)
    rt = Radiation(t)
    dc, do = to_device(c), to_device(alloc_outputs(jm * im, nlay))
    rt.run_device(dc, do)
    h = rt.history_fields(dc, do)
    torch.cuda.synchronize()
    fields = {k: H.grid_field(h[k].cpu().numpy(), jm, im) for k in HISTORY_KEYS}
    dims = dict(jm=jm, im=im, layers=nlay, nl=6, ntrace=6)
    with open(tmp_path / "fort.11", "wb") as f:
        H.write_header(f, **dims)
        H.write_step(f, dims, **fields)
    _, steps = H.read_history(str(tmp_path / "fort.11"))
    for k in HISTORY_KEYS:
        assert np.array_equal(steps[0][k], fields[k])
        assert steps[0][k][2, 5] == h[k][5 * jm + 2].item()  # (J=3, I=6) is column (I-1)*L_J + J-1
