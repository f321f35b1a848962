"""DSOLFLUX and COMP3's surface VIS flux on the GPU (k_dsolflux), from the radiation's
own DETAU/DIFFVT outputs: bit for bit against the reference's and the oracle's."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _run(t, c):
    import torch
    from mgcm_rt import Radiation, alloc_outputs, to_device
    rt = Radiation(t)
    dc = to_device(c)
    do = to_device(alloc_outputs(int(c["ncol"]), int(c["nlay"])))
    rt.run_device(dc, do)
    s = rt.dsolflux(dc, do)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in s.items()}, {k: v.cpu().numpy() for k, v in do.items()}


def test_dsolflux_matches_reference_fixture():
    from mgcm_rt import synth
    f = np.load(os.path.join(GOLDEN, "column_ref24.npz"))
    t = synth.make_tables(int(f["seed"]), 24, weights="reference")
    c = {k[3:]: f[k] for k in f.files if k.startswith("in_")}
    c["ncol"], c["nlay"] = 32, 24
    s, o = _run(t, c)
    assert np.array_equal(s["directsol"], np.load(os.path.join(GOLDEN, "dsolflux_ref.npz"))["directsol"])
    assert np.array_equal(s["dnvflux"], o["diffvt"] + s["directsol"])


def test_dsolflux_matches_oracle_on_real_tables(oracle, tmp_path):
    """The reference's k-distribution (optically thin bands give large direct fluxes)."""
    from mgcm_rt import realdata, synth
    real = np.load(os.path.join(GOLDEN, "column_real24.npz"))
    realdata.write_ktable(str(tmp_path / realdata.VIS_FILE), real["raw_v"], real["fzero_v"])
    realdata.write_ktable(str(tmp_path / realdata.IR_FILE), real["raw_i"], real["fzero_i"])
    t = realdata.load_tables(str(tmp_path), 24)
    c = synth.make_columns(71, 2000, 24)
    s, o = _run(t, c)
    want = oracle.dsolflux(t, c["acosz"], o["detau"])
    assert np.array_equal(s["directsol"], want)
    assert np.max(want) > 10.0 and np.sum(want == 0.0) > 0  # lit and dark columns both present


def test_surface_outputs_of_the_run_calls():
    """directsol/dnvflux requested in mgcm_rt_fluxes equal the separate entry's, for the
    host path, the device path and a batch."""
    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    t, c = synth.make_workload(72, 700, 24)
    rt = Radiation(t)
    sep, o = _run(t, c)
    h = alloc_outputs(700, 24)
    h["directsol"], h["dnvflux"] = np.zeros(700), np.zeros(700)
    rt.run_host(c, out=h)
    assert np.array_equal(h["directsol"], sep["directsol"]) and np.array_equal(h["dnvflux"], sep["dnvflux"])
    grids, outs = [], []
    for i, n in enumerate((700, 333)):
        cc = c if i == 0 else synth.make_columns(73, n, 24)
        grids.append(to_device(cc))
        d = to_device(alloc_outputs(n, 24))
        d["directsol"] = torch.zeros(n, dtype=torch.float64, device="cuda")
        d["dnvflux"] = torch.zeros(n, dtype=torch.float64, device="cuda")
        outs.append(d)
    rt.run_batched(grids, outs, FP64)
    torch.cuda.synchronize()
    assert np.array_equal(outs[0]["directsol"].cpu().numpy(), sep["directsol"])
    assert np.array_equal(outs[0]["dnvflux"].cpu().numpy(), sep["dnvflux"])


def test_dsolflux_host_equals_device_entry():
    """mgcm_rt_dsolflux_host (host arrays: COMP3's per-step TEMPGR loop in one call)
    gives mgcm_rt_dsolflux's bits, sun-down columns included."""
    import torch
    from mgcm_rt import Radiation, alloc_outputs, synth, to_device
    t, c = synth.make_workload(81, 400, 24)
    c["acosz"][::7] = 0.0
    rt = Radiation(t)
    dc, do = to_device(c), to_device(alloc_outputs(400, 24))
    rt.run_device(dc, do)
    dev = rt.dsolflux(dc, do)
    torch.cuda.synchronize()
    host = rt.dsolflux_host(c["acosz"], do["detau"].cpu().numpy(), do["diffvt"].cpu().numpy())
    for k in ("directsol", "dnvflux"):
        assert np.array_equal(host[k], dev[k].cpu().numpy()), k
    assert not np.any(host["directsol"][::7])
