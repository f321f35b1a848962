"""Prescribed (Conrath) dust on the GPU (k_dust_tauref): bit for bit against the
reference's outputs and the oracle, and fed into the radiation."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_device_dust_matches_reference_fixture():
    from mgcm_rt import dust
    fx = np.load(os.path.join(GOLDEN, "dust_ref.npz"))
    for ci in (0, 1):
        ptrop = float(fx[f"case{ci}"][1])
        o = dust.tauref(fx[f"case{ci}_prdst"], float(fx["rptau"]), ptrop, fx[f"case{ci}_tautot"],
                        fx[f"case{ci}_conr"], fx[f"case{ci}_p"], fx["sigma"], 24)
        assert np.array_equal(o["tauref"], fx[f"case{ci}_tauref"]), ci
        assert np.array_equal(o["taudst"], fx[f"case{ci}_taudst"]), ci


@pytest.mark.parametrize("nlay,ncol,seed", [(24, 4096, 61), (36, 2048, 62), (1, 64, 63), (99, 256, 64), (24, 1003, 65),
                                             (2, 1, 66)])
def test_device_dust_matches_oracle(oracle, nlay, ncol, seed):
    from mgcm_rt import dust, synth
    rng = np.random.default_rng(seed)
    sigma = synth.sigma_levels(nlay)
    prdst = dust.pressure_grid(7.01, 0.0008)
    tautot, p = rng.uniform(0.05, 3.0, ncol), rng.uniform(3.0, 11.0, ncol)
    conr = np.array([dust.conrath_nu(360.0 * rng.random(), int(rng.integers(1, 37)), 36) for _ in range(ncol)])
    got = dust.tauref(prdst, 6.1, 0.0008, tautot, conr, p, sigma, nlay)
    want = oracle.dust_columns(prdst, 6.1, 0.0008, tautot, conr, p, sigma, nlay)
    assert np.array_equal(got["tauref"], want["tauref"])
    assert np.array_equal(got["taudst"], want["taudst"])


def test_dust_tauref_into_radiation(oracle):
    """TAUREF from the device's Conrath profile drives the radiation like the oracle's."""
    from mgcm_rt import Radiation, dust, synth
    from test_gpu_plain import compare
    t, c = synth.make_workload(64, 300, 24)
    prdst = dust.pressure_grid(7.01, 0.0008)
    rng = np.random.default_rng(64)
    tautot = rng.uniform(0.1, 1.0, 300)
    conr = np.full(300, 0.03)
    psurf = c["plev"][:, -1] - 0.0008  # P(J,I) = surface pressure - PTROP
    tr = dust.tauref(prdst, 6.1, 0.0008, tautot, conr, psurf, synth.sigma_levels(24), 24)["tauref"]
    c["tauref"] = tr
    compare(Radiation(t).run_host(c), oracle.run(t, c))


def test_dust_host_entry_equals_device_entry():
    """mgcm_rt_dust_tauref_host equals mgcm_rt_dust_tauref on device tensors bit for bit."""
    import torch
    from mgcm_rt import dust, synth
    rng = np.random.default_rng(65)
    nlay, ncol = 36, 500
    sigma = synth.sigma_levels(nlay)
    prdst = dust.pressure_grid(7.01, 0.0008)
    tautot, p = rng.uniform(0.05, 3.0, ncol), rng.uniform(3.0, 11.0, ncol)
    conr = rng.uniform(0.01, 0.05, ncol)
    host = dust.tauref(prdst, 6.1, 0.0008, tautot, conr, p, sigma, nlay)
    d = {k: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).cuda()
         for k, v in (("tautot", tautot), ("conr", conr), ("p", p), ("sigma", sigma), ("prdst", prdst))}
    tr = torch.empty((ncol, 2 * nlay + 4), dtype=torch.float64, device="cuda")
    td = torch.empty((ncol, prdst.size), dtype=torch.float64, device="cuda")
    dust.tauref_device(d, 6.1, 0.0008, nlay, tr, td)
    torch.cuda.synchronize()
    assert np.array_equal(host["tauref"], tr.cpu().numpy())
    assert np.array_equal(host["taudst"], td.cpu().numpy())
