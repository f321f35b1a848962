"""Prescribed (Conrath) dust on the CPU: the oracle's DUSTPROFILE + FILLTAUCUM and the
library's pressure grid against the reference's outputs (dust_ref.npz; the reference
build itself where it exists)."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HAVE_REF = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libmgcm_ref.so"))


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLDEN, "dust_ref.npz"))


def test_pressure_grid_three_ways(fx, oracle):
    from mgcm_rt import dust, load_library
    lib = load_library()
    lib.mgcm_rt_dust_grid.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    for ci in (0, 1):
        psf, ptrop = fx[f"case{ci}"][:2]
        want = fx[f"case{ci}_prdst"]
        assert np.array_equal(oracle.dust_grid(psf, ptrop), want)
        assert np.array_equal(dust.pressure_grid(psf, ptrop), want)
        got = np.zeros(100)
        assert lib.mgcm_rt_dust_grid(psf, ptrop, 100, got.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        assert np.array_equal(got, want)
    bad = np.zeros(200)
    assert lib.mgcm_rt_dust_grid(6.1, 0.0005, 200, bad.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == -1


def test_oracle_bit_exact_vs_reference_fixture(fx, oracle):
    from mgcm_rt import dust
    for ci in (0, 1):
        psf, ptrop, vary, vpout = fx[f"case{ci}"]
        conr = np.array([dust.conrath_nu(vpout, int(j), 36, bool(vary)) for j, _ in fx[f"case{ci}_cols"]])
        assert np.array_equal(conr, fx[f"case{ci}_conr"])
        o = oracle.dust_columns(fx[f"case{ci}_prdst"], float(fx["rptau"]), ptrop, fx[f"case{ci}_tautot"], conr,
                                fx[f"case{ci}_p"], fx["sigma"], 24)
        for k in ("taudst", "tauref", "taucum"):
            assert np.array_equal(o[k], fx[f"case{ci}_{k}"]), (ci, k)


@pytest.mark.skipif(not HAVE_REF, reason="the reference build exists only in the build container")
def test_oracle_vs_reference_build_many_columns(oracle):
    from oracle import Reference
    from mgcm_rt import dust, synth
    ref = Reference()
    rng = np.random.default_rng(12)
    sigma = synth.sigma_levels(24)
    tautot = rng.uniform(0.05, 2.0, (60, 36))
    cols = [(j, i) for j in range(1, 37, 5) for i in range(1, 61, 11)]
    p = {c: rng.uniform(3.0, 11.0) for c in cols}
    for vary, vpout in ((False, 0.0), (True, 120.0)):
        r = ref.dust(6.1, 0.0005, 6.1, 0.03, vary, vpout, tautot, p, sigma, cols)
        conr = np.array([dust.conrath_nu(vpout, j, 36, vary) for j, _ in cols])
        o = oracle.dust_columns(r["prdst"], 6.1, 0.0005, np.array([tautot[i - 1, j - 1] for j, i in cols]), conr,
                                np.array([p[c] for c in cols]), sigma, 24)
        for k in ("taudst", "tauref", "taucum"):
            assert np.array_equal(o[k], r[k]), (vary, k)
