"""FILLPT (code/fillpt.f90:30-72), the caller-side builder of the radiation grid's
PLEV/TLEV/PMID/TMID, as the column generator restates it (mgcm_rt.synth.fillpt):
bit for bit against the reference's own FILLPT, from the committed fixture
(tests/golden/fillpt_ref.npz, written by the reference build) and, where the
reference build is present, live on the generator's own columns."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN

OUT = ("plev", "tlev", "pmid", "tmid")


def same_bits(a, b) -> int:
    return int(np.sum(np.asarray(a).view(np.int64) != np.asarray(b).view(np.int64)))


def test_fillpt_matches_reference_fixture():
    from mgcm_rt import synth
    f = np.load(os.path.join(GOLDEN, "fillpt_ref.npz"))
    got = synth.fillpt(f["pl"], float(f["ptrop"]), f["tg"], f["tstrat"], f["tl"])
    assert {k: same_bits(got[k], f["out_" + k]) for k in OUT} == {k: 0 for k in OUT}


def test_make_columns_profiles_are_the_reference_fillpt():
    """The PLEV/TLEV/PMID/TMID every test and the bench feed the path are the reference's
    FILLPT of the generator's GCM-side fields, bit for bit (24 layers: the build's size)."""
    from oracle import REF_SO, Reference
    if not os.path.exists(REF_SO):
        pytest.skip("reference Fortran build not present on this host")
    from mgcm_rt import synth
    seed, ncol, nlay = 2024, 2000, 24
    rng = np.random.Generator(np.random.PCG64(seed + 7919))  # make_columns' stream
    g = synth.gcm_levels(rng, ncol, nlay)
    want = Reference().fillpt(g["pl"], synth.PTROP, g["tg"], g["tstrat"], g["tl"])
    c = synth.make_columns(seed, ncol, nlay)
    assert {k: same_bits(c[k], want[k]) for k in OUT} == {k: 0 for k in OUT}


def test_fillpt_closed_forms():
    """What FILLPT fixes without the reference: the cap levels, the copies, TG at the
    bottom, and odd levels exactly log-p interpolated between their neighbours."""
    from mgcm_rt import synth
    rng = np.random.Generator(np.random.PCG64(9))
    for nlay in (1, 2, 16, 36):
        g = synth.gcm_levels(rng, 50, nlay)
        f = synth.fillpt(g["pl"], synth.PTROP, g["tg"], g["tstrat"], g["tl"])
        nlev = 2 * nlay + 3
        assert np.all(f["plev"][:, :2] == synth.PTROP / 2) and np.array_equal(f["plev"][:, 2:], g["pl"][:, 2:])
        assert np.all(f["tlev"][:, :3] == g["tstrat"][:, None]) and np.array_equal(f["tlev"][:, -1], g["tg"])
        assert np.array_equal(f["tlev"][:, 3:nlev - 1:2], g["tl"][:, 3:nlev - 1:2])
        for k in range(4, nlev - 2, 2):  # 0-based odd levels K = 5..L_LEVELS-2
            lp = np.log(f["plev"][:, [k - 1, k, k + 1]])
            w = (lp[:, 1] - lp[:, 2]) / (lp[:, 0] - lp[:, 2])
            lin = f["tlev"][:, k + 1] + (f["tlev"][:, k - 1] - f["tlev"][:, k + 1]) * w
            assert np.allclose(f["tlev"][:, k], lin, rtol=1e-13, atol=0)
        assert np.array_equal(f["tmid"][:, 2:nlev - 1:2], f["tlev"][:, 2:nlev - 1:2])
        assert np.array_equal(f["tmid"][:, 3:nlev - 1:2], f["tlev"][:, 2:nlev - 2:2])
        assert np.array_equal(f["pmid"][:, 3:nlev - 1:2], f["plev"][:, 2:nlev - 2:2])
        assert f["tmid"][0, 0] == f["tlev"][0, 1] and np.array_equal(f["pmid"][:, -1], f["plev"][:, -1])
