"""The reference's k-coefficient data files (LAGINTERP's reads, code/laginterp.f90:59-69):
the library's reader (host code, no GPU), the record writer, and the oracle on the
reference's k-distribution against the reference's own outputs (column_real24.npz)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

REF_DATA = "/root/reference/run/data"
OUT_KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
            "diffvt", "detau", "cumtauv", "hr")
IN_KEYS = ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz")


@pytest.fixture(scope="module")
def real():
    return np.load(os.path.join(GOLDEN, "column_real24.npz"))


def fixture_columns(real) -> dict:
    c = {k: real["in_" + k] for k in IN_KEYS}
    c["ncol"], c["nlay"] = int(real["ncol"]), int(real["nlay"])
    return c


def test_reader_round_trip(tmp_path, real):
    from mgcm_rt import read_ktable, realdata
    for name, nb in (("raw_v", 7), ("raw_i", 5)):
        p = str(tmp_path / name)
        fz = real["fzero_v"] if nb == 7 else real["fzero_i"]
        realdata.write_ktable(p, real[name], fz)
        assert os.path.getsize(p) == (8 + 8 * 7 * 11 * 10 * nb * 17) + (8 + 8 * nb)
        raw, z = read_ktable(p, nb)
        assert np.array_equal(raw, real[name]) and np.array_equal(z, fz)


def _write_records(path, records):
    with open(path, "wb") as f:
        for r in records:
            n = np.array([len(r)], dtype="<u4").tobytes()
            f.write(n + r + n)


def test_reader_skips_the_tail_of_longer_records(tmp_path, real):
    """Fortran's READ takes the list's length from a longer record and skips the rest."""
    from mgcm_rt import read_ktable
    k = real["raw_i"].astype("<f8").tobytes()
    z = real["fzero_i"].astype("<f8").tobytes()
    p = str(tmp_path / "long")
    _write_records(p, [k + b"\x01" * 24, z + b"\x02" * 8, b"third record"])
    raw, fz = read_ktable(p, 5)
    assert np.array_equal(raw, real["raw_i"]) and np.array_equal(fz, real["fzero_i"])


def test_reader_errors(tmp_path, real):
    from mgcm_rt import RadiationError, read_ktable, realdata
    good = str(tmp_path / "good")
    realdata.write_ktable(good, real["raw_i"], real["fzero_i"])
    data = open(good, "rb").read()
    cases = {
        "missing": None,
        "empty": b"",
        "truncated": data[: len(data) // 2],
        "no_second_record": data[: 8 + 8 * 7 * 11 * 10 * 5 * 17],
        "bad_trailing_marker": data[:-1] + b"\x7f",
    }
    for name, blob in cases.items():
        p = str(tmp_path / name)
        if blob is not None:
            open(p, "wb").write(blob)
        with pytest.raises(RadiationError) as e:
            read_ktable(p, 5)
        assert e.value.status == -7, name
    # an IR file read as the VIS table: its first record is too short
    with pytest.raises(RadiationError) as e:
        read_ktable(good, 7)
    assert e.value.status == -7
    for nb in (0, 8):
        with pytest.raises(RadiationError) as e:
            read_ktable(good, nb)
        assert e.value.status == -1


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's data files exist only in the build container")
def test_reference_files_match_fixture(real):
    """The library reads the reference's own files to the fixture's arrays, and the
    writer reproduces the files byte for byte."""
    from mgcm_rt import read_ktable, realdata
    for fname, nb, key, zkey in ((realdata.VIS_FILE, 7, "raw_v", "fzero_v"), (realdata.IR_FILE, 5, "raw_i", "fzero_i")):
        raw, z = read_ktable(os.path.join(REF_DATA, fname), nb)
        assert np.array_equal(raw, real[key]) and np.array_equal(z, real[zkey])


def oracle_real_tables(oracle, real, nlay: int) -> dict:
    """The reference's start-up tables restated by the oracle from the fixture's raw records."""
    from mgcm_rt import realdata, synth
    f = np.load(os.path.join(GOLDEN, "tables_ref.npz"))
    gw = synth.GWEIGHT_REF
    t = synth.base_tables(nlay, np.tile(gw, (7, 1)), np.tile(gw, (5, 1)))
    pint, co2v, co2i = oracle.laginterp(realdata.PGREF11, real["raw_v"], real["raw_i"])
    t.update(co2v=co2v, co2i=co2i, pfgasref=pint, tgasref=realdata.TGASREF.copy(), wrefh2o=real["wrefh2o"],
             fzerov=real["fzero_v"], fzeroi=real["fzero_i"], planckir=f["planckir"], dwni=f["dwni"],
             solarf=real["solarf"], tauray=real["tauray"], sol=real["solarf"] / synth.RSDIST)
    return t


def test_oracle_matches_reference_on_real_tables(oracle, real):
    """The oracle's LAGINTERP and column path on the reference's k-distribution against the
    reference's own LAGINTERP and OPTCV/SFLUXV/OPTCI/SFLUXI (shared Gauss weights)."""
    from mgcm_rt import realdata
    t = oracle_real_tables(oracle, real, 24)
    assert np.array_equal(t["pfgasref"], np.load(os.path.join(GOLDEN, "tables_ref.npz"))["pfgasref"])
    assert np.array_equal(realdata.PGREF11, np.load(os.path.join(GOLDEN, "tables_ref.npz"))["pgref"])
    got = oracle.run(t, fixture_columns(real), nthreads=2)
    for k in OUT_KEYS:
        tol = 1e-11 if k == "hr" else 1e-13
        assert rel_err(got[k], real["out_" + k]) < tol, (k, rel_err(got[k], real["out_" + k]))
    # the real k-distribution skips Gauss points (TLIMIT) and the last VIS band (FZEROV = 0.94)
    n = oracle.count(t, fixture_columns(real))
    full = 32 * (7 * 17 + 5 * 17)
    assert n["sqrts"] / 25 < 0.9 * full


@pytest.mark.skipif(not (os.path.isdir(REF_DATA) and os.path.exists(os.path.join(
    os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libmgcm_ref.so"))),
    reason="the reference's build and data files exist only in the build container")
def test_oracle_laginterp_bit_exact_vs_reference_build(oracle, real):
    """Every entry of the refined k-tables: the oracle against the reference's own LAGINTERP."""
    from oracle import Reference
    from mgcm_rt import realdata
    tab = Reference().tables(os.path.dirname(REF_DATA))
    pint, co2v, co2i = oracle.laginterp(realdata.PGREF11, real["raw_v"], real["raw_i"])
    assert np.array_equal(pint, tab["pfgasref"])
    assert np.array_equal(co2v, tab["co2v"]), int(np.sum(co2v != tab["co2v"]))
    assert np.array_equal(co2i, tab["co2i"]), int(np.sum(co2i != tab["co2i"]))
