"""Aerosol producers (CPU): the library's reader of INITCLD's optical-property files
and its size grid, and the oracle's OPT_DST/OPT_CLD against the reference's outputs
(aerosol_ref.npz; the reference build itself where it exists)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

REF_RUN = "/root/reference/run"
HAVE_REF = os.path.isdir(REF_RUN) and os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libmgcm_ref.so"))
KEYS = ("qxv", "qsv", "gv", "qxi", "qsi", "gi", "qextref", "tauref", "tau2")


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLDEN, "aerosol_ref.npz"))


def write_files(d, fx):
    for k in fx.files:
        if k.startswith("file_"):
            open(os.path.join(d, k[5:]), "wb").write(fx[k].tobytes())


def parse_fixed(blob: bytes, n: int, w: int) -> np.ndarray:
    """Independent reading of '(n(Ew.d,x))' records: REAL(4) values."""
    rows = []
    for line in blob.decode().splitlines():
        rows.append([np.float32(line[j * (w + 1):j * (w + 1) + w]) for j in range(n)])
    return np.array(rows, dtype=np.float64)


def test_reader_matches_independent_parse_and_grid(tmp_path, fx, oracle):
    from mgcm_rt import aerosol_table_views, read_aerosol_tables
    write_files(str(tmp_path), fx)
    v = aerosol_table_views(read_aerosol_tables(str(tmp_path)))
    _, radb = oracle.aer_grid()
    assert np.array_equal(v["radb_rt"], radb)
    dv = parse_fixed(fx["file_Dust_vis_wolff2010_JD_12bands.dat"].tobytes(), 7, 11)   # (20*3, 7)
    di = parse_fixed(fx["file_Dust_ir_wolff2010_JD_12bands.dat"].tobytes(), 5, 11)
    cv = parse_fixed(fx["file_waterCoated_vis_JD_12bands.dat"].tobytes(), 7, 12)      # (15*20*3, 7)
    ci = parse_fixed(fx["file_waterCoated_ir_JD_12bands.dat"].tobytes(), 5, 12)
    # records: bin-major then (Qext, Qscat, g); clouds core-ratio-major, then bin
    assert np.array_equal(v["q_v_dst"], dv.reshape(20, 3, 7).transpose(1, 2, 0))
    assert np.array_equal(v["q_i_dst"], di.reshape(20, 3, 5).transpose(1, 2, 0))
    assert np.array_equal(v["q_v_cld"], cv.reshape(15, 20, 3, 7).transpose(2, 3, 1, 0))
    assert np.array_equal(v["q_i_cld"], ci.reshape(15, 20, 3, 5).transpose(2, 3, 1, 0))


def test_reader_errors(tmp_path, fx):
    from mgcm_rt import RadiationError, read_aerosol_tables
    with pytest.raises(RadiationError) as e:
        read_aerosol_tables(str(tmp_path / "missing"))
    assert e.value.status == -7
    write_files(str(tmp_path), fx)
    p = tmp_path / "Dust_ir_wolff2010_JD_12bands.dat"
    blob = p.read_bytes()
    p.write_bytes(blob[: len(blob) // 2])  # truncated
    with pytest.raises(RadiationError) as e:
        read_aerosol_tables(str(tmp_path))
    assert e.value.status == -7
    p.write_bytes(blob.replace(b"E-03", b"X-03", 1))  # not a number
    with pytest.raises(RadiationError) as e:
        read_aerosol_tables(str(tmp_path))
    assert e.value.status == -7


def test_oracle_producers_bit_exact_vs_reference_fixture(tmp_path, fx, oracle):
    from mgcm_rt import aerosol_table_views, read_aerosol_tables
    write_files(str(tmp_path), fx)
    v = aerosol_table_views(read_aerosol_tables(str(tmp_path)))
    for name, cloud in (("dst", False), ("cld", True)):
        o = oracle.opt_aer(cloud, v["radb_rt"], v["q_v_" + name], v["q_i_" + name], fx["qtrace"], fx["pl"])
        for k in KEYS:
            assert np.array_equal(o[k], fx[f"{name}_{k}"]), (name, k)
    assert int(np.sum(fx["dst_tauref"] > 0)) > 0 and int(np.sum(fx["cld_tauref"] > 0)) > 0


@pytest.mark.skipif(not HAVE_REF, reason="the reference build and data exist only in the build container")
def test_reader_and_oracle_vs_reference_build(oracle):
    from oracle import Reference
    from mgcm_rt import aerosol_table_views, read_aerosol_tables, synth
    ref = Reference()
    tb = ref.initcld(REF_RUN)
    v = aerosol_table_views(read_aerosol_tables(os.path.join(REF_RUN, "data")))
    assert np.array_equal(v["radb_rt"], tb["radb_rt"])
    for a, b in (("q_v_dst", "qv_dst"), ("q_i_dst", "qi_dst"), ("q_v_cld", "qv_cld"), ("q_i_cld", "qi_cld")):
        assert np.array_equal(v[a], tb[b]), a
    qtr = synth.make_tracers(109, 24, 24)
    pl = synth.make_columns(109, 24, 24)["plev"]
    for name, cloud in (("dst", False), ("cld", True)):
        o = oracle.opt_aer(cloud, v["radb_rt"], v["q_v_" + name], v["q_i_" + name], qtr, pl)
        for i in range(24):
            r = ref.opt_aer(cloud, qtr[i], pl[i])
            for k in KEYS:
                assert np.array_equal(o[k][i], r[k]), (name, i, k)
