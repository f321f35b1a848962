"""Fortran host binding (fortran/mgcm_rt_mod.f90 + mgcm_rt_driver.f90).

CPU: the driver program was built against libmgcm_rt.so and answers its usage
line without touching a device; the case file round-trips.  GPU: a batch run
through the Fortran program equals the Python binding of the same C ABI bit
for bit, and a bad layer count comes back as a status, not an abort."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

DRIVER = os.path.join(PKG, "fortran", "mgcm_rt_driver")


def test_driver_built_and_usage():
    assert os.path.exists(DRIVER), "build() builds the Fortran driver"
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stdout


def test_case_file_layout(tmp_path):
    from mgcm_rt import casefile, synth
    t, c = synth.make_workload(5, 3, 24)
    p = str(tmp_path / "case.bin")
    casefile.write_case(p, t, c, precision=1)
    raw = np.fromfile(p, dtype=np.uint8)
    hdr = np.frombuffer(raw[:12].tobytes(), dtype=np.int32)
    assert hdr.tolist() == [3, 24, 1]
    nlev = 51
    n_tab = 7 * 51 * 10 * 7 * 17 + 7 * 51 * 10 * 5 * 17 + 51 + 7 + 10 + 17 * 7 + 17 * 5 + 7 + 5 + 7 + 5 + 5 * 8501
    n_aer = 6 * 7 + 6 * 5
    n_col = 3 * (5 * nlev + 2 * (nlev + 1) + 3)
    assert raw.size == 12 + 8 * (n_tab + n_aer + 2 + 7 + n_col)
    body = np.frombuffer(raw[12:12 + 8 * 7 * 51 * 10 * 7 * 17].tobytes(), dtype=np.float64)
    assert np.array_equal(body, t["co2v"].ravel())


@pytest.mark.gpu
@pytest.mark.parametrize("nlay,precision", [(24, 0), (36, 0), (24, 1)])
def test_fortran_driver_matches_python_binding(tmp_path, nlay, precision):
    from mgcm_rt import Radiation, casefile, synth
    t, c = synth.make_workload(17, 300, nlay, clouds=(nlay == 36))
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    casefile.write_case(case, t, c, precision)
    r = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    status, got = casefile.read_result(res, 300, nlay)
    assert status == 0
    want = Radiation(t).run_host(c, precision=precision)
    for k in casefile.RESULT_KEYS:
        assert np.array_equal(got[k], want[k]), k


@pytest.mark.gpu
def test_fortran_driver_reports_bad_layer_count(tmp_path):
    from mgcm_rt import casefile, synth
    t, _ = synth.make_workload(18, 2, 24)
    nlev = 2 * 100 + 3  # arrays sized for the header; the library rejects nlay > 99 before reading them
    c = {"ncol": 2, "nlay": 100, "als": np.zeros(2), "albi": np.zeros(2), "acosz": np.zeros(2)}
    c.update({k: np.zeros((2, nlev)) for k in ("plev", "tlev", "pmid", "tmid", "qh2o")})
    c.update({k: np.zeros((2, nlev + 1)) for k in ("tauref", "taurefcld")})
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    casefile.write_case(case, t, c)
    r = subprocess.run([DRIVER, case, res], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    status, _ = casefile.read_result(res, 2, 100)
    assert status == -2
    assert "layer count" in r.stdout


def test_fortran_driver_reports_missing_data_files(tmp_path):
    """RADSETUP from a data directory without the k-coefficient files: a status, no abort
    (the files are read before any device work, so this runs without a GPU)."""
    from mgcm_rt import casefile, synth
    t, c = synth.make_workload(19, 2, 24)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    casefile.write_case(case, t, c)
    r = subprocess.run([DRIVER, case, res, str(tmp_path / "no_such_dir")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 1
    status, _ = casefile.read_result(res, 2, 24)
    assert status == -7
    assert "data file" in r.stdout


@pytest.mark.gpu
def test_fortran_driver_on_reference_data_files(tmp_path):
    """The driver sets the k-tables up from the reference's data files (written from the
    fixture's records) and reproduces the reference's outputs on its own k-distribution."""
    from mgcm_rt import casefile, realdata
    from conftest import GOLDEN, rel_err
    real = np.load(os.path.join(GOLDEN, "column_real24.npz"))
    d = tmp_path / "data"
    d.mkdir()
    realdata.write_ktable(str(d / realdata.VIS_FILE), real["raw_v"], real["fzero_v"])
    realdata.write_ktable(str(d / realdata.IR_FILE), real["raw_i"], real["fzero_i"])
    t = realdata.load_tables(str(d), 24)
    t["tauray"], t["sol"] = real["tauray"], real["solarf"] / 2.4277
    # the case file carries other k-tables; the driver replaces them from the files
    t["co2v"], t["co2i"] = t["co2v"] * 0.5, t["co2i"] * 2.0
    t["fzerov"], t["fzeroi"] = np.zeros(7), np.zeros(5)
    c = {k: real["in_" + k] for k in casefile.COLUMN_KEYS}
    c["ncol"], c["nlay"] = 32, 24
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "res.bin")
    casefile.write_case(case, t, c)
    r = subprocess.run([DRIVER, case, res, str(d)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    status, got = casefile.read_result(res, 32, 24)
    assert status == 0
    for k in casefile.RESULT_KEYS:
        e = rel_err(got[k], real["out_" + k])
        assert e <= 1e-10, (k, e)


@pytest.mark.gpu
def test_fortran_driver_history_records(tmp_path):
    """The driver fills the history's REAL*4 flux fields as COMP3 does and writes
    MHISTV's two radiation records with Fortran unformatted I/O; the records read
    back with mgcm_rt.history equal the device's k_history fields bit for bit."""
    import torch
    from mgcm_rt import HISTORY_KEYS, Radiation, alloc_outputs, casefile, synth, to_device
    from mgcm_rt import history as H
    t, c = synth.make_workload(19, 250, 24)
    c["acosz"][::6] = 0.0
    case, res, hist = str(tmp_path / "case.bin"), str(tmp_path / "res.bin"), str(tmp_path / "fort.11")
    casefile.write_case(case, t, c, 0)
    r = subprocess.run([DRIVER, case, res, "-", hist], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(hist, "rb") as f:
        rec1, rec2 = H.read_record(f), H.read_record(f)
        assert H.read_record(f) is None
    f1 = np.frombuffer(rec1, "<f4").reshape(4, 250)
    f2 = np.frombuffer(rec2, "<f4").reshape(3, 250)
    rt = Radiation(t)
    dc, do = to_device(c), to_device(alloc_outputs(250, 24))
    rt.run_device(dc, do)
    h = rt.history_fields(dc, do)
    torch.cuda.synchronize()
    got = dict(zip(HISTORY_KEYS[:7], list(f1) + list(f2)))
    for k in HISTORY_KEYS[:7]:
        assert np.array_equal(got[k], h[k].cpu().numpy()), k
    assert not np.any(got["fuptopv"][::6]) and np.all(got["fuptopir"] > 0)


@pytest.mark.gpu
def test_fortran_per_column_shim_equals_batch(tmp_path):
    """COMP3's loop shape: mgcm_rt_column once per column, DETAU/CUMTAUV as the
    module arrays indexed (J, I), gives the batch call's outputs bit for bit."""
    from mgcm_rt import casefile, synth
    t, c = synth.make_workload(23, 40, 24, clouds=True)
    c["acosz"][::5] = 0.0
    case = str(tmp_path / "case.bin")
    casefile.write_case(case, t, c, 0)
    outs = {}
    for mode in ("batch", "column"):
        res = str(tmp_path / f"res_{mode}.bin")
        args = [DRIVER] + (["--per-column"] if mode == "column" else []) + [case, res]
        r = subprocess.run(args, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        status, outs[mode] = casefile.read_result(res, 40, 24)
        assert status == 0
    for k in casefile.RESULT_KEYS:
        if k == "hr":
            continue
        assert np.array_equal(outs["column"][k], outs["batch"][k]), k
    assert not np.any(outs["column"]["fluxdnv"][::5])  # sun-down columns


@pytest.mark.gpu
def test_fortran_tiles_equal(tmp_path):
    """The column kernel's tiles chosen from Fortran (mgcm_rt_set_tile through
    --tile) give the same bits, batch and per-column calls alike."""
    from mgcm_rt import casefile, synth
    t, c = synth.make_workload(24, 13, 24)
    case = str(tmp_path / "case.bin")
    casefile.write_case(case, t, c, 0)
    outs = {}
    for mode in ("batch", "column"):
        for tile in (1, 2, 4, 12):
            res = str(tmp_path / f"res_{mode}_{tile}.bin")
            args = [DRIVER, "--tile", str(tile)] + (["--per-column"] if mode == "column" else []) + [case, res]
            r = subprocess.run(args, capture_output=True, text=True, timeout=300)
            assert r.returncode == 0, r.stdout + r.stderr
            status, outs[mode, tile] = casefile.read_result(res, 13, 24)
            assert status == 0
    for mode in ("batch", "column"):
        for tile in (2, 4, 12):
            for k in casefile.RESULT_KEYS:
                if mode == "column" and k == "hr":
                    continue
                assert np.array_equal(outs[mode, tile][k], outs[mode, 1][k]), (mode, tile, k)


@pytest.mark.gpu
def test_fortran_driver_active_dust_and_clouds(tmp_path):
    """COMP3's active dust and clouds from Fortran: INITCLD's tables read from the
    reference's optical-property files, OPT_DST and OPT_CLD on host arrays
    (mgcm_rt_opt_aer over mgcm_rt_opt_*_host), their optics through the aerosol
    argument of mgcm_rt_radiation.  Equal bit for bit to the same chain through
    the Python binding."""
    from mgcm_rt import Radiation, casefile, read_aerosol_tables, synth
    from test_aerosol_producers import write_files
    from conftest import GOLDEN
    fx = np.load(os.path.join(GOLDEN, "aerosol_ref.npz"))
    aerdir = tmp_path / "aer"
    aerdir.mkdir()
    write_files(str(aerdir), fx)
    nlay, ncol = 24, 120
    t, c = synth.make_workload(29, ncol, nlay)
    qtr = synth.make_tracers(29, ncol, nlay)                    # [ncol][6][nlay] = Fortran (nlay, 6, ncol)
    case, res, trf = str(tmp_path / "case.bin"), str(tmp_path / "res.bin"), str(tmp_path / "qtrace.bin")
    casefile.write_case(case, t, c, 0)
    np.ascontiguousarray(qtr, dtype=np.float64).tofile(trf)
    r = subprocess.run([DRIVER, "--aerosol", trf, str(aerdir), case, res], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    status, got = casefile.read_result(res, ncol, nlay)
    assert status == 0
    rt = Radiation(t)
    rt.set_aerosol_tables(read_aerosol_tables(str(aerdir)))
    d, k = rt.opt_aerosol_host(False, c["plev"], qtr), rt.opt_aerosol_host(True, c["plev"], qtr)
    cc = dict(c)
    cc["tauref"], cc["taurefcld"] = d["tauref"], k["tauref"]
    for q in ("qxv", "qsv", "gv", "qxi", "qsi", "gi"):
        cc[q + "_dst"], cc[q + "_cld"] = d[q], k[q]
    cc["qextref_dst"], cc["qextref_cld"] = d["qextref"], k["qextref"]
    want = rt.run_host(cc)
    for key in casefile.RESULT_KEYS:
        assert np.array_equal(got[key], want[key]), key
    assert np.any(d["tauref"] > 0) and np.any(k["tauref"] > 0)  # dust and cloud are both active


@pytest.mark.gpu
def test_fortran_driver_prescribed_dust(tmp_path):
    """COMP3's prescribed-dust path from Fortran: PRDST (mgcm_rt_dust_grid) and each
    column's TAUREF (mgcm_rt_dust_tauref_host) through the module's bind(C) interfaces
    (mgcm_rt_dust_profile), then the radiation; equal bit for bit to the same chain
    through the Python binding."""
    from mgcm_rt import Radiation, casefile, dust, synth
    nlay, ncol = 24, 96
    t, c = synth.make_workload(31, ncol, nlay)
    rng = np.random.Generator(np.random.PCG64(31))
    psf, ptrop, rptau = 7.01, synth.PTROP, 6.1
    tautot = rng.uniform(0.1, 1.0, ncol)
    conr = np.array([dust.conrath_nu(0.0, j, 36, True) for j in rng.integers(1, 37, ncol)])
    p = c["plev"][:, -1] - ptrop
    sigma = synth.sigma_levels(nlay)
    case, res, df = str(tmp_path / "case.bin"), str(tmp_path / "res.bin"), str(tmp_path / "dust.bin")
    casefile.write_case(case, t, c, 0)
    np.concatenate([[psf, ptrop, rptau], tautot, conr, p, sigma]).astype(np.float64).tofile(df)
    r = subprocess.run([DRIVER, "--dust", df, case, res], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    status, got = casefile.read_result(res, ncol, nlay)
    assert status == 0
    prdst = dust.pressure_grid(psf, ptrop)
    tr = dust.tauref(prdst, rptau, ptrop, tautot, conr, p, sigma, nlay)["tauref"]
    assert np.any(tr > 0)
    want = Radiation(t).run_host(c | {"tauref": tr})
    for key in casefile.RESULT_KEYS:
        assert np.array_equal(got[key], want[key]), key
