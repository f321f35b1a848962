"""The real-data table pipeline on the GPU: the reference's k-coefficient files (written
from the fixture's records) read by the library, refined by the device LAGINTERP, the
device Planck table, and the plain FP64 path on the reference's k-distribution."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err
from test_gpu_plain import compare
from test_realdata import fixture_columns

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def real():
    return np.load(os.path.join(GOLDEN, "column_real24.npz"))


@pytest.fixture(scope="module")
def data_dir(tmp_path_factory, real):
    from mgcm_rt import realdata
    d = tmp_path_factory.mktemp("data")
    realdata.write_ktable(str(d / realdata.VIS_FILE), real["raw_v"], real["fzero_v"])
    realdata.write_ktable(str(d / realdata.IR_FILE), real["raw_i"], real["fzero_i"])
    return str(d)


def test_device_tables_match_reference_and_oracle(oracle, real, data_dir):
    from mgcm_rt import realdata
    f = np.load(os.path.join(GOLDEN, "tables_ref.npz"))
    t = realdata.load_tables(data_dir, 24)
    assert np.array_equal(t["pfgasref"], f["pfgasref"])
    assert np.array_equal(t["dwni"], f["dwni"])
    # bit for bit: a one-ulp change of PLANCKIR moves the IR fluxes on this k-distribution by ~1e-8
    assert np.array_equal(t["planckir"], f["planckir"]), int(np.sum(t["planckir"] != f["planckir"]))
    assert np.array_equal(t["fzerov"], real["fzero_v"]) and np.array_equal(t["fzeroi"], real["fzero_i"])
    # the reference's own LAGINTERP rows
    for (b, g, h, tt), want in zip(f["lag_rows"], f["lag_out"]):
        got = t["co2v"][g, b, h, :, tt] if b < 7 else t["co2i"][g, b - 7, h, :, tt]
        assert rel_err(got, want) < 1e-13
    # every entry against the oracle's LAGINTERP, which equals the reference's bit for bit
    _, co2v, co2i = oracle.laginterp(realdata.PGREF11, real["raw_v"], real["raw_i"])
    assert np.array_equal(t["co2v"], co2v), int(np.sum(t["co2v"] != co2v))
    assert np.array_equal(t["co2i"], co2i), int(np.sum(t["co2i"] != co2i))


def test_plain_path_on_real_tables_matches_reference(real, data_dir):
    """Device tables + plain FP64 path against the reference's outputs on its own k-distribution."""
    from mgcm_rt import Radiation, realdata
    t = realdata.load_tables(data_dir, 24)
    # the start-up tables are the reference's own, bit for bit (SETSPV: code/setspv.f90:67-105)
    assert np.array_equal(t["solarf"], real["solarf"]) and np.array_equal(t["tauray"], real["tauray"])
    t["sol"] = real["solarf"] / 2.4277  # SOLARF/RSDIST of the fixture's orbit (code/comp3.f:300-304)
    got = Radiation(t).run_host(fixture_columns(real))
    want = {k: real["out_" + k] for k in ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv",
                                          "nfluxtopi", "diffvt", "detau", "cumtauv", "hr")}
    errs = compare(got, want)
    print("real tables vs reference", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)


VIS_KEYS = ("fmnetv", "fluxupv", "fluxdnv", "nfluxtopv", "diffvt", "detau", "cumtauv")
IR_KEYS = ("fmneti", "fluxupi", "fluxdni", "nfluxtopi", "hr")


def column_errors(got: dict, want: dict, keys) -> np.ndarray:
    """Per column: max over keys of |got - want| / max(|want|, the column's max |want|)."""
    out = np.zeros(np.asarray(want[keys[0]]).shape[0])
    for k in keys:
        x, y = np.asarray(got[k]).reshape(len(out), -1), np.asarray(want[k]).reshape(len(out), -1)
        s = np.max(np.abs(y), axis=1, keepdims=True)
        out = np.maximum(out, np.max(np.abs(x - y) / np.maximum(np.maximum(np.abs(y), s), 1e-300), axis=1))
    return out


@pytest.mark.parametrize("nlay,ncol,seed", [(24, 1024, 31), (36, 512, 32), (16, 256, 33)])
def test_real_tables_match_oracle(oracle, real, data_dir, nlay, ncol, seed):
    """On the reference's k-distribution the IR solve is ill-conditioned in the thin top
    layers (one ulp of PMID moves its IR fluxes by up to ~1e-8 and the top heating rates
    by ~1e-5), so parity needs every operation to be the reference's, TPINDEX's LOG10
    included: VIS and IR at rel 1e-10 in every column."""
    from mgcm_rt import Radiation, realdata, synth
    t = realdata.load_tables(data_dir, nlay)
    c = synth.make_columns(seed, ncol, nlay, clouds=(nlay == 36))
    # the oracle's own LAGINTERP; the remaining tables are the device-built ones
    _, co2v, co2i = oracle.laginterp(realdata.PGREF11, real["raw_v"], real["raw_i"])
    tor = t | {"co2v": co2v, "co2i": co2i}
    want = oracle.run(tor, c)
    got = Radiation(t).run_host(c)
    ev = column_errors(got, want, VIS_KEYS)
    ei = column_errors(got, want, IR_KEYS)
    print(nlay, ncol, f"VIS max {ev.max():.1e}, IR max {ei.max():.1e}", flush=True)
    assert ev.max() <= 1e-10
    assert ei.max() <= 1e-10, np.where(ei > 1e-10)[0].tolist()


def test_mixed_path_on_real_tables(oracle, data_dir):
    """The mixed path (k-tables stored in FP32, FP64 arithmetic) on the reference's
    k-distribution, 36 layers with cloud ice, against the FP64 oracle: within twice
    the floor that rounding the tables to FP32 sets for an all-FP64 evaluation
    (plus 0.05 K/sol), and RMS below 0.05 K/sol."""
    from mgcm_rt import MIXED, Radiation, realdata, synth
    t = realdata.load_tables(data_dir, 36)
    c = synth.make_columns(41, 1024, 36, clouds=True)
    ref = oracle.run(t, c)
    t32 = dict(t)
    for k in ("co2v", "co2i"):
        t32[k] = t[k].astype(np.float32).astype(np.float64)
    floor = float(np.abs(oracle.run(t32, c)["hr"] - ref["hr"]).max())
    got = Radiation(t).run_host(c, precision=MIXED)
    err = np.abs(got["hr"] - ref["hr"])
    rms = float(np.sqrt((err ** 2).mean()))
    print(f"real tables, 36 layers: mixed max {err.max():.3e} K/sol rms {rms:.3e} (floor {floor:.3e}); "
          f"|hr| median {np.median(np.abs(ref['hr'])):.3e} max {np.abs(ref['hr']).max():.3e} K/sol", flush=True)
    assert err.max() <= 2.0 * floor + 0.05
    assert rms <= 0.05


def _reference_built() -> bool:
    from oracle import REF_SO
    return os.path.exists(REF_SO)


@pytest.mark.skipif(not _reference_built(), reason="oracle/_ref not built (needs the reference's sources)")
def test_real_tables_match_the_reference_fortran(data_dir):
    """8,100 seeded columns x 24 layers on the reference's k-distribution (the device-built
    tables, which equal the reference's bit for bit) against the reference's own Fortran
    on the same tables and columns: VIS and IR at rel 1e-10 in every column."""
    from oracle import Reference
    from mgcm_rt import Radiation, realdata, synth
    t = realdata.load_tables(data_dir, 24)
    c = synth.make_columns(35, 8100, 24)
    got = Radiation(t).run_host(c)
    want = Reference().run_grid(t, c)
    ev = column_errors(got, want, VIS_KEYS)
    ei = column_errors(got, want, IR_KEYS)
    print(f"vs the reference Fortran: VIS max {ev.max():.1e}, IR max {ei.max():.1e}", flush=True)
    assert ev.max() <= 1e-10
    assert ei.max() <= 1e-10, np.where(ei > 1e-10)[0].tolist()
