"""Aerosol producers on the GPU: OPT_DST and OPT_CLD (k_opt_aer) against the oracle, and
the chain producers -> radiation with per-level aerosol optics (active dust and clouds)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err
from test_aerosol_producers import KEYS, write_files

pytestmark = pytest.mark.gpu
# erf, log and pow are the device's own (log correctly rounded, erf/pow within an ulp or
# two of the host's), so the bar is a few ulps of the weighted sums, not bit equality
TOL = 1e-13


@pytest.fixture(scope="module")
def tables(tmp_path_factory):
    from mgcm_rt import read_aerosol_tables
    fx = np.load(os.path.join(GOLDEN, "aerosol_ref.npz"))
    d = tmp_path_factory.mktemp("aer")
    write_files(str(d), fx)
    return read_aerosol_tables(str(d))


def test_producers_match_reference_fixture(tables):
    from mgcm_rt import Radiation, synth
    fx = np.load(os.path.join(GOLDEN, "aerosol_ref.npz"))
    rt = Radiation(synth.make_tables(1, 24))
    rt.set_aerosol_tables(tables)
    for name, cloud in (("dst", False), ("cld", True)):
        o = rt.opt_aerosol_host(cloud, fx["pl"], fx["qtrace"])
        errs = {k: rel_err(o[k], fx[f"{name}_{k}"]) for k in KEYS}
        print(name, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        assert all(v <= TOL for v in errs.values()), errs
        # levels and layers without aerosol carry OPT_DST's defaults exactly
        off = fx[f"{name}_tauref"] == 0
        assert np.array_equal(o["qextref"][off], fx[f"{name}_qextref"][off])


@pytest.mark.parametrize("nlay,ncol,seed", [(24, 2048, 51), (36, 1024, 52), (64, 256, 53), (1, 64, 54), (65, 64, 55),
                                            (99, 128, 56)])
def test_producers_match_oracle(tables, oracle, nlay, ncol, seed):
    from mgcm_rt import Radiation, aerosol_table_views, synth
    v = aerosol_table_views(tables)
    rt = Radiation(synth.make_tables(1, nlay))
    rt.set_aerosol_tables(tables)
    qtr = synth.make_tracers(seed, ncol, nlay)
    pl = synth.make_columns(seed, ncol, nlay)["plev"]
    for name, cloud in (("dst", False), ("cld", True)):
        o = rt.opt_aerosol_host(cloud, pl, qtr)
        w = oracle.opt_aer(cloud, v["radb_rt"], v["q_v_" + name], v["q_i_" + name], qtr, pl)
        errs = {k: rel_err(o[k], w[k]) for k in KEYS}
        same = np.mean([np.array_equal(o[k], w[k]) for k in KEYS])
        print(nlay, ncol, name, f"max {max(errs.values()):.1e}, arrays bit-equal {same:.2f}", flush=True)
        assert all(e <= TOL for e in errs.values()), errs


def test_active_dust_and_cloud_chain_matches_oracle(tables, oracle):
    """OPT_DST and OPT_CLD feed TAUREF, TAUREFCLD and the per-level optics of the
    radiation, as COMP3 does (code/comp3.f:595-631), on the device and in the oracle."""
    from mgcm_rt import Radiation, aerosol_table_views, synth
    from test_gpu_plain import compare
    nlay, ncol, seed = 24, 512, 55
    t, c = synth.make_workload(seed, ncol, nlay)
    rt = Radiation(t)
    rt.set_aerosol_tables(tables)
    qtr = synth.make_tracers(seed, ncol, nlay)
    v = aerosol_table_views(tables)

    def chain(prod):
        d, k = prod(False), prod(True)
        cc = dict(c)
        cc["tauref"], cc["taurefcld"] = d["tauref"], k["tauref"]
        for q in ("qxv", "qsv", "gv", "qxi", "qsi", "gi"):
            cc[q[:-1] + q[-1] + "_dst"] = d[q]
            cc[q[:-1] + q[-1] + "_cld"] = k[q]
        cc["qextref_dst"], cc["qextref_cld"] = d["qextref"], k["qextref"]
        return cc

    cg = chain(lambda cloud: rt.opt_aerosol_host(cloud, c["plev"], qtr))
    co = chain(lambda cloud: oracle.opt_aer(cloud, v["radb_rt"], v["q_v_" + ("cld" if cloud else "dst")],
                                            v["q_i_" + ("cld" if cloud else "dst")], qtr, c["plev"]))
    got = rt.run_host(cg)
    # the radiation on identical inputs (the device producers' outputs): the plain bar
    errs = compare(got, oracle.run(t, cg))
    print("radiation on the device producers' outputs", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
    # the whole chain: fluxes at the plain bar; the heating, a difference of net fluxes,
    # within 4x the oracle's own response to the producers' ulp-level differences
    want = oracle.run(t, co)
    flux_keys = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi", "diffvt")
    compare(got, want, keys=flux_keys)
    own = rel_err(oracle.run(t, cg)["hr"], want["hr"])
    print(f"chain heating {rel_err(got['hr'], want['hr']):.1e}; oracle's own response {own:.1e}", flush=True)
    assert rel_err(got["hr"], want["hr"]) <= max(1e-10, 4.0 * own)


def test_host_entry_points_equal_device_entry_points(tables):
    """mgcm_rt_opt_dst_host / _cld_host (host arrays, library-managed copies) give the
    device entry points' results bit for bit, at 24 and 99 layers."""
    import torch
    from mgcm_rt import AER_OUT_KEYS, Radiation, synth
    for nlay, ncol in ((24, 300), (99, 37)):
        rt = Radiation(synth.make_tables(1, nlay))
        rt.set_aerosol_tables(tables)
        qtr = synth.make_tracers(7 + nlay, ncol, nlay)
        pl = synth.make_columns(7 + nlay, ncol, nlay)["plev"]
        ld = 2 * nlay + 4
        shapes = {"qxv": (ncol, 7, ld), "qsv": (ncol, 7, ld), "gv": (ncol, 7, ld), "qxi": (ncol, 5, ld),
                  "qsi": (ncol, 5, ld), "gi": (ncol, 5, ld), "qextref": (ncol, ld), "tauref": (ncol, ld),
                  "tau2": (ncol, 2)}
        for cloud in (False, True):
            host = rt.opt_aerosol_host(cloud, pl, qtr)
            dev = {k: torch.empty(v, dtype=torch.float64, device="cuda") for k, v in shapes.items()}
            rt.opt_aerosol(cloud, torch.from_numpy(pl).cuda(), torch.from_numpy(qtr).cuda(), dev)
            torch.cuda.synchronize()
            for k in AER_OUT_KEYS:
                assert np.array_equal(host[k], dev[k].cpu().numpy()), (nlay, cloud, k)
