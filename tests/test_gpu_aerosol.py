"""Per-column, per-level aerosol optics on the GPU (the mgcm_rt_aerosol block of
mgcm_rt_columns): what OPT_DST and OPT_CLD hand to OPTCV/OPTCI for active dust and
water-ice clouds (code/comp3.f:595-631)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import GOLDEN
from test_gpu_plain import compare

pytestmark = pytest.mark.gpu
IN_KEYS = ("plev", "tlev", "pmid", "tmid", "qh2o", "tauref", "taurefcld", "als", "albi", "acosz")


def test_per_level_aerosol_matches_reference_fixture():
    from mgcm_rt import Radiation, synth
    f = np.load(os.path.join(GOLDEN, "column_aer24.npz"))
    seed = int(f["seed"])
    t, c = synth.make_workload(seed, 32, 24, clouds=True, weights="reference")
    for k in IN_KEYS:
        assert np.array_equal(c[k], f["in_" + k]), k
    got = Radiation(t).run_host(c | synth.make_aerosol(seed, 32, 24))
    errs = compare(got, {k: f["out_" + k] for k in got})
    print("per-level aerosol vs reference", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)


@pytest.mark.parametrize("nlay,dust,cloud,seed", [(36, True, True, 41), (24, True, False, 42), (16, False, True, 43)])
def test_per_level_aerosol_matches_oracle(oracle, nlay, dust, cloud, seed):
    """Dust and cloud blocks independently optional (the other from the tables)."""
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(seed, 300, nlay, clouds=True)
    ca = c | synth.make_aerosol(seed, 300, nlay, dust=dust, cloud=cloud)
    got = Radiation(t).run_host(ca)
    errs = compare(got, oracle.run(t, ca))
    print(nlay, dust, cloud, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)


def test_constant_aerosol_block_equals_tables():
    """The tables' constants passed as per-level arrays give the same bits as no block."""
    from mgcm_rt import AER_KEYS, Radiation, synth
    t, c = synth.make_workload(44, 200, 24, clouds=True)
    ld = 2 * 24 + 4
    src = {"qxv_dst": "qxvdst", "qsv_dst": "qsvdst", "gv_dst": "gvdst", "qxi_dst": "qxidst", "qsi_dst": "qsidst",
           "gi_dst": "gidst", "qxv_cld": "qxvcld", "qsv_cld": "qsvcld", "gv_cld": "gvcld", "qxi_cld": "qxicld",
           "qsi_cld": "qsicld", "gi_cld": "gicld"}
    aer = {k: np.ascontiguousarray(np.broadcast_to(np.asarray(t[v])[None, :, :ld], (200,) + np.asarray(t[v])[:, :ld].shape))
           for k, v in src.items()}
    aer["qextref_dst"] = np.full((200, ld), float(np.asarray(t["qextrefdst"])[0]))
    aer["qextref_cld"] = np.full((200, ld), float(np.asarray(t["qextrefcld"])[0]))
    assert set(aer) == set(AER_KEYS)
    rt = Radiation(t)
    a, b = rt.run_host(c), rt.run_host(c | aer)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_batched_grids_with_and_without_aerosol():
    """A batch mixing grids with and without per-level aerosol equals the single-grid runs."""
    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    t, _ = synth.make_workload(45, 8, 24)
    rt = Radiation(t)
    grids, outs, singles = [], [], []
    for i, (n, nl, with_aer) in enumerate([(300, 24, True), (257, 24, False), (129, 24, True)]):
        c = synth.make_columns(450 + i, n, nl, clouds=True)
        if with_aer:
            c = c | synth.make_aerosol(450 + i, n, nl)
        grids.append(to_device(c))
        outs.append(to_device(alloc_outputs(n, nl)))
        singles.append(rt.run_host(c))
    rt.run_batched(grids, outs, FP64)
    torch.cuda.synchronize()
    for o, s in zip(outs, singles):
        for k in s:
            assert np.array_equal(o[k].cpu().numpy(), s[k]), k
