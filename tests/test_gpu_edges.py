"""Edge cases of the plain FP64 path on the GPU, through the C ABI:
the reference's own edge fixture (sun down at 0 and 9.99999e-6, grazing sun at
1e-5, H2O above the water grid, T at the ends of the k-table and Planck
ranges, a TLIMIT-skipped band, thick dust, cloud); the smallest and largest
layer counts; one-column, empty and non-multiple-of-4 grids; the range flag;
determinism; status codes."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu

KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
        "diffvt", "detau", "cumtauv", "hr")


def _fixture(name):
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    from test_oracle import load_fixture
    return load_fixture(name)


def test_reference_edge_fixture():
    from mgcm_rt import Radiation
    t, c, want = _fixture("edge24.npz")
    got = Radiation(t).run_host(c)
    for k in KEYS:
        assert rel_err(got[k], want[k]) <= 1e-10, (k, rel_err(got[k], want[k]))
    down = c["acosz"] < np.float32(1e-5)
    assert down[0] and down[1] and not down[5]
    assert not np.any(got["detau"][down]) and not np.any(got["fmnetv"][down])


@pytest.mark.parametrize("nlay,ncol", [(1, 37), (16, 5), (64, 9), (24, 1), (36, 2), (65, 6), (80, 3), (99, 11)])
def test_layer_and_column_extremes(oracle, nlay, ncol):
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(60 + nlay, ncol, nlay, clouds=(nlay == 36))
    got = Radiation(t).run_host(c)
    want = oracle.run(t, c)
    for k in KEYS:
        assert rel_err(got[k], want[k]) <= 1e-10, (nlay, ncol, k)


def test_tpindex_pressure_cells_on_grid_points(oracle):
    """TPINDEX's pressure cell (code/tpindex.f90:93-106) where a sub-level's log10(PMID)
    equals a grid value: PFGASREF moved onto the log10 (host libm, which the device's
    equals bit for bit) of sub-level pressures of the grid, ascending -- the kernel
    counts the points the pressure fails -- and with the point after one of them set a
    hair below it, so the grid no longer ascends and the kernel scans as the reference
    does.  Against the oracle's scan at rel 1e-10."""
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(75, 64, 24)
    pf = np.array(t["pfgasref"], dtype=np.float64)
    _, lg, _ = oracle.libm(np.ascontiguousarray(c["pmid"][:, 1:]).ravel())
    lg = np.unique(lg)
    placed = []
    for i in range(1, len(pf) - 1):  # PFGASREF(2..L_PINT-1), the scanned points
        inside = lg[(lg > pf[i - 1]) & (lg < pf[i + 1])]
        if inside.size:
            pf[i] = inside[np.argmin(np.abs(inside - pf[i]))]
            placed.append(i)
    assert len(placed) >= 5 and np.all(np.diff(pf) > 0)
    i = next(j for j in placed[len(placed) // 2:] if j + 1 < len(pf) - 1)
    down = pf.copy()
    down[i + 1] = pf[i] - 2.0 ** -40
    for label, grid in (("ascending", pf), ("not ascending", down)):
        tt = dict(t, pfgasref=grid)
        got = Radiation(tt).run_host(c)
        want = oracle.run(tt, c)
        print(label, "max rel", f"{max(rel_err(got[k], want[k]) for k in KEYS):.1e}", flush=True)
        for k in KEYS:
            assert rel_err(got[k], want[k]) <= 1e-10, (label, k)


def test_empty_grid():
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(70, 3, 24)
    empty = {k: (v[:0] if isinstance(v, np.ndarray) else v) for k, v in c.items()}
    empty["ncol"] = 0
    out = Radiation(t).run_host(empty)
    assert out["hr"].shape == (0, 25)


def test_range_flag_reported():
    from mgcm_rt import Radiation, RadiationError, synth
    t, c = synth.make_workload(71, 8, 24)
    rt = Radiation(t)
    c["tlev"][3, 10] = 950.0          # above the Planck table (50..900 K)
    with pytest.raises(RadiationError) as e:
        rt.run_host(c)
    assert e.value.status == -5
    t2, c2 = synth.make_workload(72, 8, 24)
    rt.run_host(c2)                    # the flag was cleared by the check
    c2["tlev"][5, 4] = np.nan
    with pytest.raises(RadiationError) as e:
        rt.run_host(c2)
    assert e.value.status == -5


@pytest.mark.parametrize("key,where", [("plev", (2, 7)), ("tlev", (1, 30)), ("pmid", (3, 12)), ("tmid", (4, 40)),
                                       ("qh2o", (5, 20)), ("tauref", (6, 0)), ("tauref", (6, 51)),
                                       ("taurefcld", (0, 33)), ("taurefcld", (7, 51)), ("als", (1,)),
                                       ("albi", (2,)), ("acosz", (3,))])
def test_nan_input_sets_range_flag(key, where):
    """A NaN in any column input the path reads -- every level of PLEV, TLEV, PMID, TMID,
    QH2O it uses, TAUREF/TAUREFCLD including levels 1 and L_LEVELS+1, and the column's
    ALS/ALBI/ACOSZ -- is reported as MGCM_RT_ERR_RANGE; the next clean call is OK."""
    from mgcm_rt import Radiation, RadiationError, synth
    t, c = synth.make_workload(73, 8, 24)
    rt = Radiation(t)
    bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.items()}
    bad[key][where] = np.nan
    with pytest.raises(RadiationError) as e:
        rt.run_host(bad)
    assert e.value.status == -5
    rt.run_host(c)


def test_range_flag_of_an_unchecked_device_run_is_kept():
    """A range error of an asynchronous mgcm_rt_run that no check has read yet is reported
    by the next check -- here the one mgcm_rt_run_host makes -- and then cleared;
    mgcm_rt_check_stream reads it waiting on the run's stream only."""
    import torch
    from mgcm_rt import FP64, Radiation, RadiationError, alloc_outputs, synth, to_device
    t, c = synth.make_workload(74, 16, 24)
    rt = Radiation(t)
    bad = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in c.items()}
    bad["tlev"][3, 10] = 950.0
    s = torch.cuda.Stream()
    dc, do = to_device(bad), to_device(alloc_outputs(16, 24))
    rt.run_device(dc, do, FP64, s.cuda_stream)   # no check: the flag stays pending
    s.synchronize()
    with pytest.raises(RadiationError) as e:
        rt.run_host(c)                            # clean call, pending flag reported
    assert e.value.status == -5
    rt.run_host(c)                                # cleared by that check
    rt.run_device(dc, do, FP64, s.cuda_stream)
    with pytest.raises(RadiationError) as e:
        rt.check_stream(s.cuda_stream)
    assert e.value.status == -5
    rt.check_stream(s.cuda_stream)


def test_deterministic():
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(73, 101, 24)
    rt = Radiation(t)
    a, b = rt.run_host(c), rt.run_host(c)
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


def test_status_codes():
    from mgcm_rt import Radiation, RadiationError, synth
    t, c = synth.make_workload(74, 4, 24)
    rt = Radiation(t)
    for prec in (7, 0x30, 0x40):
        with pytest.raises(RadiationError) as e:
            rt.run_host(c, precision=prec)
        assert e.value.status == -6
    for nl in (0, 100):
        bad = dict(c)
        bad["nlay"] = nl
        with pytest.raises(RadiationError) as e:
            rt.run_host(bad)
        assert e.value.status == -2
    assert rt.lib.mgcm_rt_run(rt.ctx, None, None, 0, None) == -1
    assert b"Planck" in rt.lib.mgcm_rt_strerror(-5)


def test_two_million_column_grid_is_a_tiled_copy(oracle):
    """2,097,152 columns (32 copies of a 65,536-column grid) in one call: every
    copy's outputs equal the first copy's bit for bit (no index wraps or
    overlaps at this size), and sampled columns of the first copy equal the
    oracle.  About 23 GB of device memory."""
    import torch
    from mgcm_rt import Radiation, alloc_outputs, synth, to_device
    base_n, reps, nlay = 65536, 32, 24
    t, c = synth.make_workload(77, base_n, nlay)
    c["acosz"][::9] = 0.0
    rt = Radiation(t)
    db = to_device(c)
    big = {k: (v.repeat((reps,) + (1,) * (v.dim() - 1)) if isinstance(v, torch.Tensor) else v) for k, v in db.items()}
    big["ncol"] = base_n * reps
    del db
    proto = alloc_outputs(1, nlay)
    # -1 marks unwritten outputs; DETAU/CUMTAUV of sun-down columns stay as given (0, the oracle's)
    out = {k: torch.full((base_n * reps,) + v.shape[1:], 0.0 if k in ("detau", "cumtauv") else -1.0,
                         dtype=torch.float64, device="cuda") for k, v in proto.items()}
    rt.run_device(big, out)
    torch.cuda.synchronize()
    rt.check()
    for k, v in out.items():
        tiles = v.view((reps, base_n) + v.shape[1:])
        assert bool(torch.all(tiles == tiles[:1])), k
    idx = np.r_[0:8, base_n - 8:base_n, np.random.default_rng(3).choice(base_n, 48, replace=False)]
    sub = {k: (v[idx] if isinstance(v, np.ndarray) and v.shape[:1] == (base_n,) else v) for k, v in c.items()}
    sub["ncol"] = len(idx)
    want = oracle.run(t, sub)
    for k in KEYS:
        got = out[k][(reps - 1) * base_n:][idx].cpu().numpy()  # the last copy
        assert rel_err(got, want[k]) <= 1e-10, k
