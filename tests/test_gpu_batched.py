"""Ragged batches through mgcm_rt_run_batched (C ABI): every grid equals its own
single-grid mgcm_rt_run bit for bit and the oracle within 1e-10; the cached
plan and HIP graph are replayed for a batch of the same shape with new data and
rebuilt for a new shape; empty grids and one-column grids are accepted."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
        "diffvt", "detau", "cumtauv", "hr")
SHAPES = [(37, 16), (1, 20), (100, 24), (0, 30), (64, 33), (5, 48), (23, 64), (9, 24), (7, 99), (12, 65), (3, 81)]


def _batch(seed, shapes, weights="dirichlet"):
    from mgcm_rt import synth
    t = synth.make_tables(seed, 24, weights=weights)
    cols = [synth.make_columns(seed * 100 + i, n, nlay) if n else None for i, (n, nlay) in enumerate(shapes)]
    return t, cols


def _empty(nlay):
    from mgcm_rt import synth
    c = synth.make_columns(1, 1, nlay)
    return {k: (v[:0] if isinstance(v, np.ndarray) else v) for k, v in c.items()} | {"ncol": 0}


def _run(rt, cols, shapes, precision=0):
    import torch
    from mgcm_rt import alloc_outputs, to_device
    dc = [to_device(c if c is not None else _empty(nlay)) for c, (_, nlay) in zip(cols, shapes)]
    do = [to_device(alloc_outputs(n, nlay)) for n, nlay in shapes]
    rt.run_batched(dc, do, precision, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rt.check()
    return [{k: v.cpu().numpy() for k, v in o.items()} for o in do]


def test_batched_matches_single_and_oracle(oracle):
    from mgcm_rt import Radiation
    t, cols = _batch(41, SHAPES)
    rt = Radiation(t)
    outs = _run(rt, cols, SHAPES)
    for (n, nlay), c, o in zip(SHAPES, cols, outs):
        if not n:
            continue
        single = rt.run_host(c)
        for k in KEYS:
            assert np.array_equal(o[k], single[k]), (n, nlay, k)
        from mgcm_rt import synth
        want = oracle.run(synth.make_tables(41, nlay), c)  # same k-tables, per-level arrays for nlay
        for k in KEYS:
            assert rel_err(o[k], want[k]) <= 1e-10, (n, nlay, k)


def test_batched_graph_replay_and_replan():
    from mgcm_rt import Radiation
    t, cols = _batch(42, SHAPES)
    _, cols2 = _batch(43, SHAPES)
    rt = Radiation(t)
    first = _run(rt, cols, SHAPES)
    second = _run(rt, cols2, SHAPES)   # same shape: plan and graph replayed, new pointers
    again = _run(rt, cols, SHAPES)
    for a, b in zip(first, again):
        for k in KEYS:
            assert np.array_equal(a[k], b[k])
    for (n, _), c2, o2 in zip(SHAPES, cols2, second):
        if n:
            single = rt.run_host(c2)
            assert np.array_equal(o2["fmnetv"], single["fmnetv"])
    shapes3 = [(11, 40), (3, 17)]      # new shape: re-planned
    _, cols3 = _batch(44, shapes3)
    o3 = _run(rt, cols3, shapes3)
    for c, o in zip(cols3, o3):
        assert np.array_equal(o["hr"], rt.run_host(c)["hr"])


def test_batched_rejects_bad_layer_count():
    from mgcm_rt import Radiation, RadiationError
    t, cols = _batch(45, [(4, 24)])
    rt = Radiation(t)
    bad = dict(cols[0])
    bad["nlay"] = 100
    with pytest.raises(RadiationError) as e:
        _run(rt, [bad], [(4, 24)])
    assert e.value.status == -2


# Per-class work lists longer than the resident capacity: the <= 24-layer class
# has 50,000 columns, 12,500 4-column items -- more than 4x the 2,048 one-wave
# blocks that two waves per SIMD keep resident on 256 CUs, so every persistent
# block of every tuner candidate walks several items -- and 5,556 VIS (9-column)
# and 4,167 IR (12-column) continuum-point items, more than the 4,096 blocks of
# those launches.  The 36- and 60-layer classes exceed one resident round too.
AT_SHAPE = [(25000, 24), (25000, 20), (8000, 36), (5000, 60), (3000, 16)]


def test_batched_at_shape_multi_item(oracle):
    """The persistent blocks' multi-item loop (staging reuse, the per-item barrier) and
    the tuner's candidates under a check: every grid equals its own single-grid run bit
    for bit, sampled columns match the oracle within 1e-10, and the DSOLFLUX/DNVFLUX of
    the grids that ask for them (inside the batch's graph) equal mgcm_rt_dsolflux's."""
    import torch
    from mgcm_rt import Radiation, alloc_outputs, synth, to_device
    t, cols = _batch(46, AT_SHAPE)
    rt = Radiation(t)
    dc = [to_device(c) for c in cols]
    do = [to_device(alloc_outputs(n, nlay)) for n, nlay in AT_SHAPE]
    for i in (0, 2, 3):  # surface fluxes for some grids only
        do[i]["directsol"] = torch.full((AT_SHAPE[i][0],), -1.0, dtype=torch.float64, device="cuda")
        do[i]["dnvflux"] = torch.full_like(do[i]["directsol"], -1.0)
    s = torch.cuda.current_stream().cuda_stream
    rt.run_batched(dc, do, 0, s)
    torch.cuda.synchronize()
    rt.check()
    tune = rt.batched_tuning()
    print(tune, flush=True)
    c24 = [c for c in tune if c["max_layers"] == 24][0]
    assert c24["items"] == 12500 and c24["blocks"] < c24["items"] * c24["tile"], c24  # several items per block
    assert len(c24["candidates"]) >= 8 and c24["tuned_items"] > 0, c24
    assert {c["tile"] for c in c24["candidates"]} == {1, 2, 4}, c24
    rng = np.random.Generator(np.random.PCG64(5))
    for i, ((n, nlay), c, o) in enumerate(zip(AT_SHAPE, cols, do)):
        got = {k: v.cpu().numpy() for k, v in o.items()}
        single = rt.run_host(c)
        for k in KEYS:
            assert np.array_equal(got[k], single[k]), (n, nlay, k)
        if "directsol" in got:
            ref = rt.dsolflux(dc[i], {k: torch.from_numpy(v).cuda() for k, v in single.items()}, s)
            torch.cuda.synchronize()
            assert np.array_equal(got["directsol"], ref["directsol"].cpu().numpy()), (n, nlay)
            assert np.array_equal(got["dnvflux"], ref["dnvflux"].cpu().numpy()), (n, nlay)
        pick = np.sort(rng.choice(n, 48, replace=False))
        sub = {k: (v[pick] if isinstance(v, np.ndarray) and v.shape[:1] == (n,) else v) for k, v in c.items()}
        sub["ncol"] = len(pick)
        want = oracle.run(synth.make_tables(46, nlay), sub)
        for k in KEYS:
            assert rel_err(got[k][pick], want[k]) <= 1e-10, (n, nlay, k)
    # a replay of the same shape (graph, tuned launch shapes) gives the same bits
    first = [{k: v.clone() for k, v in o.items()} for o in do]
    rt.run_batched(dc, do, 0, s)
    torch.cuda.synchronize()
    for a, b in zip(first, do):
        for k in a:
            assert torch.equal(a[k], b[k]), k
