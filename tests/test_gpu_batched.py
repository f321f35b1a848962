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
