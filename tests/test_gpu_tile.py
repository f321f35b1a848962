"""The column kernel's tiles -- 1, 2 or 4 band slots per column, i.e. 4, 2 or 1
columns per wave, and 12: one band per wave, 4 columns per wave, the bands'
sums added by a finishing kernel -- through the C ABI.  Every tile gives the
same bits on every output: grids whose column count is no multiple of 4, the reference's edge
fixture (sun down and grazing, a TLIMIT-skipped band, thick dust, cloud), one
part only (VIS or IR), both precisions, the smallest and largest layer
counts, single-grid and batched calls; the oracle within 1e-10; the automatic
choice per grid size."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu

KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
        "diffvt", "detau", "cumtauv", "hr")
TILES = (1, 2, 4, 12)
BATCH_TILES = (1, 2, 4)  # the batched path's tiles; 12 is for single-grid calls


def _per_tile(rt, c, precision=0):
    out = {}
    for s in TILES:
        rt.set_tile(s)
        out[s] = rt.run_host(c, precision)
    rt.set_tile(0)
    return out


def _same(out, keys=KEYS, label=""):
    for s in TILES[1:]:
        for k in keys:
            assert np.array_equal(out[s][k], out[1][k]), (label, s, k)


@pytest.mark.parametrize("nlay,ncol", [(24, 37), (16, 5), (36, 2), (99, 11), (65, 6), (24, 1), (1, 3), (48, 10)])
def test_tiles_bit_identical_and_oracle(oracle, nlay, ncol):
    from mgcm_rt import Radiation, synth
    print(f"tiles nlay={nlay} ncol={ncol}", flush=True)
    t, c = synth.make_workload(300 + nlay, ncol, nlay, clouds=(nlay in (36, 48)))
    out = _per_tile(Radiation(t), c)
    _same(out, label=(nlay, ncol))
    want = oracle.run(t, c)
    for k in KEYS:
        assert rel_err(out[4][k], want[k]) <= 1e-10, (nlay, ncol, k)


def test_tiles_edge_fixture():
    from mgcm_rt import Radiation
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)
    from test_oracle import load_fixture
    t, c, want = load_fixture("edge24.npz")
    out = _per_tile(Radiation(t), c)
    _same(out, label="edge24")
    for k in KEYS:
        assert rel_err(out[2][k], want[k]) <= 1e-10, k


@pytest.mark.parametrize("part", ["vis", "ir"])
def test_tiles_one_part(part):
    from mgcm_rt import SKIP_IR, SKIP_VIS, Radiation, synth
    t, c = synth.make_workload(311, 23, 24)
    out = _per_tile(Radiation(t), c, SKIP_IR if part == "vis" else SKIP_VIS)
    _same(out, label=part)


def test_tiles_mixed_precision():
    from mgcm_rt import MIXED, Radiation, synth
    t, c = synth.make_workload(312, 29, 36, clouds=True)
    out = _per_tile(Radiation(t), c, MIXED)
    _same(out, label="mixed")


def test_tiles_sun_down_everywhere():
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(313, 9, 24)
    c["acosz"][:] = 0.0
    out = _per_tile(Radiation(t), c)
    _same(out, label="night")
    assert not np.any(out[4]["fmnetv"]) and np.any(out[4]["fmneti"])


def test_auto_tile_choice():
    from mgcm_rt import Radiation, RadiationError, synth
    t, _ = synth.make_workload(314, 1, 24)
    rt = Radiation(t)
    assert rt.tile_for(64800, 24) == 1
    assert rt.tile_for(540, 24) == 12
    assert rt.tile_for(2160, 24) == 12  # the reference's 36 x 60 grid
    picks = [rt.tile_for(n, 24) for n in (100, 540, 1080, 2160, 4320, 8100, 16200, 64800)]
    assert picks == sorted(picks, reverse=True), picks  # fewer slots as grids grow
    rt.set_tile(2)
    assert rt.tile_for(64800, 24) == 2
    rt.set_tile(12)
    assert rt.tile_for(64800, 24) == 12
    with pytest.raises(RadiationError) as e:
        rt.set_tile(3)
    assert e.value.status == -1


@pytest.mark.parametrize("tile", BATCH_TILES)
def test_batched_pinned_tile(tile):
    """The batched kernels of each tile (the tuner restricted to it) against
    single-grid runs, bit for bit, on a ragged batch with partial 4-column items."""
    import torch
    from mgcm_rt import Radiation, alloc_outputs, synth, to_device
    shapes = [(37, 16), (1, 20), (101, 24), (6, 33), (23, 64), (7, 99)]
    t = synth.make_tables(315, 24)
    cols = [synth.make_columns(3150 + i, n, nlay) for i, (n, nlay) in enumerate(shapes)]
    rt = Radiation(t)
    rt.set_tile(tile)
    dc = [to_device(c) for c in cols]
    do = [to_device(alloc_outputs(n, nlay)) for n, nlay in shapes]
    rt.run_batched(dc, do, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rt.check()
    assert all(c["tile"] == tile for c in rt.batched_tuning())
    rt.set_tile(1)
    for (n, nlay), c, o in zip(shapes, cols, do):
        single = rt.run_host(c)
        for k in KEYS:
            assert np.array_equal(o[k].cpu().numpy(), single[k]), (tile, n, nlay, k)


def test_batched_with_one_band_tile_pinned():
    """Pinning the one-band tile (single-grid calls only) leaves the batched tuner its own
    choice among its tiles; the batch still equals the single-grid runs bit for bit."""
    import torch
    from mgcm_rt import Radiation, alloc_outputs, synth, to_device
    shapes = [(37, 16), (101, 24), (23, 36)]
    t = synth.make_tables(316, 24)
    cols = [synth.make_columns(3160 + i, n, nlay) for i, (n, nlay) in enumerate(shapes)]
    rt = Radiation(t)
    rt.set_tile(12)
    dc = [to_device(c) for c in cols]
    do = [to_device(alloc_outputs(n, nlay)) for n, nlay in shapes]
    rt.run_batched(dc, do, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rt.check()
    assert all(c["tile"] in BATCH_TILES for c in rt.batched_tuning())
    for (n, nlay), c, o in zip(shapes, cols, do):
        single = rt.run_host(c)  # at tile 12
        for k in KEYS:
            assert np.array_equal(o[k].cpu().numpy(), single[k]), (n, nlay, k)


def _reference_built() -> bool:
    import os
    from oracle import REF_SO
    return os.path.exists(REF_SO)


@pytest.mark.skipif(not _reference_built(), reason="oracle/_ref not built (needs the reference's sources)")
def test_reference_grid_every_tile_against_the_reference_fortran():
    """The reference's own 36 x 60 grid (2,160 columns, 24 layers, its Gauss weights) at
    every tile against the reference's Fortran on the same columns: rel 1e-10 everywhere,
    and the tiles bit-identical."""
    from oracle import Reference
    from mgcm_rt import Radiation, synth
    t = synth.make_tables(36, 24, weights="reference")
    c = synth.make_columns(36, 2160, 24)
    want = Reference().run_grid(t, c)
    rt = Radiation(t)
    outs = []
    for tile in TILES:
        rt.set_tile(tile)
        got = rt.run_host(c)
        worst = max(rel_err(got[k], want[k]) for k in want if k in got)
        print("tile", tile, f"{worst:.1e}", flush=True)
        assert worst <= 1e-10, (tile, worst)
        outs.append(got)
    for k in outs[0]:
        assert all(np.array_equal(outs[0][k], o[k]) for o in outs[1:]), k
