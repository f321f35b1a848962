"""The algorithmic op counts that bench.py prices its roofline with: recounted by
the oracle's op counter on the same sample (the first 1,024 columns of each
tabulated workload), so the committed table cannot drift from the restatement."""
from __future__ import annotations

import importlib.util
import os

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_ops_per_column_table_matches_the_op_counter(oracle, bench):
    from mgcm_rt import synth
    assert bench.OPS_PER_COLUMN
    for (seed, nlay), want in bench.OPS_PER_COLUMN.items():
        t, c = synth.make_workload(seed, 64800, nlay)  # bench.py's default grid
        sub = {k: (v[:1024] if isinstance(v, np.ndarray) else v) for k, v in c.items()}
        sub["ncol"] = 1024
        got = oracle.count(t, sub)
        for k, v in want.items():
            assert got[k] / 1024 == v, (seed, nlay, k, got[k] / 1024, v)
