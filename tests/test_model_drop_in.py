"""The GCM built with COMP3's radiation calls moved into glue routines
(oracle/build_model.sh): the twin whose glue runs the reference's own OPTCV/SFLUXV
and OPTCI/SFLUXI behaves exactly as the unmodified GCM -- the rewrite the MI355X
glue relies on changes nothing.  The same holds for COMP3's column loop split in
two passes around one radiation batch (fortran/comp3_split.f90), the batch run
by the reference's routines.  CPU only; skipped where the model builds are absent."""
from __future__ import annotations

import numpy as np
import pytest

import model_run


@pytest.mark.skipif(not model_run.built("gcm_ref", "gcm_refglue"), reason="GCM builds not present (oracle/build_model.sh)")
def test_rewritten_comp3_equals_unmodified_gcm(tmp_path):
    a = model_run.run_model("gcm_ref", str(tmp_path / "ref"), dump=False)
    b = model_run.run_model("gcm_refglue", str(tmp_path / "glue"))
    assert len(a["steps"]) == len(b["steps"]) == 3
    diffs = model_run.history_diff(a, b)
    bad = [(t, k, v) for t, row in enumerate(diffs) for k, v in row.items() if v != 0.0]
    assert not bad, bad[:10]
    passes = model_run.split_passes(b["dump"])
    print(len(passes), "radiation passes,", [len(p) for p in passes], "calls", flush=True)
    assert len(passes) >= 3


@pytest.mark.skipif(not model_run.built("gcm_ref", "gcm_refglue", "gcm_refsplit"),
                    reason="GCM builds not present (oracle/build_model.sh)")
def test_two_pass_comp3_equals_unmodified_gcm(tmp_path):
    """The two-pass COMP3: pass 1 runs every column's part before the radiation and
    saves its state, one batch radiates every column, pass 2 restores each column and
    runs the rest (its radiation call sites check, bit for bit, that they get the
    inputs pass 1 recorded).  With the reference's routines in the batch the model
    is the unmodified GCM bit for bit: history and every radiation call."""
    a = model_run.run_model("gcm_ref", str(tmp_path / "ref"), dump=False)
    g = model_run.run_model("gcm_refglue", str(tmp_path / "glue"))
    b = model_run.run_model("gcm_refsplit", str(tmp_path / "split"))
    diffs = model_run.history_diff(a, b)
    bad = [(t, k, v) for t, row in enumerate(diffs) for k, v in row.items() if v != 0.0]
    assert not bad, bad[:10]
    assert len(g["dump"]) == len(b["dump"]) > 10000
    for x, y in zip(g["dump"], b["dump"]):
        assert x[:3] == y[:3] and np.array_equal(x[3], y[3]), (x[:3], y[:3])
