"""The GCM built with COMP3's radiation calls moved into glue routines
(oracle/build_model.sh): the twin whose glue runs the reference's own OPTCV/SFLUXV
and OPTCI/SFLUXI behaves exactly as the unmodified GCM -- the rewrite the MI355X
glue relies on changes nothing.  CPU only; skipped where the model builds are absent."""
from __future__ import annotations

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
