"""The in-model drop-in: the whole GCM (the tutorial case -- active dust, 36 x 60
grid, 24 layers -- for one model hour: 30 physics steps, 5 COMP3 radiation passes of
2,100 columns) with COMP3's radiation calls served by the library on the MI355X
(fortran/comp3_rt.f90, via mgcm_rt_mod), against the same GCM on the reference's
own radiation routines (oracle/build_model.sh).  Every radiation call of the first
pass -- both runs then start from the same state -- matches to rel 1e-10: fluxes,
DETAU/CUMTAUV (DSOLFLUX's inputs), TEMPGR's DNVFLUX/DNDIFFV/DNIRFLUX and the net-flux
differences QRAD is made of.  Later passes start from states that differ by the
first pass's rounding; their drift, and the history files' drift after the hour,
are measured against control runs of the reference GCM whose radiation fluxes are
moved by random one-ulp amounts: the GCM's own amplification of rounding-level
differences, which the drift must not exceed tenfold."""
from __future__ import annotations

import numpy as np
import pytest

import model_run

pytestmark = pytest.mark.gpu


def record_errors(ref: list, got: list) -> dict:
    """Per quantity: max over the pass of |got - ref| / max(|ref|, the call's scale)
    with the scale the record's largest |net flux| (fluxes, differences) or the
    quantity's largest |value| (optical depths)."""
    assert [r[:3] for r in ref] == [g[:3] for g in got]
    worst = {}
    for r, g in zip(ref, got):
        a, b = model_run.groups(r[2], r[3]), model_run.groups(g[2], g[3])
        sflux = max(float(np.max(np.abs(a["fmnet"]))), float(np.max(np.abs(a["fluxup"]))), 1e-300)
        for k in a:
            s = sflux if k not in ("cumtauv", "detau") else max(float(np.max(np.abs(a[k]))), 1e-300)
            e = float(np.max(np.abs(b[k] - a[k]) / np.maximum(np.abs(a[k]), s)))
            worst[k] = max(worst.get(k, 0.0), e)
    return worst


@pytest.mark.skipif(not model_run.built("gcm_refglue", "gcm_mi355x"),
                    reason="GCM builds not present (oracle/build_model.sh)")
def test_gcm_with_mi355x_radiation(tmp_path, monkeypatch):
    ref = model_run.run_model("gcm_refglue", str(tmp_path / "ref"))
    got = model_run.run_model("gcm_mi355x", str(tmp_path / "mi355x"), timeout=600)
    # controls: the reference GCM with every radiation flux moved by a random -1/0/+1 ulp
    ctls = []
    for seed in (1, 2):
        monkeypatch.setenv("MGCM_COMP3_PERTURB", f"random:{seed}")
        ctls.append(model_run.run_model("gcm_refglue", str(tmp_path / f"control{seed}")))
    monkeypatch.delenv("MGCM_COMP3_PERTURB")
    pr, pg = model_run.split_passes(ref["dump"]), model_run.split_passes(got["dump"])
    assert len(pr) == len(pg) == 5, (len(pr), len(pg))
    ours = [max(record_errors(a, b).values()) for a, b in zip(pr, pg)]
    ctl = [max(max(record_errors(a, b).values()) for b in (model_run.split_passes(c["dump"])[n] for c in ctls))
           for n, a in enumerate(pr)]
    for n in range(len(pr)):
        e = record_errors(pr[n], pg[n])
        print(f"radiation pass {n + 1} ({len(pr[n])} calls): max {ours[n]:.1e} (controls {ctl[n]:.1e})",
              {k: f"{v:.1e}" for k, v in e.items()}, flush=True)
    assert ours[0] <= 1e-10, ours[0]
    for n in range(1, len(pr)):
        assert ours[n] <= 10.0 * ctl[n], (n + 1, ours[n], ctl[n])
    drift = model_run.history_diff(ref, got)
    cdrift = [model_run.history_diff(ref, c) for c in ctls]
    for t, row in enumerate(drift):
        cmax = max(max(cd[t].values()) for cd in cdrift)
        print(f"history output {t}: max rel diff {max(row.values()):.1e} (controls {cmax:.1e})",
              {k: f"{v:.1e}" for k, v in row.items() if v > 0}, flush=True)
        assert max(row.values()) <= 10.0 * cmax + 1e-7, (t, max(row.values()), cmax)


@pytest.mark.skipif(not model_run.built("gcm_mi355x", "gcm_mi355x_split"),
                    reason="GCM builds not present (oracle/build_model.sh)")
def test_gcm_two_pass_batched_equals_per_column(tmp_path):
    """COMP3's column loop in two passes around ONE mgcm_rt_radiation call per
    radiation step (all 2,100 columns, VIS and IR together) gives the GCM of the
    per-column glue (two library calls per column) bit for bit -- every radiation
    call's outputs and the history after the hour: columns are independent, and the
    library's tiles and one-part runs give the same bits.  The wall times of both
    forms and of the unmodified reference GCM (one host core) are printed."""
    import time
    t0 = time.perf_counter()
    per = model_run.run_model("gcm_mi355x", str(tmp_path / "percol"), timeout=600)
    t1 = time.perf_counter()
    bat = model_run.run_model("gcm_mi355x_split", str(tmp_path / "batch"), timeout=600)
    t2 = time.perf_counter()
    model_run.run_model("gcm_ref", str(tmp_path / "ref"), dump=False)
    t3 = time.perf_counter()
    print(f"one model hour: per-column glue {t1 - t0:.2f} s, two-pass batch {t2 - t1:.2f} s, "
          f"the reference GCM on one host core {t3 - t2:.2f} s", flush=True)
    assert len(per["dump"]) == len(bat["dump"]) > 10000
    for x, y in zip(per["dump"], bat["dump"]):
        assert x[:3] == y[:3] and np.array_equal(x[3], y[3]), (x[:3], y[:3])
    diffs = model_run.history_diff(per, bat)
    bad = [(t, k, v) for t, row in enumerate(diffs) for k, v in row.items() if v != 0.0]
    assert not bad, bad[:10]
