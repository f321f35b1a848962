"""The history file (MHISTH/MHISTV records) and COMP3's radiation diagnostics:
the writer and reader against each other and against the reference's own reader
(tutorial/historymod.f90 + htest.f90, built into oracle/_ref by build_ref.sh),
and the oracle's diagnostics against what the formulas fix."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HTEST = os.path.join(ROOT, "oracle", "_ref", "htest")
DIMS = dict(jm=4, im=3, layers=5, nl=6, ntrace=6)


def _fields(rng, d):
    from mgcm_rt import history as H
    ji = (d["jm"], d["im"])
    f = {n: (rng.random(ji) * 300.0).astype(np.float32) for n in H.RADIATION_FIELDS}
    f.update(p=rng.random(ji) * 8, t=rng.random(ji + (d["layers"],)) * 200, dheat=rng.normal(0, 30, ji + (d["layers"],)),
             qtrace=rng.random(ji + (d["layers"], d["ntrace"])), stemp=rng.random(ji + (d["nl"],)) * 250,
             tausurf=rng.random(ji), tau=float(rng.integers(1, 1000)), ptrop=2e-4, nc3=3, ncycle=7)
    return f


def _write(path, steps, rng):
    from mgcm_rt import history as H
    hdr = dict(DIMS, runnum=12.5, sdepth=np.arange(2 * DIMS["nl"] + 1) * 0.1, grav=3.72,
               topog=rng.normal(0, 3000, (DIMS["jm"], DIMS["im"])),
               npcflag=rng.random((DIMS["jm"], DIMS["im"])) > 0.5)
    with open(path, "wb") as f:
        H.write_header(f, **hdr)
        for s in steps:
            H.write_step(f, DIMS, **s)
    return hdr


def test_history_round_trip(tmp_path):
    from mgcm_rt import history as H
    rng = np.random.default_rng(5)
    steps = [_fields(rng, DIMS) for _ in range(3)]
    hdr = _write(tmp_path / "fort.11", steps, rng)
    head, got = H.read_history(str(tmp_path / "fort.11"))
    assert head["jm"] == 4 and head["im"] == 3 and head["version"] == H.VERSION
    assert head["runnum"] == 12.5 and head["grav"] == np.float32(3.72)
    assert np.array_equal(head["npcflag"], hdr["npcflag"])
    assert np.array_equal(head["topog"], hdr["topog"].astype(np.float32))
    assert len(got) == 3
    for s, g in zip(steps, got):
        for k in H.RADIATION_FIELDS + ("p", "t", "dheat", "qtrace", "stemp", "tausurf"):
            assert np.array_equal(g[k], np.asarray(s[k]).astype(np.float32)), k
        assert g["nc3"] == 3 and g["ncycle"] == 7 and g["tau"] == np.float32(s["tau"])
        assert not np.any(g["u"]) and not np.any(g["geop"])  # absent fields are zeros
    # record sizes: 3 header + 21 per step, each REAL*4 array 4 bytes per value
    n = 0
    with open(tmp_path / "fort.11", "rb") as f:
        while H.read_record(f) is not None:
            n += 1
    assert n == 3 + 21 * 3


def test_history_rejects_bad_shapes_and_truncation(tmp_path):
    from mgcm_rt import history as H
    with open(tmp_path / "h", "wb") as f:
        with pytest.raises(ValueError):
            H.write_step(f, DIMS, fuptopv=np.zeros((3, 4)))
    rng = np.random.default_rng(6)
    _write(tmp_path / "fort.11", [_fields(rng, DIMS)], rng)
    data = (tmp_path / "fort.11").read_bytes()
    (tmp_path / "cut").write_bytes(data[:-10])
    with pytest.raises(ValueError):
        H.read_history(str(tmp_path / "cut"))


def test_grid_field_order():
    """Columns ordered J fastest map onto the history's (L_J, L_I) arrays."""
    from mgcm_rt import history as H
    col = np.arange(12)
    g = H.grid_field(col, 4, 3)
    for i in range(3):
        for j in range(4):
            assert g[j, i] == i * 4 + j


@pytest.mark.skipif(not os.path.exists(HTEST), reason="oracle/_ref/htest not built (reference absent)")
@pytest.mark.parametrize("rec,j,i", [(1, 1, 1), (2, 2, 3), (3, 4, 2)])
def test_reference_htest_reads_the_written_history(tmp_path, rec, j, i):
    """The reference's htest (historymod's readheader/skiprecords/readvariables)
    opens the file and prints the radiation fields of the chosen record and point."""
    rng = np.random.default_rng(7)
    steps = [_fields(rng, DIMS) for _ in range(3)]
    _write(tmp_path / "fort.11_0001", steps, rng)
    r = subprocess.run([HTEST], input=f"fort.11_0001\n{rec}\n{j},{i},1\n", capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stderr
    s = steps[rec - 1]
    printed = dict(re.findall(r"(\w+)\(J,I\)\s*=\s*(-?[0-9.]+)", r.stdout))
    for k in ("fuptopv", "fdntopv", "fupsurfv", "fdnsurfv", "fuptopir", "fupsurfir", "fdnsurfir"):
        assert printed[k] == f"{float(s[k][j - 1, i - 1]):.5f}", (k, printed[k])
    assert printed["SURFALB"] == f"{float(s['surfalb'][j - 1, i - 1]):.4f}"
    assert printed["P"] == f"{float(np.float32(s['p'][j - 1, i - 1])):.4f}"
    tau = re.search(r"TAU =\s*([0-9.]+)", r.stdout).group(1)
    assert tau == f"{s['tau']:.2f}"


def _oracle_grid(oracle, seed=3, ncol=64, nlay=24):
    from mgcm_rt import synth
    t, c = synth.make_workload(seed, ncol, nlay)
    c["acosz"][::5] = 0.0  # sun-down columns
    return t, c, oracle.run(t, c)


def test_oracle_history_fields(oracle):
    """REAL*4 fields are the nearest floats of the fluxes; VIS ones and surfalb
    follow the sun-down rule."""
    from oracle import history_fields
    t, c, o = _oracle_grid(oracle)
    prev = np.full(64, -7.0, np.float32)
    h = history_fields(c, o, surfalb=prev)
    up = c["acosz"] >= np.float32(1e-5)
    assert up.sum() < 64 and up.sum() > 0
    for k, src, lev in (("fuptopv", "fluxupv", 0), ("fdntopv", "fluxdnv", 0), ("fupsurfv", "fluxupv", -1),
                        ("fdnsurfv", "fluxdnv", -1), ("fuptopir", "fluxupi", 0), ("fupsurfir", "fluxupi", -1),
                        ("fdnsurfir", "fluxdni", -1)):
        d = o[src][:, lev]
        f = h[k].astype(np.float64)
        vis = k.endswith("v")
        sel = up if vis else np.ones(64, bool)
        # nearest: no float32 lies strictly closer to the double than the stored one
        lo = np.nextafter(h[k], np.float32(-np.inf)).astype(np.float64)
        hi = np.nextafter(h[k], np.float32(np.inf)).astype(np.float64)
        assert np.all(np.abs(f - d)[sel] <= np.abs(lo - d)[sel]) and np.all(np.abs(f - d)[sel] <= np.abs(hi - d)[sel])
        if vis:
            assert not np.any(h[k][~up])
    assert np.array_equal(h["surfalb"][~up], prev[~up])
    assert np.array_equal(h["surfalb"][up], c["als"][up].astype(np.float32))


def test_oracle_nlte_heating_limits(oracle):
    """hr_nlte: equal to hr where no sunlight reaches a layer, tending to hr as the
    non-LTE factor tends to 1, and the closed form of the difference."""
    from mgcm_rt import synth
    from oracle import history_fields
    t, c, o = _oracle_grid(oracle)
    nlay = 24
    sigma = synth.sigma_levels(nlay)
    p = c["plev"][:, -1] - 2.0e-4
    up = c["acosz"] >= np.float32(1e-5)
    h = history_fields(c, o, p=p, sigma=sigma, ptrop=2.0e-4)
    assert np.array_equal(h["hr_nlte"][~up], o["hr"][~up])  # sun down: SUNTOT = 0
    # factor -> 1 for huge PNLTE (22000 * 1e12 >> 1/eps): hr_nlte -> hr.  The
    # stratosphere's PNLTE is PTROP (SIGMA(2) = 0), so PTROP is the huge one.
    big = history_fields(c, o, p=np.full(64, 2e12), sigma=sigma, ptrop=1e12)["hr_nlte"]
    scale = np.max(np.abs(o["hr"]), axis=1, keepdims=True)
    assert np.max(np.abs(big - o["hr"]) / scale) < 1e-12
    # closed form: hr_nlte - hr = -SUNTOT/(1 + 22000 PNLTE) * g/(cp*100*dp) * 88775
    fv = o["fmnetv"]
    sun = np.concatenate([fv[:, :1] - o["nfluxtopv"][:, None], np.diff(fv, axis=1)], axis=1)
    sun[~up] = 0.0
    plev = c["plev"]
    dp = np.stack([plev[:, 0] + plev[:, 1]] + [plev[:, 2 * l + 2] - plev[:, 2 * l] for l in range(1, nlay + 1)], 1)
    pn = sigma[2 * np.arange(nlay + 1) + 1][None, :] * (p[:, None] - 2.0e-4) + 2.0e-4
    want = -sun / (1.0 + 22000.0 * pn) * 3.72 / (735.94 * 100.0 * dp) * 88775.0
    diff = h["hr_nlte"] - o["hr"]
    assert np.max(np.abs(diff - want) / scale) < 1e-12
    assert np.max(np.abs(want[up]) / scale[up]) > 1e-3  # the factor matters at the top
