"""The learned IR surrogate on the GPU: matrix-core GEMM and feature kernels
against numpy, device gradients against the CPU reference MLP, training that
learns (ablation against the method without the network), the one-part modes
of the plain path it relies on, and the exact fallback."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi", "diffvt")


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t if dtype is None else t.to(dtype)).to("cuda:0")


def test_gemm_matches_numpy():
    import ctypes
    import torch
    import mlp_oracle as mo
    from mgcm_rt import load_library
    from mgcm_rt.surrogate import _bind
    lib = _bind(load_library())
    rng = np.random.Generator(np.random.PCG64(1))
    for m, n, k in ((100, 70, 96), (64, 64, 32), (1, 513, 160)):
        a = mo.bf16(rng.normal(0, 1, (m, k)).astype(np.float32))
        b = mo.bf16(rng.normal(0, 1, (k, n)).astype(np.float32))
        da, db = _dev(a, torch.bfloat16), _dev(b, torch.bfloat16)
        dc = torch.zeros((m, n), dtype=torch.float32, device="cuda:0")
        assert lib.mgcm_sg_gemm_check(m, n, k, ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(db.data_ptr()),
                                      ctypes.c_void_p(dc.data_ptr())) == 0
        want = a.astype(np.float64) @ b.astype(np.float64)
        err = np.abs(dc.cpu().numpy() - want).max() / np.abs(want).max()
        assert err < 1e-5, (m, n, k, err)


def test_features_match_numpy():
    import torch
    from mgcm_rt import synth, to_device
    from mgcm_rt.surrogate import NIN, features, raw_features
    c = synth.make_columns(3, 200, 36, clouds=True)
    raw = raw_features(c)
    mean, std = raw.mean(axis=0), raw.std(axis=0) + 1e-3
    x = torch.empty((200, NIN), dtype=torch.float32, device="cuda:0")
    features(to_device(c), _dev(mean, torch.float32), _dev(1.0 / std, torch.float32), x)
    torch.cuda.synchronize()
    want = ((raw - mean) / std).astype(np.float32)
    assert np.allclose(x.cpu().numpy(), want, rtol=1e-4, atol=1e-4)


def test_device_gradients_match_reference_mlp():
    import mlp_oracle as mo
    import torch
    from mgcm_rt.surrogate import Net
    rng = np.random.Generator(np.random.PCG64(2))
    net = Net(64, seed=7)
    p = net.params()
    x = rng.normal(0, 1, (64, mo.NIN)).astype(np.float32)
    t = rng.normal(0, 1, (64, mo.NOUT)).astype(np.float32)
    y = torch.empty((64, mo.NOUT), dtype=torch.float32, device="cuda:0")
    net.forward(_dev(x), y)
    yo, _, _ = mo.forward(p, x.astype(np.float64), exact=False)
    assert np.abs(y.cpu().numpy() - yo).max() <= 2e-2 * np.abs(yo).max()
    g = net.gradients(_dev(x), _dev(t)).astype(np.float64)
    _, go = mo.loss_and_grads(p, x.astype(np.float64), t.astype(np.float64), exact=False)
    o = 0
    for k, n in mo.SHAPES:
        for size in (k * n, n):
            a, b = g[o:o + size], go[o:o + size]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            assert rel < 2e-2, (k, n, size, rel)
            o += size
    assert np.all(g[:160 * 512].reshape(160, 512)[147:] == 0.0)


def test_one_part_modes_match_full_run():
    from mgcm_rt import FP64, SKIP_IR, SKIP_VIS, Radiation, synth
    t, c = synth.make_workload(9, 200, 36, clouds=True)
    rt = Radiation(t)
    full = rt.run_host(c)
    ir = rt.run_host(c, precision=FP64 | SKIP_VIS)
    vis = rt.run_host(c, precision=FP64 | SKIP_IR)
    for k in ("fmneti", "fluxupi", "fluxdni", "nfluxtopi"):
        assert np.array_equal(ir[k], full[k]), k
    for k in ("fmnetv", "fluxupv", "fluxdnv", "nfluxtopv", "diffvt", "detau", "cumtauv"):
        assert np.array_equal(vis[k], full[k]), k
    assert not np.any(ir["fmnetv"]) and not np.any(vis["fmneti"])
    scale = np.abs(full["hr"]).max(axis=1, keepdims=True)
    assert np.abs(vis["hr"] + ir["hr"] - full["hr"]).max() <= 1e-12 * scale.max()


def _trained(ncol=65536, epochs=15, seed=11):
    import torch
    from mgcm_rt import Radiation, synth, to_device
    from mgcm_rt.surrogate import Net, exact_ir_heating, features, make_stats, raw_features, train
    t = synth.make_tables(seed, 36)
    c = synth.make_columns(seed, ncol, 36)
    rt = Radiation(t)
    dc = to_device(c)
    h = exact_ir_heating(rt, dc)
    stats = make_stats(raw_features(c), h.cpu().numpy(), "cuda:0")
    x = torch.empty((ncol, 147), dtype=torch.float32, device="cuda:0")
    features(dc, stats["mean"], stats["inv_std"], x)
    ty = ((torch.asinh(h / 10.0) - stats["tmean"]) / stats["tstd"]).float().contiguous()
    net = Net(512, seed=seed)
    hist = train(net, x, ty, epochs=epochs, batch=512, lr=2e-3, seed=seed)
    return rt, t, stats, net, hist, x, ty


def test_training_learns_and_beats_the_method_without_it():
    """Ablation: held-out error of the trained network against what the method
    gives without it (the training-set mean of every output)."""
    import torch
    from mgcm_rt import synth, to_device
    from mgcm_rt.surrogate import exact_ir_heating, features
    rt, t, stats, net, hist, _, _ = _trained()
    assert hist[-1] < 0.6 * hist[0], hist
    c = synth.make_columns(99, 2048, 36)
    dc = to_device(c)
    h = exact_ir_heating(rt, dc)
    x = torch.empty((2048, 147), dtype=torch.float32, device="cuda:0")
    features(dc, stats["mean"], stats["inv_std"], x)
    y = torch.empty((2048, 36), dtype=torch.float32, device="cuda:0")
    for b in range(0, 2048, 512):
        net.forward(x[b:b + 512], y[b:b + 512])
    ty = (torch.asinh(h / 10.0) - stats["tmean"]) / stats["tstd"]
    err_net = float((y.double() - ty).pow(2).mean())
    err_mean = float(ty.pow(2).mean())      # the mean predictor outputs 0 in standardised space
    print(f"held-out MSE (standardised asinh): net {err_net:.4f}, mean predictor {err_mean:.4f}", flush=True)
    assert err_net < 0.4 * err_mean


def test_fallback_gives_exact_heating_where_flagged():
    import torch
    from mgcm_rt import synth, to_device
    from mgcm_rt.surrogate import Net, Surrogate, exact_ir_heating
    rt, t, stats, net, _, _, _ = _trained(ncol=2048, epochs=3)
    net2 = Net(512, seed=12)
    c = synth.make_columns(55, 300, 36)
    c["qh2o"][:20] = 0.3                     # H2O vmr far above the training range (inside the tables): must fall back
    dc = to_device(c)
    exact = exact_ir_heating(rt, dc)
    sg = Surrogate(stats, [net, net2], threshold=float("inf"))
    h, bad = sg.heating(rt, dc)
    assert bool(bad[:20].all())
    assert torch.equal(h[bad], exact[bad])
    pred, _, _ = sg.predict(dc)
    assert torch.equal(h[~bad], pred[~bad])
    sg0 = Surrogate(stats, [net, net2], threshold=0.0)
    h0, bad0 = sg0.heating(rt, dc)
    assert bool(bad0.all()) and torch.equal(h0, exact)



def test_total_heating_both_fallback_schedules():
    """The learned form's whole call: with every column falling back it makes one
    combined plain run, whose heating it returns bit for bit; with every column
    trusted it makes a VIS-only run and adds the network's IR heating; in between
    (below the switch) the fallback columns get the VIS-only run's heating plus
    their exact IR heating."""
    import torch
    from mgcm_rt import FP64, SKIP_IR, alloc_outputs, synth, to_device
    from mgcm_rt.surrogate import Net, Surrogate, exact_ir_heating
    rt, t, stats, net, _, _, _ = _trained(ncol=2048, epochs=3)
    net2 = Net(512, seed=13)
    c = synth.make_columns(56, 300, 36)
    c["qh2o"][:20] = 0.3                     # far outside the training range: falls back
    dc = to_device(c)
    full = to_device(alloc_outputs(300, 36))
    rt.run_device(dc, full, FP64)
    vis = to_device(alloc_outputs(300, 36))
    rt.run_device(dc, vis, FP64 | SKIP_IR)
    exact = exact_ir_heating(rt, dc)
    pred, _, _ = Surrogate(stats, [net, net2], threshold=float("inf")).predict(dc)
    out = to_device(alloc_outputs(300, 36))
    h, bad, combined = Surrogate(stats, [net, net2], threshold=0.0).total_heating(rt, dc, out)
    assert combined and bool(bad.all()) and torch.equal(h, full["hr"][:, 1:])
    out = to_device(alloc_outputs(300, 36))
    h, bad, combined = Surrogate(stats, [net, net2], threshold=float("inf")).total_heating(rt, dc, out)
    # columns outside the (small) training run's feature range fall back too
    assert bool(bad[:20].all()) and int(bad.sum()) < 150 and not combined, int(bad.sum())
    assert torch.equal(h[~bad], vis["hr"][~bad][:, 1:] + pred[~bad])
    assert torch.equal(h[bad], vis["hr"][bad][:, 1:] + exact[bad])
    # the VIS-only heating plus the exact IR heating is the plain heating to rounding
    hb, fb = h[bad], full["hr"][bad][:, 1:]
    assert float((hb - fb).abs().max()) <= 1e-9 * float(fb.abs().max())
