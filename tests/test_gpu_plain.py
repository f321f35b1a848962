"""FP64 plain path on the GPU against the oracle (rel <= 1e-10), through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10  # plain FP64 path vs the reference's FP64 arithmetic
KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
        "diffvt", "detau", "cumtauv", "hr")


def compare(got, want, keys=KEYS, tol=TOL):
    errs = {k: rel_err(got[k], want[k]) for k in keys}
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"rel errors above {tol}: {bad}"
    return errs


@pytest.mark.parametrize("nlay,weights,seed", [(24, "dirichlet", 11), (24, "reference", 12), (36, "dirichlet", 13)])
def test_plain_matches_oracle(oracle, nlay, weights, seed):
    from mgcm_rt import Radiation, synth
    t, c = synth.make_workload(seed, 300, nlay, clouds=(nlay == 36), weights=weights)
    want = oracle.run(t, c)
    rt = Radiation(t)
    got = rt.run_host(c)
    errs = compare(got, want)
    print(nlay, weights, {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)


def libm_arguments(n: int, seed: int) -> np.ndarray:
    """Arguments of the FP64 path's transcendentals: exp arguments in [-40, 40] and over
    its whole finite range; positive doubles of every binade (random bits), pressures
    1e-7..1e5 mbar (TPINDEX's LOG10), values near 1 (the log's polynomial path)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    q = n // 5
    bits = rng.integers(0x0010000000000000, 0x7FEFFFFFFFFFFFFF, q, dtype=np.int64)
    return np.concatenate([rng.uniform(-40, 40, q), rng.uniform(-745, 709, q), bits.view(np.float64),
                           10 ** rng.uniform(-7, 5, q), rng.uniform(0.9, 1.1, n - 4 * q),
                           [0.0, 1.0, -35.0, 35.0, 1e-300, 5e-324, 2.2250738585072014e-308, 0.5, 2.0]])


def test_device_libm_matches_host_libm(oracle):
    """exp, log10 and log bit-identical to the host C library (the functions the
    reference's EXP, LOG10 and LOG call) on 12 M arguments; sqrt correctly rounded."""
    from mgcm_rt import libm_check
    x = libm_arguments(12_000_000, 77)
    e, l, q, ln = libm_check(x)
    he, hl, hln = oracle.libm(x)
    assert np.array_equal(e.view(np.int64), he.view(np.int64)), int(np.sum(e != he))
    pos = x > 0
    nb10 = int(np.sum(l[pos].view(np.int64) != hl[pos].view(np.int64)))
    nbln = int(np.sum(ln[pos].view(np.int64) != hln[pos].view(np.int64)))
    assert nb10 == 0 and nbln == 0, (nb10, nbln)
    nn = x >= 0
    assert np.array_equal(q[nn], np.sqrt(x[nn]))
    print("exp/log10/log: 0 differences of", x.size, flush=True)


def test_run_captured_in_a_hip_graph():
    """mgcm_rt_run needs no host synchronisation or allocation after its first call at a
    size, so a caller can capture it (here in a torch CUDA graph, i.e. a hipGraph) and
    replay it; every replay, with new inputs copied in between, equals the eager run."""
    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    t, c = synth.make_workload(81, 900, 24)
    rt = Radiation(t)
    dc = to_device(c)
    do = to_device(alloc_outputs(900, 24))
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        rt.run_device(dc, do, FP64, s.cuda_stream)  # first call at this size: workspace
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        rt.run_device(dc, do, FP64, s.cuda_stream)
    for k in range(3):
        ck = synth.make_columns(82 + k, 900, 24)
        want = rt.run_host(ck)
        for key in dc:
            if hasattr(dc[key], "copy_"):
                dc[key].copy_(torch.from_numpy(np.ascontiguousarray(ck[key], dtype=np.float64)))
        for key, v in do.items():  # DETAU/CUMTAUV of sun-down columns keep what they held
            v.fill_(0.0 if key in ("detau", "cumtauv") else -1.0)
        g.replay()
        torch.cuda.synchronize()
        for key in want:
            assert np.array_equal(do[key].cpu().numpy(), want[key]), (k, key)


def test_two_contexts_on_two_streams_and_two_host_threads():
    """Contexts share no mutable state: two of them (different tables and layer
    counts) interleaved on two streams, and driven from two host threads at once
    (ctypes releases the GIL), give each context's serial results bit for bit."""
    import threading

    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    ta, ca = synth.make_workload(91, 3000, 24)
    tb, cb = synth.make_workload(92, 2000, 36, clouds=True)
    ra, rb = Radiation(ta), Radiation(tb)
    want = {"a": ra.run_host(ca), "b": rb.run_host(cb)}
    da, db = to_device(ca), to_device(cb)
    oa, ob = to_device(alloc_outputs(3000, 24)), to_device(alloc_outputs(2000, 36))
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        ra.run_device(da, oa, FP64, sa.cuda_stream)
        rb.run_device(db, ob, FP64, sb.cuda_stream)
    torch.cuda.synchronize()
    for name, o in (("a", oa), ("b", ob)):
        for k in KEYS:
            assert np.array_equal(o[k].cpu().numpy(), want[name][k]), (name, k)
    got, errors = {}, []

    def work(name, rt, c):
        try:
            got[name] = [rt.run_host(c) for _ in range(3)]
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(e)
    threads = [threading.Thread(target=work, args=("a", ra, ca)), threading.Thread(target=work, args=("b", rb, cb))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for name in ("a", "b"):
        for o in got[name]:
            for k in KEYS:
                assert np.array_equal(o[k], want[name][k]), (name, k)
