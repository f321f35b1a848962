"""Radiative columns/s of the MI355X column radiation on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--ncol 64800] [--nlay 24]

A step is one mgcm_rt_run over the whole grid: 64,800 seeded synthetic columns
(180x360), 24 layers, 7 VIS + 5 IR bands x 17 Gauss points, inputs resident in
HBM.  Prints one JSON line.  Besides the plain FP64 path it reports the
mixed-precision and surrogate forms when they are built, each with its error
against the FP64 path, the FP64 kernel's roofline position, the reference's
own Fortran (oracle/_ref, built from its sources) timed on one host core over a
bounded sample of the same columns, and the C restatement with OpenMP on all
host cores over the whole grid.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "legacy-mars-global-climate-model_amd")]

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector/matrix peak
HBM_PEAK_GBS = 8000.0
SECONDS_PER_SOL = 88775.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--ncol", type=int, default=64800)
    p.add_argument("--nlay", type=int, default=24)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--cpu-sample", type=int, default=64800,
                   help="columns of the same grid the CPU baseline runs (the whole grid, about 14 s on one core)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-forms", action="store_true", help="only the FP64 headline path")
    p.add_argument("--forms", default="", help="comma-separated subset of the forms (default: all)")
    p.add_argument("--sg-train", type=int, default=2000000, help="surrogate training columns")
    p.add_argument("--sg-epochs", type=int, default=10)
    p.add_argument("--copy-calib", action="store_true",
                   help="also run a 512 MiB device copy (known HBM bytes) for scaling FETCH_SIZE/WRITE_SIZE")
    return p.parse_args()


def timed_steps(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall / steps * 1e3


# Algorithmic work of the plain step per column: executed FP64 +-*/, exp, sqrt
# and log, counted by the CPU restatement's op counter on the first 1,024
# columns of the seeded workload, keyed by (seed, layers).  Every column of
# these workloads executes the same counts.  tests/test_bench_ops.py recounts
# them with the op counter.
OPS_PER_COLUMN = {
    (2024, 24): {"flops": 984769.0, "exps": 25517.0, "sqrts": 5100.0, "logs": 100.0},
}


def _sample(cols, n):
    sub = {k: (v[:n] if isinstance(v, np.ndarray) else v) for k, v in cols.items()}
    sub["ncol"] = min(n, int(cols["ncol"]))
    return sub


def cpu_baseline(seed, cols, nsample, keep=None):
    """The reference's own Fortran (oracle/_ref, built from its sources) on one core.

    It runs with the reference's single GWEIGHT set (one SFLUXV/SFLUXI per
    column, its native cost); the GPU time does not depend on the weights.
    With keep (a dict), its tables, columns and outputs are stored there for
    the parity check against the device on the same inputs.
    """
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Reference
    from mgcm_rt import synth
    ref = Reference()
    if ref.nlay != int(cols["nlay"]):
        raise RuntimeError(f"reference build has {ref.nlay} layers")
    t = synth.make_tables(seed, int(cols["nlay"]), weights="reference")
    ref.run_grid(t, _sample(cols, 64))
    sub = _sample(cols, nsample)
    t0 = time.perf_counter()
    out = ref.run_grid(t, sub)
    dt = time.perf_counter() - t0
    if keep is not None:
        keep.update(tables=t, cols=sub, out=out)
    return {"value": sub["ncol"] / dt, "unit": "columns/s", "cores": 1, "kind": "reference",
            "sample": f"first {sub['ncol']} of the {cols['ncol']} seeded columns, {int(cols['nlay'])} layers, "
                      f"reference OPTCV/SFLUXV/OPTCI/SFLUXI (amdflang -O2 build), one core, {dt:.2f} s wall"}


def cpu_port(tables, cols, keep=None):
    """The C restatement with OpenMP on every host core, over the whole grid.  With
    keep (a dict), its outputs are stored there for the parity check."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    o = Oracle()
    o.run(tables, _sample(cols, 64), nthreads=threads)
    t0 = time.perf_counter()
    out = o.run(tables, cols, nthreads=threads)
    dt = time.perf_counter() - t0
    if keep is not None:
        keep.update(out)
    return {"value": int(cols["ncol"]) / dt, "unit": "columns/s", "cores": threads, "kind": "port",
            "sample": f"all {cols['ncol']} columns, C restatement with OpenMP, {dt:.2f} s wall"}


PARITY_KEYS = ("fmnetv", "fluxupv", "fluxdnv", "fmneti", "fluxupi", "fluxdni", "nfluxtopv", "nfluxtopi",
               "diffvt", "detau", "cumtauv", "hr")


def parity(got: dict, want: dict) -> dict:
    """Max over the grid of |got - want| / max(|want|, the column's max |want|), per output."""
    errs = {}
    for k in PARITY_KEYS:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        g2, w2 = g.reshape(g.shape[0], -1), w.reshape(w.shape[0], -1)
        s = np.max(np.abs(w2), axis=1, keepdims=True)
        errs[k] = float(np.max(np.abs(g2 - w2) / np.maximum(np.maximum(np.abs(w2), s), 1e-300)))
    return {"max_rel_err": max(errs.values()), "per_output": errs, "tolerance": 1e-10,
            "pass": all(v <= 1e-10 for v in errs.values())}


def library_sha256():
    import hashlib
    from mgcm_rt import LIB_PATH
    with open(LIB_PATH, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def profile_figures(ncol, nlay):
    """Figures of the committed rocprofv3 summaries of this step, used only when they
    were measured on the library that is loaded (their library_sha256 line): HBM bytes
    per step (FETCH_SIZE/WRITE_SIZE, calibrated) and the FP64 VALU operations the
    counters saw (SQ_INSTS_VALU_FLOPS_FP64)."""
    sha = library_sha256()
    res = {"traffic": None, "counter_fp64_flop": None, "profile_library_matches": False}
    for name, key, tag in ((f"k_rt_fp64_{ncol}x{nlay}_hbm.txt", "traffic", "traffic_bytes_per_launch"),
                           (f"k_rt_fp64_{ncol}x{nlay}.txt", "counter_fp64_flop", "counter_fp64_flop_per_step")):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        lines = open(path).read().splitlines()
        stamp = [ln.split("=")[1].strip() for ln in lines if ln.startswith("library_sha256")]
        if not stamp or stamp[0] != sha:
            continue
        res["profile_library_matches"] = True
        for ln in lines:
            if ln.startswith(tag):
                res[key] = float(ln.split("=")[1].split()[0])
    return res


def _precision_forms():
    """The reduced-precision forms measured on the mixed workload: the mixed path
    (k-tables stored in FP32) and the precision study's components of the
    FP32/FP64 split, one at a time and together, plus FP64 with a cheaper division."""
    from mgcm_rt import MIXED, P_FASTDIV, P_INTERP32, P_TAB32, P_TAU32, P_TRANS32, STUDY
    return (("tables_fp32 (the mixed path)", MIXED),
            ("k_interpolation_fp32 (tables and arithmetic)", STUDY | P_TAB32 | P_INTERP32),
            ("optical_depths_fp32", STUDY | P_TAU32),
            ("transmissions_fp32", STUDY | P_TRANS32),
            ("all_four_fp32 (the configuration's split)", STUDY | P_TAB32 | P_INTERP32 | P_TAU32 | P_TRANS32),
            ("fp64_fast_division", STUDY | P_FASTDIV),
            ("fp64_fast_division_tables_fp32", STUDY | P_FASTDIV | P_TAB32))


def form_mixed(args, rt_unused, stream, tables=None, label="cloud ice"):
    """Mixed precision on 64,800 columns x 36 layers with cloud ice, against the
    FP64 path on the same columns (itself within 1e-10 of the oracle): per form,
    ms per step, speedup, max and RMS heating error (K/sol) and max flux error
    (W/m^2).  The headline value is the fastest form within the 0.05 K/sol bar
    that beats FP64, or -- when none does -- the mixed path, flagged as missing
    the bar."""
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    ncol, nlay, bar = 64800, 36, 0.05
    t, c = synth.make_workload(args.seed + 1, ncol, nlay, clouds=True)
    if tables is not None:
        t = tables
    rt = Radiation(t)
    dc = to_device(c)
    d64, dm = to_device(alloc_outputs(ncol, nlay)), to_device(alloc_outputs(ncol, nlay))
    sh = stream.cuda_stream
    ms64, _ = timed_steps(lambda: rt.run_device(dc, d64, FP64, sh), args.steps, args.warmup, stream)
    rt.check()
    forms = {}
    for name, mode in _precision_forms():
        ms, _ = timed_steps(lambda: rt.run_device(dc, dm, mode, sh), args.steps, args.warmup, stream)
        rt.check()
        err = (dm["hr"] - d64["hr"]).abs()
        flux = max(float((dm[k] - d64[k]).abs().max()) for k in ("fmnetv", "fmneti", "fluxupv", "fluxdnv",
                                                                 "fluxupi", "fluxdni"))
        forms[name] = {"ms_per_step": ms, "speedup_vs_fp64": ms64 / ms, "max_err_K_per_sol": float(err.max()),
                       "rms_err_K_per_sol": float(err.pow(2).mean().sqrt()), "max_flux_err_W_m2": flux,
                       "meets_bar": bool(float(err.max()) <= bar and ms < ms64)}
    ok = [n for n, f in forms.items() if f["meets_bar"]]
    head = min(ok, key=lambda n: forms[n]["ms_per_step"]) if ok else _precision_forms()[0][0]
    f = forms[head]
    out = {"workload": f"{ncol} columns x {nlay} layers, {label}", "form": head, "value": ncol / (f["ms_per_step"] / 1e3),
           "unit": "columns/s", "ms_per_step": f["ms_per_step"], "fp64_ms_per_step": ms64,
           "speedup_vs_fp64": f["speedup_vs_fp64"], "max_err_K_per_sol": f["max_err_K_per_sol"],
           "rms_err_K_per_sol": f["rms_err_K_per_sol"], "max_flux_err_W_m2": f["max_flux_err_W_m2"],
           "bar_K_per_sol": bar, "meets_bar": f["meets_bar"],
           "median_abs_hr_K_per_sol": float(d64["hr"].abs().median()),
           "max_abs_hr_K_per_sol": float(d64["hr"].abs().max()), "forms": forms}
    del d64, dm, dc
    return out


def _real_tables(nlay):
    """The reference's k-distribution: the records of run/data/CO2H2O_V_12_95_INTEL
    and CO2H2O_IR_12_95_INTEL (carried as data by tests/golden/column_real24.npz)
    written back to files, read by the library, refined by the device LAGINTERP."""
    import tempfile

    from mgcm_rt import realdata
    f = np.load(os.path.join(ROOT, "tests", "golden", "column_real24.npz"))
    with tempfile.TemporaryDirectory() as d:
        realdata.write_ktable(os.path.join(d, realdata.VIS_FILE), f["raw_v"], f["fzero_v"])
        realdata.write_ktable(os.path.join(d, realdata.IR_FILE), f["raw_i"], f["fzero_i"])
        return realdata.load_tables(d, nlay)


def form_real(args, rt_unused, stream):
    """The plain FP64 path on the reference's own k-distribution (_real_tables);
    the same seeded columns as the headline."""
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    t = _real_tables(args.nlay)
    c = synth.make_columns(args.seed, args.ncol, args.nlay)
    rt = Radiation(t)
    dc = to_device(c)
    do = to_device(alloc_outputs(args.ncol, args.nlay))
    ms, _ = timed_steps(lambda: rt.run_device(dc, do, FP64, stream.cuda_stream), args.steps, args.warmup, stream)
    rt.check()
    return {"workload": f"{args.ncol} columns x {args.nlay} layers, the reference's k-distribution "
                        "(CO2H2O_*_12_95_INTEL via LAGINTERP)",
            "value": args.ncol / (ms / 1e3), "unit": "columns/s", "ms_per_step": ms}


def form_mixed_real(args, rt_unused, stream):
    """The same forms on the reference's own k-distribution (_real_tables), in the
    mixed workload's shape (the same seeded columns as form_mixed)."""
    return form_mixed(args, rt_unused, stream, tables=_real_tables(36),
                      label="cloud ice, the reference's k-distribution")


def form_batched(args, rt, stream):
    """512 grids, U{2000..20000} columns, U{16..64} layers, in one
    mgcm_rt_run_batched call (plan, tuning and graph capture on the first call,
    then graph replays).  Inputs of grids with equal layer counts share one
    column pool; outputs are per grid.  After the timed replays, 64 sampled
    columns (8 from each of 8 grids spread over the layer range) are checked
    against the oracle at rel 1e-10, and the tuner's choices are reported."""
    import torch
    from mgcm_rt import FP64, alloc_outputs, synth, to_device
    shapes = synth.ragged_shapes(args.seed, 512)
    host_pools = synth.ragged_pools(args.seed, shapes)
    pools = {nl: to_device(p) for nl, p in host_pools.items()}
    grids, outs = [], []
    for n, nl in shapes:
        g = {k: (v[:n] if hasattr(v, "shape") else v) for k, v in pools[nl].items()}
        g["ncol"] = n
        grids.append(g)
        proto = alloc_outputs(1, nl)
        outs.append({k: torch.empty((n,) + v.shape[1:], dtype=torch.float64, device="cuda")
                     for k, v in proto.items()})
    sh = stream.cuda_stream
    t0 = time.perf_counter()
    rt.run_batched(grids, outs, FP64, sh)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    steps = max(1, min(args.steps, 3))
    ms, _ = timed_steps(lambda: rt.run_batched(grids, outs, FP64, sh), steps, 0, stream)
    rt.check()
    ncol = sum(n for n, _ in shapes)
    out = {"workload": f"512 grids, {ncol} columns, 16..64 layers", "value": ncol / (ms / 1e3),
           "unit": "columns/s", "ms_per_batch": ms, "steps": steps, "first_call_s": first,
           "column_layers_per_s": sum(n * (nl + 1) for n, nl in shapes) / (ms / 1e3),
           "tuning": rt.batched_tuning()}
    try:  # sampled parity against the oracle
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        o = Oracle()
        order = sorted(range(len(shapes)), key=lambda i: shapes[i][1])
        worst = 0.0
        rng = np.random.Generator(np.random.PCG64(args.seed))
        for i in [order[j] for j in np.linspace(0, len(order) - 1, 8).astype(int)]:
            n, nl = shapes[i]
            pick = np.sort(rng.choice(n, 8, replace=False))
            sub = {k: (v[pick] if isinstance(v, np.ndarray) and v.ndim and v.shape[0] >= n else v)
                   for k, v in host_pools[nl].items()}
            sub["ncol"] = len(pick)
            want = o.run(synth.make_tables(args.seed, nl), sub)
            got = {k: outs[i][k][torch.from_numpy(pick).cuda()].cpu().numpy() for k in PARITY_KEYS}
            worst = max(worst, parity(got, want)["max_rel_err"])
        out["parity_sampled"] = {"columns": 64, "max_rel_err": worst, "tolerance": 1e-10, "pass": worst <= 1e-10}
    except Exception as e:  # noqa: BLE001
        out["parity_sampled"] = {"error": f"{type(e).__name__}: {e}"[:200]}
    del grids, outs, pools
    return out


SWEEP = [(24, 540), (24, 1080), (24, 2160), (24, 4320), (24, 8100), (24, 64800),
         (36, 2160), (36, 8100), (36, 64800)]


def form_small_grid(args, rt, stream):
    """The reference's own CPU case on the GPU: 2,160 columns (a 36 x 60 grid) x
    24 layers, one single-grid call per step, at each column-kernel tile (1, 2
    or 4 band slots per column: 4, 2 or 1 columns per wave; 12: one band per
    wave and a finishing kernel) and at the library's automatic choice.  Every tile's outputs are compared bit for bit
    with the one-slot tile's, and the automatic tile's with the oracle at rel
    1e-10.  At 4 columns per wave this grid is 540 waves, a quarter of the
    device's resident slots.  "sweep" repeats the per-tile timing over SWEEP's
    (layers, columns) shapes, with the automatic choice of each."""
    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    if args.nlay != 24:
        return {"skipped": "the context's tables are built for another layer count"}
    sh = stream.cuda_stream

    def per_tile(ctx, dc, do, steps):
        res, first, same = {}, None, True
        for tile in (1, 2, 4, 12):
            ctx.set_tile(tile)
            ms, _ = timed_steps(lambda: ctx.run_device(dc, do, FP64, sh), steps, 3, stream)
            got = {k: do[k].cpu().numpy() for k in PARITY_KEYS}
            if first is None:
                first = got
            else:
                same = same and all(np.array_equal(got[k], first[k]) for k in PARITY_KEYS)
            res[str(tile)] = ms
        ctx.set_tile(0)
        return res, same

    ncol = 2160
    c = synth.make_columns(args.seed, ncol, 24)
    dc, do = to_device(c), to_device(alloc_outputs(ncol, 24))
    tiles, same = per_tile(rt, dc, do, 50)
    auto = rt.tile_for(ncol, 24)
    ms, _ = timed_steps(lambda: rt.run_device(dc, do, FP64, sh), 50, 5, stream)
    rt.check()
    out = {"workload": f"{ncol} columns (36 x 60) x 24 layers, plain FP64, one call per step",
           "auto_tile": auto, "ms_per_step": ms, "value": ncol / (ms / 1e3), "unit": "columns/s",
           "speedup_vs_tile1": tiles["1"] / ms, "tiles_ms": tiles, "tiles_bit_identical": same}
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        want = Oracle().run(synth.make_tables(args.seed, 24), c)
        out["parity"] = parity({k: v.cpu().numpy() for k, v in do.items()}, want)
    except Exception as e:  # noqa: BLE001
        out["parity"] = {"error": f"{type(e).__name__}: {e}"[:200]}
    del dc, do
    sweep = []
    ctxs = {24: rt}
    for nlay, n in SWEEP:
        if nlay not in ctxs:
            ctxs[nlay] = Radiation(synth.make_tables(args.seed, nlay))
        ctx = ctxs[nlay]
        cs = synth.make_columns(args.seed, n, nlay)
        dcs, dos = to_device(cs), to_device(alloc_outputs(n, nlay))
        t, same_s = per_tile(ctx, dcs, dos, 50 if n < 20000 else 10)
        best = min(t, key=t.get)
        pick = ctx.tile_for(n, nlay)
        sweep.append({"nlay": nlay, "ncol": n, "ms": t, "auto_tile": pick, "fastest_tile": int(best),
                      "auto_vs_fastest": t[str(pick)] / t[best], "bit_identical": same_s})
        del dcs, dos
    ctxs.pop(24)
    del ctxs
    out["sweep"] = sweep
    return out


def _graph_ms(fn, stream, reps=20, replays=5):
    """GPU time of one fn() call: reps calls captured in a HIP graph, replayed;
    a kernel of a few microseconds is not hidden behind Python launch overhead."""
    import torch
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        for _ in range(reps):
            fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):  # replay() launches on the current stream
        g.replay()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(replays):
            g.replay()
        e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / (replays * reps)


def form_next_rows(args, rt_unused, stream_unused):
    """The kernels next to the radiation call, each alone on the headline grid
    (64,800 columns x 24 layers): DSOLFLUX (code/dsolflux.f), OPT_DST / OPT_CLD
    (code/opt_dst.f, code/opt_cld.f), the Conrath profile and FILLTAUCUM
    (code/dustprofile.f, code/filltaucum.f), COMP3's history diagnostics
    (code/comp3.f:654,703-747,765-771), and the start-up LAGINTERP and Planck
    table.  GPU times from HIP graphs of 20 calls; bytes = the inputs and
    outputs each call must move, against 8 TB/s."""
    import ctypes as C
    import tempfile

    import torch
    from mgcm_rt import (AER_OUT_KEYS, DP, FP64, HISTORY_KEYS, AerosolOut, Radiation, Tracers, alloc_outputs,
                         dust, laginterp, planck_table, read_aerosol_tables, realdata, synth, to_device)
    n, nlay = args.ncol, args.nlay
    nlev, nl = 2 * nlay + 3, nlay + 1
    t, c = synth.make_workload(args.seed, n, nlay)
    rt = Radiation(t)
    dc, do = to_device(c), to_device(alloc_outputs(n, nlay))
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()  # graphs are captured on a stream of their own, not the legacy default one
    sh = stream.cuda_stream
    rt.run_device(dc, do, FP64, sh)
    torch.cuda.synchronize()
    out = {}

    def entry(name, ms, nbytes, what):
        gbs = nbytes / (ms / 1e3) / 1e9
        out[name] = {"what": what, "ms": ms, "columns_per_s": n / (ms / 1e3), "bytes": nbytes, "GBps": gbs,
                     "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}

    # DSOLFLUX + DNVFLUX
    ds = torch.empty(n, dtype=torch.float64, device="cuda")
    dn = torch.empty_like(ds)
    sol = np.ascontiguousarray(rt.sol, dtype=np.float64)
    solp = sol.ctypes.data_as(DP)
    args_ds = (rt.ctx, n, solp, dc["acosz"].data_ptr(), do["detau"].data_ptr(), do["diffvt"].data_ptr(),
               ds.data_ptr(), dn.data_ptr(), C.c_void_p(sh))
    assert rt.lib.mgcm_rt_dsolflux(*args_ds) == 0
    entry("dsolflux", _graph_ms(lambda: rt.lib.mgcm_rt_dsolflux(*args_ds), stream),
          n * 8 * (17 * 7 + 2 + 2), "k_dsolflux: DETAU[17][7], ACOSZ, DIFFVT in; DIRECTSOL, DNVFLUX out")

    # OPT_DST / OPT_CLD on the reference's optical tables (carried as data by the fixture)
    fx = np.load(os.path.join(ROOT, "tests", "golden", "aerosol_ref.npz"))
    with tempfile.TemporaryDirectory() as d:
        for k in fx.files:
            if k.startswith("file_"):
                with open(os.path.join(d, k[5:]), "wb") as f:
                    f.write(fx[k].tobytes())
        rt.set_aerosol_tables(read_aerosol_tables(d))
    qtr = torch.from_numpy(synth.make_tracers(args.seed, n, nlay)).cuda()
    ld = nlev + 1
    shapes = {"qxv": (n, 7, ld), "qsv": (n, 7, ld), "gv": (n, 7, ld), "qxi": (n, 5, ld), "qsi": (n, 5, ld),
              "gi": (n, 5, ld), "qextref": (n, ld), "tauref": (n, ld), "tau2": (n, 2)}
    ao_t = {k: torch.empty(v, dtype=torch.float64, device="cuda") for k, v in shapes.items()}
    tr = Tracers(n, nlay, dc["plev"].data_ptr(), qtr.data_ptr())
    ao = AerosolOut(*[ao_t[k].data_ptr() for k in AER_OUT_KEYS])
    aer_bytes = n * 8 * (6 * nlay + nlev + (12 * 3 + 2) * ld + 2)
    for name, fn in (("opt_dst", rt.lib.mgcm_rt_opt_dst), ("opt_cld", rt.lib.mgcm_rt_opt_cld)):
        call = (rt.ctx, C.byref(tr), C.byref(ao), C.c_void_p(sh))
        assert fn(*call) == 0
        entry(name, _graph_ms(lambda fn=fn, call=call: fn(*call), stream), aer_bytes,
              f"k_opt_aer: QTRACE[6][{nlay}], PL in; Qext/Qscat/g per band and level, QEXTREF, TAUREF out")

    # Conrath dust profile + FILLTAUCUM
    rng = np.random.default_rng(args.seed)
    prdst = torch.from_numpy(dust.pressure_grid(7.01, 0.0008)).cuda()
    dd = {"tautot": torch.from_numpy(rng.uniform(0.1, 1.0, n)).cuda(),
          "conr": torch.full((n,), 0.03, dtype=torch.float64, device="cuda"),
          "p": torch.from_numpy(c["plev"][:, -1] - 0.0008).cuda(),
          "sigma": torch.from_numpy(synth.sigma_levels(nlay)).cuda()}
    trd = torch.empty((n, ld), dtype=torch.float64, device="cuda")
    tdd = torch.empty((n, prdst.numel()), dtype=torch.float64, device="cuda")
    lib = rt.lib
    lib.mgcm_rt_dust_tauref.argtypes = [C.POINTER(dust.DustColumns), C.c_void_p, C.c_void_p, C.c_void_p]
    dcol = dust.DustColumns(n, nlay, *[dd[k].data_ptr() for k in ("tautot", "conr", "p", "sigma")],
                            prdst.data_ptr(), prdst.numel(), 6.1, 0.0008)
    call = (C.byref(dcol), trd.data_ptr(), tdd.data_ptr(), C.c_void_p(sh))
    assert lib.mgcm_rt_dust_tauref(*call) == 0
    entry("dust_tauref", _graph_ms(lambda: lib.mgcm_rt_dust_tauref(*call), stream),
          n * 8 * (3 + ld + prdst.numel()), "k_dust_tauref: TAUTOT, CONR, P in; TAUREF, TAUDST out")

    # history diagnostics with the non-LTE heating
    from mgcm_rt import History, column_struct, flux_struct
    hist = {k: torch.zeros(n, dtype=torch.float32, device="cuda") for k in HISTORY_KEYS}
    hr = torch.empty((n, nl), dtype=torch.float64, device="cuda")
    hs = History(*[hist[k].data_ptr() for k in HISTORY_KEYS], hr.data_ptr(), dd["p"].data_ptr(),
                 dd["sigma"].data_ptr(), 0.0008)
    cs = column_struct(dc, ptr=lambda x: x.data_ptr(), sol=rt.sol)
    fs = flux_struct(do, ptr=lambda x: x.data_ptr())
    call = (C.byref(cs), C.byref(fs), C.byref(hs), C.c_void_p(sh))
    assert lib.mgcm_rt_history_fields(*call) == 0
    entry("history_fields", _graph_ms(lambda: lib.mgcm_rt_history_fields(*call), stream),
          n * (8 * (2 * nl + 2 * nl + 2 + 1 + 1 + 6 + nl) + 4 * 8),
          "k_history: FMNETV/I, PLEV, top/surface fluxes, ACOSZ, ALS, P in; 8 REAL*4 fields and the "
          "non-LTE heating out")

    # start-up tables (host + device, once per run)
    f = np.load(os.path.join(ROOT, "tests", "golden", "column_real24.npz"))
    t0 = time.perf_counter()
    laginterp(realdata.PGREF11, f["raw_v"])
    laginterp(realdata.PGREF11, f["raw_i"])
    t1 = time.perf_counter()
    planck_table()
    t2 = time.perf_counter()
    out["startup"] = {"what": "LAGINTERP of both k-tables (host log10/pow, device interpolation) and SETSPI's "
                              "Planck table, wall time of each library call", "laginterp_ms": (t1 - t0) * 1e3,
                      "planck_table_ms": (t2 - t1) * 1e3}
    out["workload"] = f"{n} columns x {nlay} layers, seed {args.seed}"
    del dc, do, ao_t, qtr, trd, tdd, hist, hr
    return out


def form_host_arrays(args, rt, stream_unused):
    """The same grid through mgcm_rt_run_host, the entry a host with its fields in
    host memory calls (the Fortran mgcm_rt_radiation): the library copies the
    inputs in, runs, copies the outputs back and synchronises.  Wall time per call
    with numpy (pageable) arrays: the PCIe-inclusive rate.  The library copies
    chunks of columns through pinned staging with its own threads while the
    device runs the chunk before."""
    from mgcm_rt import FP64, alloc_outputs, synth
    _, c = synth.make_workload(args.seed, args.ncol, args.nlay)
    out = alloc_outputs(args.ncol, args.nlay)
    rt.run_host(c, FP64, out)
    steps = max(1, min(args.steps, 5))
    t0 = time.perf_counter()
    for _ in range(steps):
        rt.run_host(c, FP64, out)
    ms = (time.perf_counter() - t0) / steps * 1e3
    nin = args.ncol * (5 * (2 * args.nlay + 3) + 2 * (2 * args.nlay + 4) + 3 + 2 * 17 * 7) * 8  # DETAU/CUMTAUV in/out
    nout = args.ncol * (7 * (args.nlay + 1) + 3 + 2 * 17 * 7) * 8
    return {"workload": f"{args.ncol} columns x {args.nlay} layers, host (pageable) arrays in and out",
            "value": args.ncol / (ms / 1e3), "unit": "columns/s", "ms_per_call": ms,
            "bytes_in": nin, "bytes_out": nout, "copy_GBps_if_all_copy": (nin + nout) / (ms / 1e3) / 1e9}


def form_surrogate(args, rt_unused, stream):
    """Learned IR surrogate with exact fallback on 64,800 fresh columns x 36 layers.

    Training: 2,000,000 clear-sky columns whose IR heating the plain path
    computes on the device; two networks (seeds 1 and 2), 10 epochs each, by
    the library's own BF16 matrix-core forward/backward kernels and Adam.  The
    disagreement threshold is calibrated at 0.05 K/sol on 16,384 held-out
    columns.  A step = VIS part exact + features + both networks + check +
    exact IR re-solve of the flagged columns, against the plain FP64 step on
    the same columns."""
    import torch
    from mgcm_rt import FP64, Radiation, alloc_outputs, synth, to_device
    from mgcm_rt.surrogate import (Net, Surrogate, calibrate, exact_ir_heating, features, make_stats,
                                   raw_features, train)
    bar, nlay, seed = 0.05, 36, args.seed + 2
    t = synth.make_tables(seed, nlay)
    rt = Radiation(t)
    t0 = time.perf_counter()
    ntrain = args.sg_train
    c = synth.make_columns(seed, ntrain, nlay)
    dc = to_device(c)
    h = exact_ir_heating(rt, dc)
    stats = make_stats(raw_features(c), h.cpu().numpy(), "cuda:0")
    x = torch.empty((ntrain, 147), dtype=torch.float32, device="cuda:0")
    features(dc, stats["mean"], stats["inv_std"], x)
    ty = ((torch.asinh(h / 10.0) - stats["tmean"]) / stats["tstd"]).float().contiguous()
    del dc, c, h
    nets, hists = [], []
    for s in (1, 2):
        net = Net(65536, seed=s)  # activations for a whole 64,800-column grid: one launch per layer
        hists.append(train(net, x, ty, epochs=args.sg_epochs, batch=1024, lr=2e-3, seed=s))
        nets.append(net)
    del x, ty
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # calibration on held-out columns
    cv = to_device(synth.make_columns(seed + 1, 16384, nlay))
    sg = Surrogate(stats, nets, threshold=0.0)
    pv, dis, inr = sg.predict(cv)
    ev = (pv - exact_ir_heating(rt, cv)).abs().amax(dim=1)
    ev = torch.where(inr, ev, torch.full_like(ev, float("inf")))
    sg.threshold = calibrate(dis, ev, bar)
    # evaluation
    ncol = 64800
    ce = synth.make_columns(seed + 2, ncol, nlay)
    de = to_device(ce)
    d64 = to_device(alloc_outputs(ncol, nlay))
    dv = to_device(alloc_outputs(ncol, nlay))
    sh = stream.cuda_stream
    ms64, _ = timed_steps(lambda: rt.run_device(de, d64, FP64, sh), args.steps, args.warmup, stream)
    exact = exact_ir_heating(rt, de)
    res = {}

    def step():
        hr, bad, combined = sg.total_heating(rt, de, dv, sh)
        res["hr"] = hr
        res["bad"] = bad
        res["combined"] = combined
    ms, _ = timed_steps(step, args.steps, args.warmup, stream)
    rt.check()
    err = (res["hr"] - d64["hr"][:, 1:]).abs()
    pred, dise, _ = sg.predict(de)
    perr = (pred - exact).abs()
    y = torch.empty((ncol, 36), dtype=torch.float32, device="cuda:0")
    xe = torch.empty((ncol, 147), dtype=torch.float32, device="cuda:0")
    features(de, stats["mean"], stats["inv_std"], xe)
    ms_net, _ = timed_steps(lambda: nets[0].forward(xe, y, sh), args.steps, args.warmup, stream)
    ty_e = (torch.asinh(exact / 10.0) - stats["tmean"]) / stats["tstd"]
    mse_net = float((y.double() - ty_e).pow(2).mean())
    mse_mean = float(ty_e.pow(2).mean())
    flops = 2 * (147 * 512 + 2 * 512 * 512 + 512 * 36)
    return {"workload": f"{ncol} fresh clear-sky columns x {nlay} layers; trained on {ntrain} columns",
            "value": ncol / (ms / 1e3), "unit": "columns/s", "ms_per_step": ms, "fp64_ms_per_step": ms64,
            "speedup_vs_fp64": ms64 / ms, "bar_K_per_sol": bar, "threshold": sg.threshold,
            "meets_bar": bool(float(err.max()) <= bar and ms < ms64),
            "fallback_fraction": float(res["bad"].double().mean()),
            "fallback_schedule": ("one combined plain run (VIS and IR) of the grid" if res["combined"]
                                  else "a VIS-only run of the grid + an IR-only run of the fallback columns"),
            "max_err_K_per_sol": float(err.max()), "rms_err_K_per_sol": float(err.pow(2).mean().sqrt()),
            "surrogate_only": {"median_err_K_per_sol": float(perr.median()),
                               "p90_err_K_per_sol": float(perr.quantile(0.9)),
                               "max_err_K_per_sol": float(perr.max()),
                               "median_rel_err": float((perr / exact.abs().clamp(min=1.0)).median()),
                               "median_abs_heating_K_per_sol": float(exact.abs().median()),
                               "ms_one_network": ms_net,
                               "mfma_TFLOPs_one_network": flops * ncol / (ms_net / 1e3) / 1e12},
            "ablation_mse_standardised": {"trained_network": mse_net, "without_network_mean_predictor": mse_mean},
            "train_s": t_train, "final_train_loss": [hh[-1] for hh in hists]}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(columns are independent; the multi-GPU split is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    from mgcm_rt import FP64, MIXED, Radiation, alloc_outputs, synth, to_device

    tables, cols = synth.make_workload(args.seed, args.ncol, args.nlay)
    rt = Radiation(tables)
    dc = to_device(cols)
    do = to_device(alloc_outputs(args.ncol, args.nlay))
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    ms, wall_ms = timed_steps(lambda: rt.run_device(dc, do, FP64, sh), args.steps, args.warmup, stream)
    value = args.ncol / (ms / 1e3)

    fpc = OPS_PER_COLUMN.get((args.seed, args.nlay), {"flops": None, "exps": None, "sqrts": None})
    achieved = None if fpc["flops"] is None else fpc["flops"] * args.ncol / (ms / 1e3) / 1e12
    prof = profile_figures(args.ncol, args.nlay)
    traffic = prof["traffic"]
    got = {k: do[k].cpu().numpy() for k in PARITY_KEYS}
    result = {
        "metric": "radiative columns/sec (plain FP64 column radiation, 64,800 columns x 24 layers)",
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 generator, mgcm_rt/synth.py)",
        "config": {"workload": f"plain FP64 radiation step, {args.ncol} columns x {args.nlay} layers, "
                               "7 VIS + 5 IR bands x 17 Gauss points",
                   "ncol": args.ncol, "nlay": args.nlay, "seed": args.seed, "launches_per_step": 2,
                   "launches": "k_rt_clear2 (continuum point g=17, VIS and IR bands), k_rt (column kernel)"},
        "roofline": {"bound": "valu", "pipe": "FP64 vector ALU (no matrix work on this path); "
                                                "peak = MI355X FP64 vector 78.6 TFLOP/s",
                     "scope": "the whole step (its 2 launches); flops and traffic summed over them",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": None if achieved is None else achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "hbm_GBps": None if traffic is None else traffic / (ms / 1e3) / 1e9,
                     "flops_per_column": fpc["flops"], "exp_per_column": fpc["exps"],
                     "sqrt_per_column": fpc["sqrts"], "kernel_ms": ms,
                     "counter_fp64_TFLOPs": (None if prof["counter_fp64_flop"] is None
                                             else prof["counter_fp64_flop"] / (ms / 1e3) / 1e12),
                     "counter_note": "executed FP64 VALU operations per lane (SQ_INSTS_VALU_FLOPS_FP64 x 64: "
                                     "add/mul/transcendental 1, FMA 2), including the expansions of "
                                     "division, sqrt and exp; from the committed profile of the same library",
                     "profile_library_matches": prof["profile_library_matches"]},
        "wall_ms_per_step": wall_ms,
    }

    if args.copy_calib:  # known traffic for scaling the FETCH_SIZE/WRITE_SIZE counters
        x = torch.empty(64 * 2 ** 20, dtype=torch.float64, device="cuda")
        x.fill_(1.0)         # writes 512 MiB
        y = x.clone()        # reads 512 MiB, writes 512 MiB
        torch.cuda.synchronize()
        result["copy_calib_bytes"] = int(x.numel() * 8)
        del x, y

    # other forms; each is reported, or its error is, never substituted
    forms = {}
    if not args.no_forms:
        del do
        for name, fn in (("real_tables", form_real), ("host_arrays", form_host_arrays), ("small_grid", form_small_grid),
                         ("next_rows", form_next_rows),
                         ("mixed", form_mixed), ("mixed_real", form_mixed_real), ("batched", form_batched), ("surrogate", form_surrogate)):
            if args.forms and name not in args.forms.split(","):
                continue
            try:
                forms[name] = fn(args, rt, stream)
            except Exception as e:  # noqa: BLE001
                forms[name] = {"error": f"{type(e).__name__}: {e}"[:300]}
            torch.cuda.empty_cache()
    result["forms"] = forms

    if not args.no_cpu:
        refrun = {}
        try:
            result["cpu_baseline"] = cpu_baseline(args.seed, cols, args.cpu_sample, keep=refrun)
        except (FileNotFoundError, RuntimeError) as e:
            print(f"bench.py: reference build unavailable ({e}); timing the C restatement", file=sys.stderr)
            result["cpu_baseline"] = cpu_port(tables, _sample(cols, args.cpu_sample))
        if refrun:  # the device on the reference's own inputs (its Gauss weights), against its outputs
            try:
                rr = Radiation(refrun["tables"])
                dr = to_device(alloc_outputs(int(refrun["cols"]["ncol"]), args.nlay))
                rr.run_device(to_device(refrun["cols"]), dr, FP64, sh)
                torch.cuda.synchronize()
                rr.check()
                result["parity_reference"] = dict(
                    parity({k: dr[k].cpu().numpy() for k in PARITY_KEYS}, refrun["out"]),
                    against=f"the reference's Fortran (oracle/_ref) on the same {refrun['cols']['ncol']} columns, "
                            "its Gauss weights")
                del rr, dr
            except RuntimeError as e:
                result["parity_reference"] = {"error": f"{type(e).__name__}: {e}"[:300], "pass": False}
        want = {}
        result["cpu_port"] = cpu_port(tables, cols, keep=want)
        result["parity"] = parity(got, want)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
