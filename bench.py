#!/usr/bin/env python3
"""Throughput of the f-fdot acceleration search on one MI355X.

A "step" of this benchmark is one DM trial of the production workload:
2^23-sample float32 series (N(0,1) + red noise + 3 accelerated pulsars, already
resident in HBM), R2C FFT, whitening, every overlap-save step (1084 x 201 x 7734
f-fdot cells), every cell summed over 8 harmonics and tested against the stage
thresholds, candidates out (fdas_search; summed planes are stored only when a
caller asks for them with fdas_run_steps).

Prints one JSON line: cells/s (value), DM-trials/s, p50/p99 trial latency, the
roofline of the dominant kernel (HIP events on the plan's stream), and the CPU
oracle timed on a bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E


def alg_bytes(info, n: int) -> dict:
    """Algorithmic HBM bytes per DM trial, and per kernel category.

    Trial: series in 4N + spectrum out 8(N/2+1) + per step the segment reads
    4 L_h per harmonic, the sub-harmonic power planes written and re-read
    2*4*Z_h*R_h, and the summed fundamental plane written once 4*Z_0*U.
    Templates are resident constants and not counted.
    corr_sub:  segment reads + sub-plane writes.
    corr_fund: fundamental segment reads + fundamental power writes.
    harmsum:   sub-plane reads + fundamental plane read (the summed plane is
               not stored by fdas_search)."""
    H = info.numharm
    S = info.nsteps
    seg = S * sum(4 * info.fftlen[h] for h in range(H))
    sub = sum(info.sub_cells[h] for h in range(1, H))
    total = 4 * n + 8 * (n // 2 + 1) + seg + 2 * 4 * sub + 4 * info.cells
    return {
        "trial": total,
        "corr_sub": S * sum(4 * info.fftlen[h] for h in range(1, H)) + 4 * sub,
        "corr_fund": S * 4 * info.fftlen[0] + 4 * info.cells,
        "harmsum": 4 * sub + 4 * info.cells,
    }


def corr_flops(info) -> dict:
    """FFT-model FLOPs per DM trial of the fused correlation kernels: per
    (template, step) 5 L log2 L for the inverse transform + 6 L for the product,
    + 3 per kept power."""
    S, H = info.nsteps, info.numharm
    sub = sum(info.sub_cells[h] for h in range(1, H))
    f = lambda h: info.numzs[h] * (5 * info.fftlen[h] * np.log2(info.fftlen[h]) + 6 * info.fftlen[h])
    return {"corr_sub": S * sum(f(h) for h in range(1, H)) + 3 * sub,
            "corr_fund": S * f(0) + 3 * info.cells}


def measured_traffic(kernel: str):
    """HBM bytes per launch of a kernel category from the committed counter
    summary (two rocprofv3 --pmc runs of this script, FETCH_SIZE and WRITE_SIZE,
    scaled by the same counters on a copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "production_trial_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d["per_launch_bytes"][kernel]["total"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(series, n, zmax, numharm, fftlen, steps):
    """The CPU oracle (FP64 restatement, one core) on a bounded sample."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    ref = O.Search(series, n, 6.4e-5, zmax, numharm, fftlen)
    t_setup = time.perf_counter() - t0
    sel = np.linspace(0, ref.nsteps - 1, steps).astype(int)
    t0 = time.perf_counter()
    for st in sel:
        ref.step_plane(int(st))
    t_steps = time.perf_counter() - t0
    cells = len(sel) * ref.uselen * ref.numz
    ref.close()
    return {"value": cells / t_steps, "unit": "f-fdot cells/s", "cores": 1, "kind": "port",
            "sample": f"oracle/liboracle.so (FP64): spectrum+whitening+templates once "
                      f"({t_setup:.1f} s), then {len(sel)} of {ref.nsteps} steps spread over "
                      f"the trial ({t_steps:.1f} s); cells/s over the sampled steps"}


def stage_form(spec, n, zmax, numharm, fftlen, info, steps):
    """The reference's per-step stage sequence (AccelSearch.cu:217-241) on the GPU:
    host segment fetch, unfused forward FFTs, complex correlation planes written
    to memory, batched inverse FFTs, powers and harmonic sums, one step at a time."""
    import gpu_accelsearch_amd as fd
    fd.accel.fdas_set_uselen(info.uselen)
    try:
        bank = fd.createSubharmonicInfo(int(np.log2(numharm)) + 1, zmax)
        pipe = fd.StagePipeline(bank)
        sel = np.linspace(0, info.nsteps - 1, steps).astype(int)
        pipe.step(spec, info.rlo, -zmax, info.uselen, info.nbins)   # warm-up
        t0 = time.perf_counter()
        for st in sel:
            pipe.step(spec, info.rlo + int(st) * (info.uselen // 2), -zmax, info.uselen, info.nbins)
        dt = (time.perf_counter() - t0) / len(sel)
        pipe.close()
        bank.free()
    finally:
        fd.accel.fdas_set_uselen(32000)
    return {"ms_per_step": dt * 1e3, "ms_per_trial": dt * 1e3 * info.nsteps,
            "dm_trials_per_s": 1.0 / (dt * info.nsteps), "sampled_steps": int(len(sel)),
            "note": "per-trial time = sampled mean step time x steps of the trial"}


def batch_form(series, n, zmax, numharm, trials, precision, fftlen, reps=3):
    """Many DM trials in one plan, replayed from a captured hipGraph."""
    import gpu_accelsearch_amd as fd
    P = fd.SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen, precision=precision,
                        powcut=(40.0, 45.0, 55.0, 70.0, 0.0))
    with fd.BatchSearch(P, [n] * trials, graph=True) as b:
        for t in range(trials):
            b.upload(t, series)
        b.run()                       # eager run + capture
        b.run()                       # first replay
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        b.synchronize()
        dt = time.perf_counter() - t0
        return {"trials": trials, "precision": precision, "fftlen": fftlen, "graph": True,
                "dm_trials_per_s": trials * reps / dt, "ms_per_trial": dt * 1e3 / (trials * reps),
                "cells_per_s": b.info.cells * reps / dt, "device_bytes": int(b.info.device_bytes)}


def batch_workload_form(trials, precision, reps=1):
    """The batching workload: uneven lengths log-uniform in [2^21, 2^23], zmax 200,
    16 harmonics, autotuned fftlen, one plan replayed from a captured hipGraph."""
    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd import synth
    lengths, xs = synth.batch_workload(trials)
    P = fd.SearchParams(n=lengths[0], zmax=200, numharm=16, fftlen=0, precision=precision,
                        powcut=(35.0, 40.0, 50.0, 65.0, 90.0))
    rep = fd.autotune(P, fftlens=(4096, 8192, 16384))
    with fd.BatchSearch(P, lengths, graph=True) as b:
        for t, x in enumerate(xs):
            b.upload(t, x)
        b.run()                       # eager run + capture
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        b.synchronize()
        dt = time.perf_counter() - t0
        ncand = len(b.candidates())
        return {"trials": trials, "precision": precision, "fftlen": rep["fftlen"], "graph": True,
                "lengths": "log-uniform [2^21, 2^23], 10% with an injected accelerated sinusoid",
                "numharm": 16, "zmax": 200, "dm_trials_per_s": trials * reps / dt,
                "cells_per_s": b.info.cells * reps / dt, "ms_per_run": dt * 1e3 / reps,
                "candidates": ncand, "fftlen_ms_per_trial_at_first_length": rep["ms_per_trial"]}


def serving_form(count, rate, seed=6):
    """A bounded Poisson request stream through the continuous-batching server."""
    import threading
    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd import synth
    reqs = synth.serving_stream(count, rate, seed=seed)
    pool = synth.noise_pool(1 << 25)
    series = []
    for r in reqs:
        x = pool[r["offset"]:r["offset"] + r["n"]].copy()
        if r["pulsar"]:
            x += synth.pulse_train(r["n"], 0.37 * r["n"] / 2, 0.5 * r["zmax"], 0.05, 30.0).astype(np.float32)
        series.append(x)
    params = [fd.SearchParams(n=r["n"], zmax=r["zmax"], numharm=r["numharm"], fftlen=0,
                              precision=r["precision"], powcut=(40.0, 45.0, 55.0, 70.0, 90.0))
              for r in reqs]
    with fd.Server(max_active=4, chunk_steps=128) as srv:
        for z in (50, 200, 600):
            for h in (4, 8, 16):
                for prec in ("fp32", "i16"):
                    srv.prewarm(fd.SearchParams(n=1 << 20, zmax=z, numharm=h, fftlen=0, precision=prec))
        ids = [None] * count
        t0 = time.perf_counter()

        def feed():
            for i, (r, p, x) in enumerate(zip(reqs, params, series)):
                delay = r["t"] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[i] = srv.submit(p, x)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        res = [srv.wait(i)[0] for i in ids]
        wall = time.perf_counter() - t0
        st = srv.stats()
    cells = sum(r["cells"] for r in res)
    return {"requests": count, "offered_rate_per_s": rate, "wall_s": wall,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "dm_trials_per_s": st["trials_per_s"], "cells_per_s": cells / wall,
            "max_active_seen": st["max_active_seen"], "memory_peak_bytes": st["memory_peak"],
            "memory_budget_bytes": st["memory_bytes"], "failed": st["failed"],
            "mix": "n log-uniform [2^20, 2^25], zmax {50,200,600}, numharm {4,8,16}, "
                   "precision {fp32,fp16,bf16}, 5% with a pulsar; fftlen per request from the "
                   "autotuned/default table"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed DM trials")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=96, help="oracle steps sampled")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16", "i16"],
                    help="storage of spectra segments and templates (arithmetic is FP32)")
    ap.add_argument("--batch-trials", type=int, default=8,
                    help="DM trials of the batched reduced-precision form (0: skip)")
    ap.add_argument("--batch-fftlen", type=int, default=0,
                    help="fftlen of the batched reduced-precision form (0: autotuned among "
                         "4096, 8192 and 16384)")
    ap.add_argument("--stage-steps", type=int, default=12,
                    help="steps sampled for the reference-structured stage form (0: skip)")
    ap.add_argument("--batch-workload-trials", type=int, default=256,
                    help="trials of the uneven-length 16-harmonic batch workload (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the Poisson serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default correlation shapes (no autotuner launches)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB "
                         "(a kernel of known traffic for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for {args.gpus}); DM trials would "
              "shard across GPUs, one process each, with no exchange during a search", file=sys.stderr)
        return 2

    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd import AccelSearch, DeviceArray, SearchParams, synchronize
    from gpu_accelsearch_amd import synth

    n, zmax, numharm, fftlen = 1 << 23, 200, 8, 8192
    series = synth.production_series(n)
    params = SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen,
                          chunk_steps=args.chunk, precision=args.precision,
                          powcut=(40.0, 45.0, 55.0, 70.0, 0.0))
    # correlation launch shapes measured for this configuration's FFT lengths
    tune = None if args.no_tune else fd.autotune(params, fftlens=(fftlen,))
    s = AccelSearch(params)
    d_series = DeviceArray.from_host(series)
    for _ in range(args.warmup):
        s.search(d_series)
    s.synchronize()

    synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        s.search(d_series)
        s.synchronize()
        lat.append(time.perf_counter() - a)
    synchronize()
    wall = time.perf_counter() - t0

    cands = s.candidates()
    # per-kernel timing on the plan's stream (separate pass, events add records)
    s.enable_timing(True)
    prof_trials = 3
    for _ in range(prof_trials):
        s.search(d_series)
    tm = s.timing()
    s.enable_timing(False)

    info = s.info
    ab = alg_bytes(info, n)
    per_trial = wall / args.steps
    cells_per_s = info.cells / per_trial
    kern = {k: {"launches": v[0] // prof_trials, "ms_per_trial": v[1] / prof_trials}
            for k, v in tm.items()}
    dom = max(("corr_fund", "corr_sub", "harmsum"), key=lambda k: tm[k][1])
    dom_ms = tm[dom][1] / max(tm[dom][0], 1)
    dom_bytes = ab[dom] / max(tm[dom][0] // prof_trials, 1)
    achieved = dom_bytes / (dom_ms * 1e-3)
    out = {
        "metric": "f-fdot cells/s & DM-trials/s (2^23 pts, zmax200, 8 harm); p50/p99 latency",
        "value": cells_per_s,
        "unit": "f-fdot cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_trial * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": args.precision,
        "data": "synthetic: N(0,1) + red noise + 3 accelerated Gaussian-pulse pulsars, seed 2",
        "config": {"workload": "production DM trial: 2^23 samples, dt 64 us, zmax 200 "
                               "(201 templates), fftlen 8192 (uselen 7734), numharm 8, "
                               f"{info.nsteps} steps, {info.cells} cells",
                   "chunk_steps": info.chunk_steps},
        "dm_trials_per_s": 1.0 / per_trial,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": measured_traffic(dom),
                     "alg_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
        # FFT-model FLOP rate of the correlations against the FP32 vector peak (157.3
        # TFLOP/s counts packed FP32, which these kernels are built without, so 78.6 is
        # the reachable vector rate)
        "valu": {k: {"tflops": fl / (tm[k][1] / prof_trials * 1e-3) / 1e12,
                     "frac_of_157.3": fl / (tm[k][1] / prof_trials * 1e-3) / 157.3e12,
                     "frac_of_78.6_unpacked": fl / (tm[k][1] / prof_trials * 1e-3) / 78.65e12,
                     "flops_per_trial": fl}
                 for k, fl in corr_flops(info).items()},
        "trial_roofline": {"alg_bytes": ab["trial"],
                           "frac": ab["trial"] / per_trial / HBM_PEAK},
        "kernels": kern,
        "candidates": int(len(cands)),
        "autotune": None if tune is None else
        {"corr": {str(k): {"variant": v["variant"], "threads": v["threads"],
                           "steps_per_wg": v["steps_per_wg"], "split": v["split"],
                           "lds_form": v["lds_form"]}
                  for k, v in tune["corr"].items()},
         "harmsum": {k: tune["harmsum"][k] for k in ("threads", "rows", "cols")}},
        "fft_lengths": [h[2] for h in s.harmonics],
    }
    spec = s.spectrum() if args.stage_steps > 0 else None
    s.close()
    del d_series
    if args.stage_steps > 0:
        out["stage_form_fp32"] = stage_form(spec, n, zmax, numharm, fftlen, info, args.stage_steps)
    if args.batch_trials > 0:
        if args.batch_fftlen > 0:
            bt = {"fftlen": args.batch_fftlen, "ms_per_trial": None}
        else:
            rp = SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=0, precision="bf16")
            bt = fd.autotune(rp, fftlens=(4096, 8192, 16384))
        out["batch_bf16"] = batch_form(series, n, zmax, numharm, args.batch_trials, "bf16", bt["fftlen"])
        out["batch_bf16"]["fftlen_ms_per_trial"] = bt["ms_per_trial"]
        out["batch_i16"] = batch_form(series, n, zmax, numharm, args.batch_trials, "i16", bt["fftlen"])
        out["batch_fp32"] = batch_form(series, n, zmax, numharm, args.batch_trials, "fp32", fftlen)
        ref = out.get("stage_form_fp32", {}).get("dm_trials_per_s")
        out["batch_bf16"]["vs_fused_fp32_single"] = out["batch_bf16"]["dm_trials_per_s"] * per_trial
        if ref:
            out["batch_bf16"]["vs_stage_form_fp32"] = out["batch_bf16"]["dm_trials_per_s"] / ref
    if args.batch_workload_trials > 0:
        out["batch_workload_bf16"] = batch_workload_form(args.batch_workload_trials, "bf16")
    if args.serve_requests > 0:
        out["serving"] = serving_form(args.serve_requests, args.serve_rate)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(series, n, zmax, numharm, fftlen, args.cpu_steps)
        out["cpu_baseline"]["vs_gpu"] = cells_per_s / out["cpu_baseline"]["value"]
    if args.copy_probe_mib > 0:
        from gpu_accelsearch_amd import lib
        nb = args.copy_probe_mib << 20
        a, b = DeviceArray((nb,), np.uint8), DeviceArray((nb,), np.uint8)
        lib().fdas_memcpy_dtod(b.ptr, a.ptr, nb)
        synchronize()
        out["copy_probe_bytes"] = nb
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
