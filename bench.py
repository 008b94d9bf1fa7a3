#!/usr/bin/env python3
"""Throughput of the f-fdot acceleration search on one MI355X.

A "step" of this benchmark is one DM trial of the production workload:
2^23-sample float32 series (N(0,1) + red noise + 3 accelerated pulsars, already
resident in HBM), R2C FFT, whitening, every overlap-save step (1084 x 201 x 7734
f-fdot cells), every cell summed over 8 harmonics and tested against the stage
thresholds, candidates out (fdas_search; summed planes are stored only when a
caller asks for them with fdas_run_steps).  `value` is that FP32 form's cells/s.

Beside it, on the same box: the batched reduced-precision form (I16 segments and
templates, P16 power planes, hipGraph replay) against the fused FP32 form; the
16-harmonic uneven-length batch workload; the serving loop on a Poisson request
stream (at least 500 requests) and a rate sweep; the reference-structured stage
form; and the CPU baseline (the FP32 CPU form on all host cores over the whole
trial, the reference's CPU stages timed on their own, and the FP64 oracle).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12        # bytes/s, MI355X HBM3E
VALU_PEAK = 157.3e12     # FP32 vector FLOP/s, MI355X published peak
POWCUT = (40.0, 45.0, 55.0, 70.0, 0.0)


def alg_bytes(info, n: int, pbytes: int = 4) -> dict:
    """Algorithmic HBM bytes per DM trial, and per kernel category.

    Trial: series in 4N + spectrum out 8(N/2+1) + per step the segment reads
    4 L_h per harmonic, the sub-harmonic power planes written and re-read
    2*b*Z_h*R_h, and the fundamental plane written once b*Z_0*U (b = bytes per
    stored power: 4 in FP32, 2 with P16 planes).  Templates are resident constants
    and not counted.
    corr_sub:  segment reads + sub-plane writes.
    corr_fund: fundamental segment reads + fundamental power writes.
    harmsum:   sub-plane reads + fundamental plane read (the summed plane is
               not stored by fdas_search)."""
    H = info.numharm
    S = info.nsteps
    seg = S * sum(4 * info.fftlen[h] for h in range(H))
    sub = sum(info.sub_cells[h] for h in range(1, H))
    total = 4 * n + 8 * (n // 2 + 1) + seg + 2 * pbytes * sub + pbytes * info.cells
    return {
        "trial": total,
        "corr_sub": S * sum(4 * info.fftlen[h] for h in range(1, H)) + pbytes * sub,
        "corr_fund": S * 4 * info.fftlen[0] + pbytes * info.cells,
        "harmsum": pbytes * sub + pbytes * info.cells,
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


def measured_traffic(kernel: str, precision: str = "fp32"):
    """HBM bytes per launch of a kernel category from the committed counter
    summary of the production trial in this storage mode (rocprofv3 --pmc
    FETCH_SIZE, then WRITE_SIZE, scaled by the same counters on a copy of known
    size), or None where no summary was taken."""
    name = {"fp32": "production_trial_traffic.json",
            "i16": "production_trial_i16_traffic.json"}.get(precision)
    if name is None:
        return None
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            d = json.load(f)
        return d["per_launch_bytes"][kernel]["total"]
    except (OSError, KeyError, ValueError):
        return None


def host_cores() -> int:
    return len(os.sched_getaffinity(0))


def cpu_baseline(series, n, zmax, numharm, fftlen, fp64_steps):
    """FP32 CPU form (oracle/fdas_cpu32.c) over the whole trial on all host cores,
    the reference's CPU stages timed on their own, and the FP64 oracle on a
    contiguous block of steps (all cores) as a secondary figure."""
    from oracle import oracle as O
    cores = host_cores()
    t0 = time.perf_counter()
    c = O.Cpu32(series, n, 6.4e-5, zmax, numharm, fftlen)
    t_create = time.perf_counter() - t0
    steps = np.arange(c.nsteps)
    wall, hits, t_norm = c.run_steps(steps, POWCUT, threads=cores)
    cells = c.nsteps * c.uselen * c.numz
    per_trial = c.seconds["r2c"] + c.seconds["whiten"] + wall
    nsteps = c.nsteps
    c.close()
    ref = O.Search(series, n, 6.4e-5, zmax, numharm, fftlen)
    s0 = ref.nsteps // 2
    t0 = time.perf_counter()
    ref.candidates(POWCUT, 0.0, s0, fp64_steps)
    t64 = time.perf_counter() - t0
    cells64 = fp64_steps * ref.uselen * ref.numz
    ref.close()
    return {"value": cells / per_trial, "unit": "f-fdot cells/s", "cores": cores, "kind": "port",
            "impl": "fp32-omp",
            "sample": f"the whole trial on {cores} host cores: FP32 R2C + whitening + all "
                      f"{nsteps} steps (radix-4 Stockham FFTs, OpenMP over steps); the template "
                      "bank, a resident constant as on the GPU, is timed apart",
            "dm_trials_per_s": 1.0 / per_trial, "ms_per_trial": per_trial * 1e3,
            "cells_above_cuts": int(hits),
            "reference_cpu_stages_s": {
                "deredden": c_seconds(c, "whiten"), "template_bank": c_seconds(c, "templates"),
                "segment_normalisation_thread_s_per_trial": t_norm,
                "segment_normalisation_s_per_trial_on_all_cores": t_norm / cores,
                "r2c_fp32": c_seconds(c, "r2c"), "setup_wall": t_create},
            "fp64_oracle": {"value": cells64 / t64, "unit": "f-fdot cells/s", "cores": cores,
                            "sample": f"steps [{s0}, {s0 + fp64_steps}) of the trial, FP64 "
                                      "FFT correlation and harmonic sum, OpenMP over steps"}}


def c_seconds(c, k):
    return float(c.seconds[k])


def stage_form(spec, zmax, numharm, info, steps):
    """The reference's per-step stage sequence (AccelSearch.cu:217-241) through
    the reference-named entry points: host segment fetch, normalise-and-spread,
    one H2D copy, unfused forward FFTs, complex correlation planes written to
    memory, batched inverse FFTs, powers and harmonic sums, one step at a time
    (no per-step allocation; the plane stays on the device)."""
    import gpu_accelsearch_amd as fd
    fd.accel.fdas_set_uselen(info.uselen)
    try:
        bank = fd.createSubharmonicInfo(int(np.log2(numharm)) + 1, zmax)
        pipe = fd.StagePipeline(bank)
        sel = np.linspace(0, info.nsteps - 1, steps).astype(int)
        pipe.step(spec, info.rlo, -zmax, info.uselen, info.nbins, to_host=False)   # warm-up
        fd.synchronize()
        t0 = time.perf_counter()
        for st in sel:
            pipe.step(spec, info.rlo + int(st) * (info.uselen // 2), -zmax, info.uselen,
                      info.nbins, to_host=False)
        fd.synchronize()
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
                        powcut=POWCUT)
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


def _pct(v, q):
    return float(np.percentile(v, q)) if len(v) else None


def serving_run(srv, pool, reqs, rate):
    """One Poisson stream through a running server; per-request latencies."""
    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd import synth
    srv.reset_stats()
    ids = [None] * len(reqs)
    t0 = time.perf_counter()

    def feed():
        for i, r in enumerate(reqs):
            delay = r["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            x = pool[r["offset"]:r["offset"] + r["n"]]
            if r["pulsar"]:
                x = x + synth.pulse_train(r["n"], 0.37 * r["n"] / 2, 0.5 * r["zmax"], 0.05,
                                          30.0).astype(np.float32)
            p = fd.SearchParams(n=r["n"], zmax=r["zmax"], numharm=r["numharm"], fftlen=0,
                                precision=r["precision"], powcut=(40.0, 45.0, 55.0, 70.0, 90.0))
            ids[i] = srv.submit(p, x)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    res = [srv.wait(i)[0] for i in ids]
    wall = time.perf_counter() - t0
    st = srv.stats()
    lat = np.array([r["latency_ms"] for r in res])
    out = {"requests": len(reqs), "offered_rate_per_s": rate, "wall_s": wall,
           "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
           "dm_trials_per_s": len(reqs) / wall, "cells_per_s": sum(r["cells"] for r in res) / wall,
           "max_active_seen": st["max_active_seen"], "memory_peak_bytes": st["memory_peak"],
           "failed": st["failed"]}
    classes = {}
    for lo, hi, name in ((1 << 20, 1 << 22, "n<2^22"), (1 << 22, 1 << 24, "2^22<=n<2^24"),
                         (1 << 24, (1 << 25) + 1, "n>=2^24")):
        sel = [k for k, r in enumerate(reqs) if lo <= r["n"] < hi]
        classes[name] = {"requests": len(sel), "p50_ms": _pct(lat[sel], 50), "p99_ms": _pct(lat[sel], 99)}
    for z in (50, 200, 600):
        sel = [k for k, r in enumerate(reqs) if r["zmax"] == z]
        classes[f"zmax{z}"] = {"requests": len(sel), "p50_ms": _pct(lat[sel], 50),
                               "p99_ms": _pct(lat[sel], 99)}
    out["latency_by_class"] = classes
    return out


def serving_form(count, rate, sweep, sweep_count, seed=6):
    """The serving workload through the continuous-batching server: a Poisson
    stream of `count` requests at `rate` per second (p50/p99 over them), then
    shorter streams at each rate of `sweep`; the completed requests per second at
    the highest offered rates give the server's capacity."""
    import gpu_accelsearch_amd as fd
    from gpu_accelsearch_amd import synth
    pool = synth.noise_pool(1 << 25)
    with fd.Server(max_active=4, chunk_steps=128) as srv:
        for z in (50, 200, 600):
            for h in (4, 8, 16):
                for prec in ("fp32", "i16"):
                    srv.prewarm(fd.SearchParams(n=1 << 20, zmax=z, numharm=h, fftlen=0, precision=prec))
        main = serving_run(srv, pool, synth.serving_stream(count, rate, seed=seed), rate)
        runs = []
        for k, r in enumerate(sweep):
            # the same request mix at every rate (arrival times scale with 1/rate)
            s = serving_run(srv, pool, synth.serving_stream(sweep_count, r, seed=seed + 10), r)
            runs.append({key: s[key] for key in ("offered_rate_per_s", "requests", "wall_s",
                                                 "dm_trials_per_s", "p50_ms", "p99_ms",
                                                 "max_active_seen", "failed")})
        # the same mix one request at a time (submit, wait): the device time a
        # server without overlap would need, against which the capacity is judged
        srv.reset_stats()
        reqs = synth.serving_stream(sweep_count, 1.0, seed=seed + 10)
        t0 = time.perf_counter()
        for r in reqs:
            x = pool[r["offset"]:r["offset"] + r["n"]]
            if r["pulsar"]:
                x = x + synth.pulse_train(r["n"], 0.37 * r["n"] / 2, 0.5 * r["zmax"], 0.05,
                                          30.0).astype(np.float32)
            p = fd.SearchParams(n=r["n"], zmax=r["zmax"], numharm=r["numharm"], fftlen=0,
                                precision=r["precision"], powcut=(40.0, 45.0, 55.0, 70.0, 90.0))
            srv.wait(srv.submit(p, x))
        serial = time.perf_counter() - t0
        budget = srv.stats()["memory_bytes"]
    main["memory_budget_bytes"] = budget
    main["rate_sweep"] = runs
    main["capacity_dm_trials_per_s"] = max(r["dm_trials_per_s"] for r in runs + [main])
    main["serial_dm_trials_per_s"] = sweep_count / serial
    main["overlap_gain"] = main["capacity_dm_trials_per_s"] / main["serial_dm_trials_per_s"]
    main["mix"] = ("n log-uniform [2^20, 2^25], zmax {50,200,600}, numharm {4,8,16}, precision "
                   "{fp32, i16}, 5% with a pulsar; fftlen per request from the autotuned/default "
                   "table; max_active 4, 128 steps per launch group")
    return main


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed DM trials")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=64,
                    help="contiguous steps of the FP64 oracle in the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16", "i16", "i16mma"],
                    help="storage of the single-trial form (arithmetic is FP32)")
    ap.add_argument("--batch-trials", type=int, default=8,
                    help="DM trials of the batched forms (0: skip)")
    ap.add_argument("--batch-fftlen", type=int, default=0,
                    help="fftlen of the batched reduced-precision form (0: autotuned among "
                         "4096, 8192 and 16384)")
    ap.add_argument("--stage-steps", type=int, default=24,
                    help="steps sampled for the reference-structured stage form (0: skip)")
    ap.add_argument("--batch-workload-trials", type=int, default=256,
                    help="trials of the uneven-length 16-harmonic batch workload (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=500,
                    help="requests of the Poisson serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=50.0)
    ap.add_argument("--serve-sweep", default="20,100,200",
                    help="further offered rates (req/s) for the capacity sweep")
    ap.add_argument("--serve-sweep-requests", type=int, default=150)
    ap.add_argument("--lengths", default="shortest", choices=["shortest", "reference"],
                    help="sub-harmonic FFT lengths: the shortest that fit, or the "
                         "reference's powers of two")
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
                          chunk_steps=args.chunk, precision=args.precision, powcut=POWCUT,
                          fft_lengths=args.lengths)
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
    pbytes = 4 if args.precision == "fp32" else 2
    ab = alg_bytes(info, n, pbytes)
    per_trial = wall / args.steps
    cells_per_s = info.cells / per_trial
    kern = {k: {"launches": v[0] // prof_trials, "ms_per_trial": v[1] / prof_trials}
            for k, v in tm.items()}
    fl = corr_flops(info)

    def per_launch(cat):
        launches = max(tm[cat][0] // prof_trials, 1)
        ms = tm[cat][1] / max(tm[cat][0], 1)
        return launches, ms

    # the single largest kernel of the trial: the harmonic sum (one launch per
    # chunk; the correlation categories are several launches of different kernels)
    hs_launches, hs_ms = per_launch("harmsum")
    hs_bytes = ab["harmsum"] / hs_launches
    kernel_rooflines = {"harmsum": {
        "bound": "hbm", "kernel": "harmsum_kernel", "launches_per_trial": hs_launches,
        "avg_launch_ms": hs_ms, "alg_bytes_per_launch": hs_bytes,
        "achieved_GBps": hs_bytes / (hs_ms * 1e-3) / 1e9,
        "frac": hs_bytes / (hs_ms * 1e-3) / HBM_PEAK}}
    for cat, kname in (("corr_fund", "corr_wave_kernel (fundamental)"),
                       ("corr_sub", "corr_odd/corr_wave_kernel (sub-harmonics)")):
        launches, ms = per_launch(cat)
        flops = fl[cat] / launches
        by = ab[cat] / launches
        rate = flops / (ms * 1e-3)
        kernel_rooflines[cat] = {
            "bound": "valu", "kernel": kname, "launches_per_trial": launches,
            "avg_launch_ms": ms, "flops_per_launch": flops, "achieved_TFLOPs": rate / 1e12,
            "frac": rate / VALU_PEAK,
            "hbm_GBps": by / (ms * 1e-3) / 1e9, "hbm_frac": by / (ms * 1e-3) / HBM_PEAK}
    rf = kernel_rooflines["harmsum"]
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
                   "chunk_steps": info.chunk_steps, "fft_lengths": args.lengths},
        "dm_trials_per_s": 1.0 / per_trial,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                       "p99": float(np.percentile(lat, 99) * 1e3)},
        "roofline": {"bound": "hbm", "kernel": "harmsum_kernel", "achieved": rf["achieved_GBps"],
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": rf["frac"],
                     "traffic": measured_traffic("harmsum", args.precision),
                     "alg_bytes_per_launch": hs_bytes, "avg_launch_ms": hs_ms},
        "kernel_rooflines": kernel_rooflines,
        "trial_roofline": {"alg_bytes": ab["trial"], "frac": ab["trial"] / per_trial / HBM_PEAK},
        "kernels": kern,
        "candidates": int(len(cands)),
        "autotune": None if tune is None else
        {"corr": {str(k): {"variant": v["variant"], "threads": v["threads"],
                           "steps_per_wg": v["steps_per_wg"], "split": v["split"],
                           "lds_form": v["lds_form"]}
                  for k, v in tune["corr"].items()},
         "harmsum": {k: tune["harmsum"][k] for k in ("threads", "rows", "cols")}},
        "fft_lengths": [h[2] for h in s.harmonics],
        "host_cores": host_cores(),
    }
    spec = s.spectrum() if args.stage_steps > 0 else None
    s.close()
    del d_series
    if args.batch_trials > 0:
        if args.batch_fftlen > 0:
            bt = {"fftlen": args.batch_fftlen, "ms_per_trial": None}
        else:
            rp = SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=0, precision="i16")
            bt = fd.autotune(rp, fftlens=(4096, 8192, 16384))
        bi = batch_form(series, n, zmax, numharm, args.batch_trials, "i16", bt["fftlen"])
        bi["fftlen_ms_per_trial"] = bt["ms_per_trial"]
        bi["storage"] = "I16 segments and templates, P16 power planes (within the 1e-3 tolerance)"
        bf = batch_form(series, n, zmax, numharm, args.batch_trials, "fp32", fftlen)
        # the direct correlation on the matrix cores, same batch and length
        bm = batch_form(series, n, zmax, numharm, args.batch_trials, "i16mma", fftlen)
        bm["storage"] = ("I16_MMA: time-domain 15-bit segments and templates, exact int8-MFMA "
                         "correlation, P16 power planes (within the 1e-3 tolerance)")
        out["batch_i16"] = bi
        out["batch_fp32"] = bf
        out["batch_i16mma"] = bm
        out["ratio_batched_reduced_vs_fused_fp32_single"] = bi["dm_trials_per_s"] * per_trial
        out["ratio_batched_reduced_vs_batched_fp32"] = bi["dm_trials_per_s"] / bf["dm_trials_per_s"]
    if args.stage_steps > 0:
        out["stage_form_fp32"] = stage_form(spec, zmax, numharm, info, args.stage_steps)
        if "batch_i16" in out:
            out["ratio_batched_reduced_vs_stage_form_fp32"] = (
                out["batch_i16"]["dm_trials_per_s"] / out["stage_form_fp32"]["dm_trials_per_s"])
    if args.batch_workload_trials > 0:
        out["batch_workload_i16"] = batch_workload_form(args.batch_workload_trials, "i16")
    if args.serve_requests > 0:
        sweep = [float(v) for v in args.serve_sweep.split(",") if v]
        out["serving"] = serving_form(args.serve_requests, args.serve_rate, sweep,
                                      args.serve_sweep_requests)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(series, n, zmax, numharm, fftlen, args.fp64_steps)
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
