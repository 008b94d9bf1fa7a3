#!/usr/bin/env python3
"""Throughput of the f-fdot acceleration search on one MI355X.

A "step" of this benchmark is one DM trial of the production workload:
2^23-sample float32 series (N(0,1) + red noise + 3 accelerated pulsars, already
resident in HBM), R2C FFT, whitening, every overlap-save step (1084 x 201 x 7734
f-fdot cells) with 8-harmonic summing, stage thresholds and the summed plane.

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
    harmsum:   sub-plane reads + fundamental plane read + summed plane write."""
    H = info.numharm
    S = info.nsteps
    seg = S * sum(4 * info.fftlen[h] for h in range(H))
    sub = sum(info.sub_cells[h] for h in range(1, H))
    total = 4 * n + 8 * (n // 2 + 1) + seg + 2 * 4 * sub + 4 * info.cells
    return {
        "trial": total,
        "corr_sub": S * sum(4 * info.fftlen[h] for h in range(1, H)) + 4 * sub,
        "corr_fund": S * 4 * info.fftlen[0] + 4 * info.cells,
        "harmsum": 4 * sub + 8 * info.cells,
    }


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed DM trials")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle steps sampled")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="storage of spectra segments and templates (arithmetic is FP32)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB "
                         "(a kernel of known traffic for scaling profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); "
              "DM trials shard across GPUs with no exchange, see DESIGN.md", file=sys.stderr)
        return 2

    from gpu_accelsearch_amd import AccelSearch, DeviceArray, SearchParams, synchronize
    from gpu_accelsearch_amd import synth

    n, zmax, numharm, fftlen = 1 << 23, 200, 8, 8192
    series = synth.production_series(n)
    params = SearchParams(n=n, zmax=zmax, numharm=numharm, fftlen=fftlen,
                          chunk_steps=args.chunk, precision=args.precision,
                          powcut=(40.0, 45.0, 55.0, 70.0, 0.0))
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "alg_bytes_per_launch": dom_bytes, "avg_launch_ms": dom_ms},
        "trial_roofline": {"alg_bytes": ab["trial"],
                           "frac": ab["trial"] / per_trial / HBM_PEAK},
        "kernels": kern,
        "candidates": int(len(cands)),
    }
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
    s.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
