#!/usr/bin/env python3
"""Benchmark: pair-interactions per second per Heun step on one MI355X.

Default workload (the one-GPU polar-epithelium model): 200,000 cells on a jittered sphere of
radius sqrt(N/4pi)*0.8 with polarities, BinnedGabriel (c = 0.8, r_max = 1), PiecewiseLinear
(0.8, 2, 1) + bending + bidirectional polarity alignment, Heun dt = 0.05.

A step is one Heun step (grid rebuild + two pair sweeps with the fused predictor/corrector).
value = pair interactions (force-functor invocations on ordered pairs, both evaluations) in the
timed steps / timed wall time, state resident in HBM. The roofline entry is for the dominant
kernel class (a pair sweep), timed with HIP events on the engine's stream in a separate pass.
cpu_baseline is the C++/OpenMP oracle on the host cores, on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_cell(has_pol: bool, has_type: bool):
    """Compulsory HBM bytes per cell of the two Heun sweeps: s = state bytes (position 12,
    + polarity 8), a = attribute bytes (type 1).
    full sweep (substep 1): read s + a, write x~ (s) and F1 (s)        = 3s + a
    stale sweep (substep 2): read x~ (s) + a, x_n (s), F1 (s), write x_{n+1} (s) = 4s + a"""
    s = 12 + (8 if has_pol else 0)
    a = 1 if has_type else 0
    return {"full": 3 * s + a, "stale": 4 * s + a}


def cpu_baseline(w, seconds_target: float = 15.0):
    import numpy as np  # noqa: F401

    import oracle

    threads = int(os.environ.get("TECS_CPU_THREADS", min(16, os.cpu_count() or 1)))
    oracle.set_threads(threads)
    o = oracle.OracleSim(w.desc)
    o.upload(w.positions, w.polarities, w.types)
    o.take_step(w.dt, 1)  # warm-up
    s0 = o.stats()["pair_interactions"]
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.take_step(w.dt, 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds_target or steps >= 1000:
            break
    pairs = o.stats()["pair_interactions"] - s0
    return {
        "value": pairs / el,
        "unit": "pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} Heun steps of the same {w.n}-cell workload after 1 warm-up step "
        f"({el:.1f} s, C++/OpenMP oracle, {threads} threads)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the streaming copy kernel used to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs each simulation on a single device; "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from tecs_amd import GpuSimulation, abi, device_count, workloads

    if device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        import ctypes

        from tecs_amd import _lib

        ms = ctypes.c_double(0.0)
        nbytes = 1 << 30
        _lib.check(_lib.load().tecs_copy_probe(nbytes, 5, ctypes.byref(ms)))
        print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes,
                          "launches": 6, "mean_ms": ms.value,
                          "GB/s (read+write)": 2 * nbytes / (ms.value * 1e-3) / 1e9}))
        return
    w = workloads.epithelium(args.n)
    sim = GpuSimulation(w.desc)
    sim.upload(w.positions, w.polarities, w.types)
    print(f"bench: epithelium n={w.n} warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
    sim.take_step(w.dt, args.warmup)
    sim.synchronize()
    sim.reset_stats()
    t0 = time.perf_counter()
    sim.take_step(w.dt, args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    st = sim.stats()
    pairs = st["pair_interactions"]
    value = pairs / elapsed
    ms_per_step = 1e3 * elapsed / max(args.steps, 1)
    print(f"bench: timed {args.steps} steps in {elapsed:.3f} s, {pairs} pair interactions",
          file=sys.stderr, flush=True)

    # kernel timing pass (HIP events on the engine's stream), same workload continued
    k_steps = min(max(args.steps // 4, 20), 500)
    sim.set_kernel_timing(True)
    sim.take_step(w.dt, k_steps)
    sim.synchronize()
    full_ms, full_n = sim.kernel_timing(abi.KT_SWEEP_FULL)
    stale_ms, stale_n = sim.kernel_timing(abi.KT_SWEEP_STALE)
    reb_ms, reb_n = sim.kernel_timing(abi.KT_REBUILD)
    sim.set_kernel_timing(False)
    bpc = algorithmic_bytes_per_cell(True, False)
    per = {
        "full": (full_ms / max(full_n, 1), bpc["full"]),
        "stale": (stale_ms / max(stale_n, 1), bpc["stale"]),
    }
    dom = max(per, key=lambda k: per[k][0])
    avg_ms, b_cell = per[dom]
    achieved = (b_cell * w.n) / (avg_ms * 1e-3) / 1e9
    variant = abi.SWEEP_NAMES.get(sim.sweep_variant(), "?")
    kname = "k_sweep_binned" if variant == "thread" else f"k_sweep_team<{variant[4:]}>"
    kernel_names = {"full": f"{kname} Gabriel polarity mode Pred (Heun substep 1)",
                    "stale": f"{kname} Gabriel polarity mode Corr (Heun substep 2, stale grid)"}
    print(
        f"bench: sweep full {per['full'][0]*1e3:.1f} us, stale {per['stale'][0]*1e3:.1f} us, "
        f"rebuild {reb_ms/max(reb_n,1)*1e3:.1f} us per launch",
        file=sys.stderr,
        flush=True,
    )

    out = {
        "metric": "pair-interactions/sec per Heun step",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {
            "workload": "polar epithelium: 200k cells on a sphere with polarity, BinnedGabriel "
            "c=0.8 r_max=1, PiecewiseLinear(0.8,2,1)+bending+alignment, Heun dt=0.05",
            "cells": w.n,
            "pair_interactions_per_step": pairs / max(args.steps, 1),
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_names[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "avg_launch_us": avg_ms * 1e3,
            "algorithmic_bytes_per_cell": b_cell,
            "sweep_full_us": per["full"][0] * 1e3,
            "sweep_stale_us": per["stale"][0] * 1e3,
            "rebuild_us": reb_ms / max(reb_n, 1) * 1e3,
            "sweep_variant": variant,
            "autotune_ms": sim.autotune_ms(),
        },
    }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        cb = cpu_baseline(w, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
