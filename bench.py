#!/usr/bin/env python3
"""Benchmark: pair-interactions per second per Heun step on one MI355X, batched vs plain.

Default workload (the batched parameter sweep): 256 independent simulations, cell counts
log-uniform in [1k, 100k], relaxation (ball) or sorting (cube, A/B types) geometry, tau_AB
linear in [0.1, 1], BinnedAllPairs r_max = 1, Heun dt = 0.05. A step is one Heun step of every
instance: one launch per kernel for all of them, replayed from captured hipGraphs.
value = pair interactions (force-functor invocations on ordered pairs, both evaluations) of
all instances in the timed steps / timed wall time, state resident in HBM. The same instances
are then run one simulation at a time ("plain"), each warmed past its graph capture, and the
ratio is reported. A short served run (Poisson arrivals of the three models, continuous
batching under an HBM budget) adds p50/p99 request latencies.

The roofline entry is for the dominant kernel class of the batched step (a pair sweep), timed
with HIP events on the batch's stream in a pass right after the timed steps. cpu_baseline is
the C++/OpenMP oracle on the host cores allowed to this process, on a bounded stratified sample
of the same instances.
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


def measured_traffic(workload: str, variant: str, phase: str):
    """HBM bytes per launch of a sweep as rocprofv3 measured them: profiles/traffic.txt holds
    FETCH_SIZE and WRITE_SIZE per dispatch, already scaled by the factors the same counters
    read for the k_copy_probe streaming copy. None when the file has no line for this
    workload, sweep variant and phase."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split("#", 1)[0].split()
            if len(parts) == 5 and parts[:3] == [workload, variant, phase]:
                return {"fetch": int(parts[3]), "write": int(parts[4])}
    return None


# VALU issue peak from the published FP32 vector peak (157.3 TFLOP/s = 256 CUs x 128 FMA lanes
# x 2 x 2.4 GHz): 4 SIMDs of 32 lanes per CU, a wavefront64 instruction in 2 cycles. (At 4 cycles
# the sweeps' counted instructions would need more than their measured launch times.)
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 2


def measured_valu(workload: str, variant: str, phase: str):
    """VALU wavefront-instructions per launch of a sweep as rocprofv3 counted them
    (profiles/valu.txt), or None."""
    path = os.path.join(ROOT, "profiles", "valu.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split("#", 1)[0].split()
            if len(parts) == 4 and parts[:3] == [workload, variant, phase]:
                return int(parts[3])
    return None


def algorithmic_bytes_per_cell(has_pol: bool, has_type: bool):
    """Compulsory HBM bytes per cell of the two Heun sweeps: s = state bytes (position 12,
    + polarity 8), a = attribute bytes (type 1).
    full sweep (substep 1): read s + a, write x~ (s) and F1 (s)        = 3s + a
    stale sweep (substep 2): read x~ (s) + a, x_n (s), F1 (s), write x_{n+1} (s) = 4s + a"""
    s = 12 + (8 if has_pol else 0)
    a = 1 if has_type else 0
    return {"full": 3 * s + a, "stale": 4 * s + a}


_ALLOWED_CPUS = []  # the process's CPU count, taken before the OpenMP runtime binds this thread


def _cpu_setup():
    """Thread count and binding of the CPU baseline: every core this process may run on
    (OMP_NUM_THREADS when the environment sets it), OMP_PROC_BIND=close, OMP_PLACES=cores,
    set before the OpenMP runtime of the oracle library starts. Returns (threads, host info).
    The count is taken once: with OMP_PROC_BIND the runtime pins the calling thread to one core,
    after which its affinity mask no longer shows the process's CPUs."""
    if not _ALLOWED_CPUS:
        try:
            _ALLOWED_CPUS.append(len(os.sched_getaffinity(0)))
        except AttributeError:
            _ALLOWED_CPUS.append(os.cpu_count() or 1)
    allowed = _ALLOWED_CPUS[0]
    threads = int(os.environ.get("TECS_CPU_THREADS", os.environ.get("OMP_NUM_THREADS", allowed)))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    import oracle

    oracle.set_threads(threads)
    host = {"cpu_model": model, "cpus_allowed": allowed, "omp_proc_bind": os.environ["OMP_PROC_BIND"],
            "omp_places": os.environ["OMP_PLACES"]}
    return threads, host


def _median(xs):
    a = sorted(xs)
    m = len(a) // 2
    return a[m] if len(a) % 2 else 0.5 * (a[m - 1] + a[m])


def cpu_baseline(w, seconds_target: float = 15.0, warmup: int = 10, min_steps: int = 100):
    """The oracle on one workload: `warmup` untimed steps, then timed steps until
    seconds_target has passed and at least min_steps ran (or 1000 steps); value = the median of
    the per-step rates (pair interactions of the step / its time)."""
    threads, host = _cpu_setup()
    import oracle
    from tecs_amd import abi

    o = oracle.OracleSim(w.desc)
    o.upload(w.positions, w.polarities, w.types)
    o.take_step(w.dt, warmup)
    rates = []
    t0 = time.perf_counter()
    while True:
        before = o.stats()["pair_interactions"]
        t1 = time.perf_counter()
        o.take_step(w.dt, 1)
        dt_s = time.perf_counter() - t1
        rates.append((o.stats()["pair_interactions"] - before) / dt_s)
        el = time.perf_counter() - t0
        if (el >= seconds_target and len(rates) >= min_steps) or len(rates) >= 1000 or el >= 4 * seconds_target:
            break
    return {
        "value": _median(rates),
        "unit": "pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median over {len(rates)} timed {'Heun' if int(w.desc.integrator) == abi.HEUN else 'Euler'} "
        f"steps of the same {w.n}-cell workload after "
        f"{warmup} warm-up steps ({el:.1f} s, C++/OpenMP oracle, {threads} threads)",
        **host,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-variant", default="auto", choices=["auto", "thread", "team4", "team8", "team16"],
                    help="pin the pair-sweep kernel variant instead of autotuning it")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the streaming copy kernel used to calibrate HBM byte counters")
    ap.add_argument("--workload", default="sweep",
                    choices=["epithelium", "relaxation", "sorting", "sweep", "serve", "spheroid", "queries",
                             "snapshots"],
                    help="sweep (default): 256 mixed-size instances, batched vs one simulation at a "
                    "time, plus a short served run; epithelium: 200k cells, BinnedGabriel + polarity, "
                    "Heun; relaxation: 10k cells, BinnedAllPairs, Euler; sorting: 4M cells, "
                    "BinnedAllPairs + type table, Heun; serve: served requests only; spheroid: NaiveDelaunay spheroids of 100 cells, "
                    "batched; queries: grid queries on the epithelium's grid; snapshots: the sorting "
                    "tissue stepped with HDF5 snapshots, asynchronous hand-off vs download + write")
    ap.add_argument("--instances", type=int, default=256, help="spheroid workload: batch size")
    ap.add_argument("--walls", action="store_true",
                    help="epithelium / relaxation / sorting with a generic wall force (a floor just "
                    "inside the lowest cells, strength 1): the step takes the separate-update path")
    ap.add_argument("--links", type=int, default=0,
                    help="epithelium / relaxation / sorting with this many links between random cells "
                    "and a random neighbour within distance 1 (strength 0.1)")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="sweep workload: steps per instance for the one-at-a-time comparison")
    ap.add_argument("--requests", type=int, default=48, help="served requests: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="served requests: arrivals per second")
    ap.add_argument("--no-plain", action="store_true",
                    help="sweep / serve: skip the one-simulation-at-a-time runs")
    ap.add_argument("--no-serve", action="store_true", help="sweep: skip the served run")
    ap.add_argument("--snapshot-every", type=int, default=50,
                    help="snapshots workload: steps between snapshots")
    ap.add_argument("--symmetric", action="store_true",
                    help="epithelium / relaxation / sorting with the symmetric variant of their pair "
                    "finder (one functor call per unordered pair)")
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
    if args.workload == "sweep":
        return bench_sweep(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "spheroid":
        return bench_spheroid(args)
    if args.workload == "queries":
        return bench_queries(args)
    if args.workload == "snapshots":
        return bench_snapshots(args)
    sym = {"finder": abi.BINNED_GABRIEL_SYMMETRIC if args.workload == "epithelium"
           else abi.BINNED_ALL_PAIRS_SYMMETRIC} if args.symmetric else {}
    if args.workload == "relaxation":
        w = workloads.relaxation(10_000 if args.n == 200_000 else args.n, **sym)
    elif args.workload == "sorting":
        w = workloads.sorting(4_000_000 if args.n == 200_000 else args.n, **sym)
    else:
        w = workloads.epithelium(args.n, **sym)
    return bench_single(args, w)


WORKLOAD_TEXT = {
    "epithelium": "polar epithelium: {n} cells on a sphere with polarity, BinnedGabriel c=0.8 "
    "r_max=1, PiecewiseLinear(0.8,2,1)+bending+alignment, Heun dt=0.05",
    "relaxation": "passive relaxation: {n} cells uniform in a ball of radius 0.5*cbrt(N), "
    "BinnedAllPairs r_max=1, PiecewiseLinear(0.8,2,1), Euler dt=0.1",
    "sorting": "large single tissue: {n} cells uniform in a cube at 1.5 cells/unit^3, types A/B, "
    "BinnedAllPairs r_max=1, PiecewiseLinear(0.8,2,tau) tau_AB=0.3, Heun dt=0.05",
}


def bench_snapshots(args):
    """Simulation::write on the 4M-cell sorting tissue (48 MB of positions a snapshot): K Heun
    steps with no snapshot, with a snapshot every E steps through the asynchronous hand-off
    (gather kernel, pinned copy on a side stream, writer thread; timed until the last file is
    written), and with the synchronous pattern of the reference (download to the host, then
    write the datasets, simulation.hpp:607-632)."""
    import shutil
    import tempfile

    from tecs_amd import GpuSimulation, workloads
    from tecs_amd.io import Writer

    n = 4_000_000 if args.n == 200_000 else args.n
    w = workloads.sorting(n)
    K = args.steps if args.steps != 2000 else 500
    E = max(1, args.snapshot_every)
    tmp = tempfile.mkdtemp(prefix="tecs_snap_")
    sim = GpuSimulation(w.desc)
    sim.upload(w.positions, w.polarities, w.types)
    sim.take_step(w.dt, max(1, args.warmup))
    sim.synchronize()
    out = {}
    try:
        for mode in ("none", "async", "sync", "none"):
            print(f"bench: snapshots mode {mode}, {K} steps, every {E}", file=sys.stderr, flush=True)
            writer = Writer(f"{tmp}/{mode}") if mode != "none" else None
            n_snap = 0
            t0 = time.perf_counter()
            done = 0
            while done < K:
                k = min(E, K - done)
                sim.take_step(w.dt, k)
                done += k
                if mode == "async":
                    sim.write(writer, done * w.dt)
                    n_snap += 1
                elif mode == "sync":
                    pos, pol, _ = sim.download()
                    writer.write(done * w.dt, n_snap, pos, pol)
                    n_snap += 1
            sim.synchronize()
            if writer is not None:
                writer.flush()
                writer.close()
            t = time.perf_counter() - t0
            out.setdefault(mode, []).append({"seconds": t, "snapshots": n_snap})
            for f in os.listdir(tmp):
                os.remove(os.path.join(tmp, f))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    t_none = min(r["seconds"] for r in out["none"])
    t_async = out["async"][0]["seconds"]
    t_sync = out["sync"][0]["seconds"]
    snaps = out["async"][0]["snapshots"]
    mb = n * 12 / 1e6
    print(f"bench: snapshots none {t_none * 1e3 / K:.3f} ms/step, async {t_async * 1e3 / K:.3f}, "
          f"sync {t_sync * 1e3 / K:.3f}", file=sys.stderr, flush=True)
    res = {
        "metric": "ms per Heun step with a snapshot every E steps (asynchronous hand-off)",
        "value": t_async * 1e3 / K,
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": t_async * 1e3 / K,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {
            "workload": f"large single tissue: {n} cells, BinnedAllPairs + type table, Heun dt=0.05, "
            f"an HDF5 snapshot (positions {mb:.0f} MB) every {E} steps",
            "snapshots": snaps,
            "ms_per_step_no_snapshots": t_none * 1e3 / K,
            "ms_per_step_async": t_async * 1e3 / K,
            "ms_per_step_sync_download_write": t_sync * 1e3 / K,
            "ms_per_snapshot_async": (t_async - t_none) * 1e3 / snaps,
            "ms_per_snapshot_sync": (t_sync - t_none) * 1e3 / snaps,
            "parallelism": "single GPU",
        },
        "roofline": None,
    }
    print(json.dumps(res), flush=True)


def bench_single(args, w):
    from tecs_amd import GpuSimulation, abi

    sim = GpuSimulation(w.desc)
    if args.sweep_variant != "auto":
        sim.set_sweep_variant({v: k for k, v in abi.SWEEP_NAMES.items()}[args.sweep_variant])
    sim.upload(w.positions, w.polarities, w.types)
    extra = args.walls or args.links > 0
    if args.walls:
        # a floor just inside the lowest cells: z >= zmin + 0.5 (strength 1)
        zmin = float(w.positions[:, 2].min())
        sim.set_walls([[0.0, 0.0, 1.0, -(zmin + 0.5), 1.0]])
    if args.links > 0:
        import numpy as np

        # each of `links` random cells linked to a random cell within distance 1 of it (the grid
        # query of the first step's grid), as the protrusion examples link neighbours
        g = np.random.default_rng(7)
        sim.take_step(w.dt, 1)
        a_idx = g.choice(w.n, size=min(args.links, w.n), replace=False)
        x, _, _ = sim.download()
        b_idx = sim.random_neighbour(x[a_idx], 1.0, g.random(len(a_idx)).astype(np.float32))
        keep = b_idx >= 0
        sim.set_links(np.stack([a_idx[keep], b_idx[keep]], axis=1), 0.1)
    warm = max(args.warmup, MIN_WARMUP)
    print(f"bench: {w.name} n={w.n} warmup={warm} steps={args.steps}", file=sys.stderr, flush=True)
    sim.take_step(w.dt, warm)
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
    has_pol = bool(w.desc.fields & abi.FIELD_POLARITY)
    has_type = bool(w.desc.fields & abi.FIELD_TYPE)
    bpc = algorithmic_bytes_per_cell(has_pol, has_type)
    symmetric = w.desc.pair_finder in abi.SYMMETRIC_FINDERS
    if symmetric or extra:
        # Delta sweep into a zeroed buffer: read s + a, the delta written once (s); the
        # integrator update is its own kernel
        s_bytes = 12 + (8 if has_pol else 0)
        bpc = {"full": 2 * s_bytes + (1 if has_type else 0), "stale": 2 * s_bytes + (1 if has_type else 0)}
    elif w.desc.integrator == abi.EULER:
        # Euler sweep: read s + a, write x' (s) and the old velocity (12)
        s_bytes = 12 + (8 if has_pol else 0)
        bpc = {"full": 2 * s_bytes + (1 if has_type else 0) + 12, "stale": 0}
    per = {
        "full": (full_ms / max(full_n, 1), bpc["full"]),
        "stale": (stale_ms / max(stale_n, 1), bpc["stale"]),
    }
    dom = max(per, key=lambda k: per[k][0])
    avg_ms, b_cell = per[dom]
    achieved = (b_cell * w.n) / (avg_ms * 1e-3) / 1e9
    variant = abi.SWEEP_NAMES.get(sim.sweep_variant(), "?")
    kname = "k_sweep_binned" if variant == "thread" else f"k_sweep_team<{variant[4:]}>"
    fname = abi.FINDER_NAMES[w.desc.pair_finder]
    if symmetric or extra:
        kernel_names = {"full": f"{kname} {fname} mode Delta",
                        "stale": f"{kname} {fname} mode Delta (stale grid)"}
    elif w.desc.integrator == abi.EULER:
        kernel_names = {"full": f"{kname} {fname} mode Euler", "stale": "-"}
    else:
        kernel_names = {"full": f"{kname} {fname} mode Pred (Heun substep 1)",
                        "stale": f"{kname} {fname} mode Corr (Heun substep 2, stale grid)"}
    traffic = None if (symmetric or extra) else measured_traffic(w.name, variant, dom)
    valu = None if (symmetric or extra) else measured_valu(w.name, variant, dom)
    print(
        f"bench: sweep full {per['full'][0]*1e3:.1f} us, stale {per['stale'][0]*1e3:.1f} us, "
        f"rebuild {reb_ms/max(reb_n,1)*1e3:.1f} us per launch",
        file=sys.stderr,
        flush=True,
    )

    out = {
        "metric": "pair-interactions/sec per Heun step" if w.desc.integrator == abi.HEUN
        else "pair-interactions/sec per Euler step",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {
            "workload": WORKLOAD_TEXT[w.name].format(n=w.n).replace(
                fname.replace("Symmetric", ""), fname),
            "cells": w.n,
            "pair_interactions_per_step": pairs / max(args.steps, 1),
            "walls": 1 if args.walls else 0,
            "links": args.links,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_names[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if traffic is None else traffic["fetch"] + traffic["write"],
            "traffic_fetch": None if traffic is None else traffic["fetch"],
            "traffic_write": None if traffic is None else traffic["write"],
            "traffic_over_algorithmic": None if traffic is None
            else (traffic["fetch"] + traffic["write"]) / (b_cell * w.n),
            "avg_launch_us": avg_ms * 1e3,
            "algorithmic_bytes_per_cell": b_cell,
            "sweep_full_us": per["full"][0] * 1e3,
            "sweep_stale_us": per["stale"][0] * 1e3,
            "rebuild_us": reb_ms / max(reb_n, 1) * 1e3,
            "sweep_variant": variant,
            "autotune_ms": sim.autotune_ms(),
            # VALU issue: counted instructions at the issue peak, against the launch time
            # measured here
            "valu": None if valu is None else {
                "instructions_per_launch": valu,
                "peak_instructions_per_s": VALU_PEAK_WAVE_INSTR_PER_S,
                "issue_us": valu / VALU_PEAK_WAVE_INSTR_PER_S * 1e6,
                "frac": (valu / VALU_PEAK_WAVE_INSTR_PER_S * 1e6) / (avg_ms * 1e3),
            },
        },
    }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        cb = cpu_baseline(w, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


def bench_queries(args):
    """Grid queries (grid.hpp:413-564) on the 200k-cell epithelium's grid after two steps:
    1M query points at the cells plus N(0, 0.3) jitter, get_nearest_point_index with search
    radius 1 and get_random_point_index_in_neighbourhood with radius 1 (caller draws). The
    entry points take host buffers, so the rate includes the copies of the points in and the
    answers out (16-20 bytes per query over PCIe); the kernel-only time is in profiles/."""
    import numpy as np

    from tecs_amd import GpuSimulation, workloads

    w = workloads.epithelium(args.n)
    sim = GpuSimulation(w.desc)
    sim.upload(w.positions, w.polarities, w.types)
    sim.take_step(w.dt, 2)
    sim.synchronize()
    g = workloads.rng(90)
    Q = 1_000_000
    pts = (w.positions[g.integers(0, w.n, Q)] + g.normal(0.0, 0.3, (Q, 3))).astype(np.float32)
    u = g.random(Q).astype(np.float32)
    print(f"bench: queries on epithelium n={w.n}, {Q} points", file=sys.stderr, flush=True)
    reps = max(args.steps // 200, 3)
    sim.nearest(pts, 1)
    sim.random_neighbour(pts, 1.0, u)
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.nearest(pts, 1)
    t_near = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.random_neighbour(pts, 1.0, u)
    t_rand = (time.perf_counter() - t0) / reps
    out = {
        "metric": "grid queries/sec (get_nearest_point_index, search radius 1 bin)",
        "value": Q / t_near,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": 1e3 * t_near,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {
            "workload": f"grid queries: {Q} points (cells + N(0,0.3) jitter) on the grid of the "
            f"{w.n}-cell polar epithelium (206^3 bins of edge 1), host buffers in and out",
            "queries": Q,
            "random_neighbour_queries_per_s": Q / t_rand,
            "parallelism": "single GPU",
        },
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        import oracle

        threads = int(os.environ.get("TECS_CPU_THREADS", min(16, os.cpu_count() or 1)))
        oracle.set_threads(threads)
        o = oracle.OracleSim(w.desc)
        o.upload(w.positions, w.polarities, w.types)
        o.take_step(w.dt, 2)
        sample = 200_000
        t0 = time.perf_counter()
        o.nearest(pts[:sample], 1)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": sample / el,
            "unit": "queries/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{sample} of the same nearest queries on the same grid ({el:.1f} s, C++ oracle, "
            "serial loop)",
        }
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


def cpu_baseline_instances(ws, seconds_target: float, warmup: int = 10, min_rounds: int = 100):
    """The oracle stepping a set of instances, one Heun step of each per round (round robin):
    `warmup` untimed rounds, then timed rounds until seconds_target has passed and at least
    min_rounds ran; value = the median over rounds of (pair interactions of the round / its
    time)."""
    threads, host = _cpu_setup()
    import oracle

    sims = []
    for w in ws:
        o = oracle.OracleSim(w.desc)
        o.upload(w.positions, w.polarities, w.types)
        o.take_step(w.dt, warmup)
        sims.append((o, w))
    rates = []
    t0 = time.perf_counter()
    while True:
        before = sum(o.stats()["pair_interactions"] for o, _ in sims)
        t1 = time.perf_counter()
        for o, w in sims:
            o.take_step(w.dt, 1)
        dt_s = time.perf_counter() - t1
        rates.append((sum(o.stats()["pair_interactions"] for o, _ in sims) - before) / dt_s)
        el = time.perf_counter() - t0
        if (el >= seconds_target and len(rates) >= min_rounds) or len(rates) >= 100000 or el >= 4 * seconds_target:
            break
    return {
        "value": _median(rates),
        "unit": "pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median over {len(rates)} timed rounds (one Heun step of each of {len(sims)} "
        f"instances, {sum(w.n for _, w in sims)} cells) after {warmup} warm-up rounds ({el:.1f} s, "
        f"C++/OpenMP oracle, {threads} threads)",
        **host,
    }


def bench_spheroid(args):
    """NaiveDelaunay: `instances` spheroids of 100 cells (examples/delaunay_spheroid.cpp: a ball
    of radius 3, r_max 1, PiecewiseLinear(0.5) x 3, Heun dt 0.001), stepped as one batch, and
    the first 16 stepped one at a time for comparison. A launch of the timed sweep class is the
    counting pass (k_dl_count) plus the per-cell pass (k_dl_cells); the roofline is VALU issue,
    the bound of the counting pass (scalar FP32 predicate arithmetic over LDS)."""
    from tecs_amd import GpuBatch, GpuSimulation, abi, workloads

    B = args.instances
    ws = [workloads.spheroid(100, seed_stream=100 + k) for k in range(B)]
    print(f"bench: spheroid x{B} (NaiveDelaunay, 100 cells each)", file=sys.stderr, flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions)
    dts = [w.dt for w in ws]
    b.take_step(dts, args.warmup)
    b.synchronize()
    b.reset_stats()
    t0 = time.perf_counter()
    b.take_step(dts, args.steps)
    b.synchronize()
    el = time.perf_counter() - t0
    pairs = b.stats()["pair_interactions"]
    value = pairs / el
    print(f"bench: batched {args.steps} steps in {el:.3f} s, {pairs} pair interactions",
          file=sys.stderr, flush=True)
    k_steps = min(max(args.steps // 4, 10), 200)
    b.set_kernel_timing(True)
    b.take_step(dts, k_steps)
    b.synchronize()
    full_ms, full_n = b.kernel_timing(abi.KT_SWEEP_FULL)
    stale_ms, stale_n = b.kernel_timing(abi.KT_SWEEP_STALE)
    b.set_kernel_timing(False)
    avg_ms = (full_ms + stale_ms) / max(full_n + stale_n, 1)
    plain_n = min(16, B)
    plain_pairs = 0
    plain_time = 0.0
    for w in ws[:plain_n]:
        s = GpuSimulation(w.desc)
        s.upload(w.positions)
        s.take_step(w.dt, 2)
        s.synchronize()
        s.reset_stats()
        t1 = time.perf_counter()
        s.take_step(w.dt, args.plain_steps)
        s.synchronize()
        plain_time += time.perf_counter() - t1
        plain_pairs += s.stats()["pair_interactions"]
        s.close()
    plain = plain_pairs / plain_time
    valu = measured_valu("spheroid", "thread", "count")
    out = {
        "metric": "pair-interactions/sec per Heun step, NaiveDelaunay batch",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {
            "workload": f"Delaunay spheroids: {B} instances of 100 cells uniform in a ball of "
            "radius 3, NaiveDelaunay r_max=1, PiecewiseLinear(0.5)x3, Heun dt=0.001",
            "cells": 100 * B,
            "instances": B,
            "pair_interactions_per_step": pairs / max(args.steps, 1),
            "plain_value": plain,
            "batched_over_plain": value / plain,
            "plain_instances": plain_n,
            "plain_steps_per_instance": args.plain_steps,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "valu",
            "kernel": "k_dl_count + k_dl_cells (NaiveDelaunay evaluation)",
            "achieved": None if valu is None else valu / (avg_ms * 1e-3),
            "peak": VALU_PEAK_WAVE_INSTR_PER_S,
            "unit": "wave-instructions/s",
            "frac": None if valu is None else (valu / (avg_ms * 1e-3)) / VALU_PEAK_WAVE_INSTR_PER_S,
            "traffic": None,
            "valu_instructions_per_launch": valu,
            "avg_launch_us": avg_ms * 1e3,
        },
    }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        cb = cpu_baseline_instances(ws, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


# untimed steps that always cover the capture of the step graphs (one period is 2 steps when
# the grid is rebuilt every step; the first step builds the grid and runs the autotuner)
MIN_WARMUP = 5


def bench_sweep(args):
    """256 independent instances (sizes log-uniform in [1k, 100k], relaxation or sorting
    geometry, tau_AB swept over [0.1, 1]), Heun dt=0.05: batched (every kernel launched once per
    substep for all instances, hipGraph replay) vs plain (each instance stepped alone, one after
    another, each warmed past its own graph capture)."""
    from tecs_amd import GpuBatch, GpuSimulation, abi, workloads

    ws = workloads.sweep(256)
    total_cells = sum(w.n for w in ws)
    warm = max(args.warmup, MIN_WARMUP)
    print(f"bench: sweep of {len(ws)} instances, {total_cells} cells, warmup {warm}, steps {args.steps}",
          file=sys.stderr, flush=True)
    b = GpuBatch([w.desc for w in ws])
    if args.sweep_variant != "auto":
        b.set_sweep_variant({v: k for k, v in abi.SWEEP_NAMES.items()}[args.sweep_variant])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    dts = [w.dt for w in ws]
    b.take_step(dts, warm)
    b.synchronize()
    b.reset_stats()
    g0 = b.graph_launches()
    t0 = time.perf_counter()
    b.take_step(dts, args.steps)
    b.synchronize()
    el = time.perf_counter() - t0
    replays = b.graph_launches() - g0
    st = b.stats()
    pairs = st["pair_interactions"]
    batched = pairs / el
    print(f"bench: batched {args.steps} steps in {el:.4f} s ({replays} graph replays)", file=sys.stderr,
          flush=True)
    # kernel timing pass: HIP events around every sweep and rebuild on the batch's stream
    k_steps = min(max(args.steps, 20), 200)
    b.set_kernel_timing(True)
    b.take_step(dts, k_steps)
    b.synchronize()
    full_ms, full_n = b.kernel_timing(abi.KT_SWEEP_FULL)
    stale_ms, stale_n = b.kernel_timing(abi.KT_SWEEP_STALE)
    reb_ms, reb_n = b.kernel_timing(abi.KT_REBUILD)
    b.set_kernel_timing(False)
    bpc = algorithmic_bytes_per_cell(False, True)
    per = {"full": (full_ms / max(full_n, 1), bpc["full"]), "stale": (stale_ms / max(stale_n, 1), bpc["stale"])}
    dom = max(per, key=lambda k: per[k][0])
    avg_ms, b_cell = per[dom]
    achieved = (b_cell * total_cells) / (avg_ms * 1e-3) / 1e9
    variant = abi.SWEEP_NAMES.get(b.sweep_variant(), "?")
    kname = "k_sweep_binned" if variant == "thread" else f"k_sweep_team<{variant[4:]}>"
    kernel_names = {"full": f"{kname} BinnedAllPairs mode Pred (Heun substep 1), all instances",
                    "stale": f"{kname} BinnedAllPairs mode Corr (Heun substep 2, stale grid), all instances"}
    traffic = measured_traffic("sweep", variant, dom)
    valu = measured_valu("sweep", variant, dom)
    autotune = b.autotune_ms()
    member_class, class_variants = b.size_classes()
    b.close()
    print(f"bench: size classes {dict(sorted((c, member_class.count(c)) for c in set(member_class)))} "
          f"variants {class_variants}", file=sys.stderr, flush=True)
    print(f"bench: sweep full {per['full'][0]*1e3:.1f} us, stale {per['stale'][0]*1e3:.1f} us, "
          f"rebuild {reb_ms/max(reb_n,1)*1e3:.1f} us per launch", file=sys.stderr, flush=True)

    plain = None
    plain_time = 0.0
    if not args.no_plain:
        # plain: one simulation on the GPU at a time, the same instances and initial state
        plain_pairs = 0
        for w in ws:
            s = GpuSimulation(w.desc)
            s.upload(w.positions, w.polarities, w.types)
            s.take_step(w.dt, warm)
            s.synchronize()
            s.reset_stats()
            t1 = time.perf_counter()
            s.take_step(w.dt, args.plain_steps)
            s.synchronize()
            plain_time += time.perf_counter() - t1
            plain_pairs += s.stats()["pair_interactions"]
            s.close()
        plain = plain_pairs / plain_time
        print(f"bench: plain {args.plain_steps} steps per instance in {plain_time:.3f} s", file=sys.stderr,
              flush=True)
    out = {
        "metric": "pair-interactions/sec per Heun step, batched vs plain GPU; p50/p99 request latency",
        "value": batched,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": 1e3 * el / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {
            "workload": "batched parameter sweep: 256 instances in one launch per kernel per substep, "
            "cells log-uniform in [1k,100k], relaxation (ball, radius 0.5*cbrt(N)) or sorting (cube, "
            "1.5 cells/unit^3, A/B types) geometry, tau_AB linear in [0.1,1], BinnedAllPairs r_max=1, "
            "PiecewiseLinear(0.8,2,tau), Heun dt=0.05, hipGraph-captured steps",
            "cells": total_cells,
            "instances": len(ws),
            "pair_interactions_per_step": pairs / max(args.steps, 1),
            "graph_replays_timed": replays,
            "plain_value": plain,
            "batched_over_plain": None if plain is None else batched / plain,
            "plain_steps_per_instance": args.plain_steps,
            "plain_timed_s": plain_time,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_names[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if traffic is None else traffic["fetch"] + traffic["write"],
            "traffic_fetch": None if traffic is None else traffic["fetch"],
            "traffic_write": None if traffic is None else traffic["write"],
            "traffic_over_algorithmic": None if traffic is None
            else (traffic["fetch"] + traffic["write"]) / (b_cell * total_cells),
            "avg_launch_us": avg_ms * 1e3,
            "algorithmic_bytes_per_cell": b_cell,
            "cells_per_launch": total_cells,
            "sweep_full_us": per["full"][0] * 1e3,
            "sweep_stale_us": per["stale"][0] * 1e3,
            "rebuild_us": reb_ms / max(reb_n, 1) * 1e3,
            "timing_launches": {"full": full_n, "stale": stale_n, "rebuild": reb_n},
            "sweep_variant": variant,
            "autotune_ms": autotune,
            # the autotuner's variant per size class (cells in [4^c, 4^(c+1))), members per class
            "size_class_variants": {str(c): v for c, v in sorted(class_variants.items())},
            "size_class_members": {str(c): member_class.count(c) for c in sorted(set(member_class))},
            # VALU issue: counted instructions at the issue peak, against the launch time
            # measured here
            "valu": None if valu is None else {
                "instructions_per_launch": valu,
                "peak_instructions_per_s": VALU_PEAK_WAVE_INSTR_PER_S,
                "issue_us": valu / VALU_PEAK_WAVE_INSTR_PER_S * 1e6,
                "frac": (valu / VALU_PEAK_WAVE_INSTR_PER_S * 1e6) / (avg_ms * 1e3),
            },
        },
    }
    if not args.no_serve:
        print("bench: served requests ...", file=sys.stderr, flush=True)
        sv = serve_run(workloads.serve_requests(args.requests, args.rate), args.rate, with_plain=not args.no_plain)
        out["serve"] = sv
        out["latency_p50_s"] = sv["latency_p50_s"]
        out["latency_p99_s"] = sv["latency_p99_s"]
    if not args.no_cpu_baseline:
        print("bench: cpu baseline ...", file=sys.stderr, flush=True)
        sample = ws[0::16] + ws[1::16]  # every 16th instance of each geometry
        cb = cpu_baseline_instances(sample, args.cpu_seconds)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = batched / cb["value"]
    print(json.dumps(out), flush=True)


def _percentiles(xs):
    import numpy as np

    a = np.asarray(xs, np.float64)
    return float(np.percentile(a, 50)), float(np.percentile(a, 99))


def serve_run(reqs, rate: float, with_plain: bool = True, chunk: int = 32) -> dict:
    """Serve the requests through the continuous-batching server (arrivals replayed in real
    time, HBM budget 200 GB) and, with_plain, the same requests one at a time in arrival order
    (each alone on the GPU). Latency = completion time - arrival time."""
    from tecs_amd import GpuSimulation
    from tecs_amd.serving import Server

    srv = Server(200 * 10**9, max_chunk_steps=chunk)
    t0 = time.perf_counter()
    ids = []
    nxt = 0
    done = set()
    last_report = t0
    while len(done) < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            r = reqs[nxt]
            ids.append(srv.submit(r.workload.desc, r.workload.positions, r.workload.polarities,
                                  r.workload.types, r.workload.dt, r.n_steps))
            nxt += 1
        running, pending = srv.advance()
        if running == 0 and pending == 0 and nxt < len(reqs):
            time.sleep(max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0)))
        for k, rid in enumerate(ids):
            if k not in done and srv.info(rid)["state"] == 2:
                done.add(k)
        if time.perf_counter() - last_report > 10:
            last_report = time.perf_counter()
            print(f"bench: serve {len(done)}/{len(reqs)} done", file=sys.stderr, flush=True)
    makespan = time.perf_counter() - t0
    st = srv.stats()
    lat = [srv.info(rid)["t_done"] - srv.info(rid)["t_submit"] for rid in ids]
    srv.close()
    p50, p99 = _percentiles(lat)
    served = st["pair_interactions"] / makespan
    print(f"bench: served makespan {makespan:.2f} s p50 {p50:.2f} s p99 {p99:.2f} s", file=sys.stderr, flush=True)
    out = {
        "workload": f"served requests: {len(reqs)} Poisson arrivals at {rate}/s, model "
        "relaxation/epithelium/sorting uniform, cells log-uniform [2k,500k], steps uniform "
        f"[100,5000], HBM budget 200 GB, continuous batching (chunks of <= {chunk} steps)",
        "value": served,
        "unit": "pair-interactions/s",
        "requests": len(reqs),
        "cells_total": int(sum(r.workload.n for r in reqs)),
        "steps_total": int(sum(r.n_steps for r in reqs)),
        "makespan_s": makespan,
        "latency_p50_s": p50,
        "latency_p99_s": p99,
        "regroups": st["regroups"],
    }
    if with_plain:
        plain_pairs = 0
        t1 = time.perf_counter()
        lat_plain = []
        by_model = {}
        for r in reqs:
            now = time.perf_counter() - t1
            if r.arrival > now:
                time.sleep(r.arrival - now)
            t_req = time.perf_counter()
            s = GpuSimulation(r.workload.desc)
            s.upload(r.workload.positions, r.workload.polarities, r.workload.types)
            s.take_step(r.workload.dt, r.n_steps)
            s.synchronize()
            s.download()
            pairs_r = s.stats()["pair_interactions"]
            plain_pairs += pairs_r
            lat_plain.append(time.perf_counter() - t1 - r.arrival)
            s.close()
            m = by_model.setdefault(r.model, {"requests": 0, "seconds": 0.0, "pair_interactions": 0})
            m["requests"] += 1
            m["seconds"] += time.perf_counter() - t_req
            m["pair_interactions"] += pairs_r
        plain_span = time.perf_counter() - t1
        q50, q99 = _percentiles(lat_plain)
        out.update({"plain_value": plain_pairs / plain_span, "plain_makespan_s": plain_span,
                    "plain_latency_p50_s": q50, "plain_latency_p99_s": q99,
                    "served_over_plain": served / (plain_pairs / plain_span), "plain_by_model": by_model})
        print(f"bench: plain makespan {plain_span:.2f} s p50 {q50:.2f} s p99 {q99:.2f} s", file=sys.stderr,
              flush=True)
    return out


def bench_serve(args):
    """Served simulation requests: Poisson arrivals, models relaxation / epithelium / sorting
    drawn uniformly, cells log-uniform in [2k, 500k], steps uniform in [100, 5000], admission by
    an HBM budget of 200 GB, continuous batching; compared with serving the same requests one
    at a time on the GPU (each runs alone, in arrival order)."""
    from tecs_amd import workloads

    reqs = workloads.serve_requests(args.requests, args.rate)
    print(f"bench: serve {len(reqs)} requests at {args.rate}/s", file=sys.stderr, flush=True)
    sv = serve_run(reqs, args.rate, with_plain=not args.no_plain)
    out = {
        "metric": "pair-interactions/sec per Heun step, served; p50/p99 request latency",
        "value": sv["value"],
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": sv["steps_total"],
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox, seed 2807)",
        "config": {**{k: v for k, v in sv.items() if k not in ("value", "unit")}, "parallelism": "single GPU"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
