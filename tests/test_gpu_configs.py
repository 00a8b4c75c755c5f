"""The benchmark configurations against the oracle, at their own sizes, through the C ABI.

  * the batched parameter sweep: 16 members of the sweep generator (relaxation and sorting
    geometry, tau_AB swept, cells up to 20,000) as one batch, 20 Heun steps, every member
    bitwise against its own oracle run (positions and old velocities);
  * served requests: 8 Poisson requests of the three models (cells 2,000-20,000, 20-100
    steps) through the continuous-batching server, each result against the oracle run of the
    same request: bitwise for relaxation and sorting, the trajectory bar for the epithelium
    (median |dx| <= 1e-5, p99 <= 1e-3, max <= 0.05, for positions and for the polarities'
    unit vectors);
  * full sizes: the 10,000-cell relaxation over all 1,000 Euler steps bitwise; the 200,000-cell
    epithelium's Gabriel pair set bit-exact and its forces within the 1e-5 bar of FP64, and 3
    Heun steps of its team sweep bitwise the thread sweep's; the
    4,000,000-cell sorting tissue's pair CSR bit-exact on a fresh and on a stale grid and one
    Heun step bitwise (int32 slot and bin arithmetic at 4M cells and 2.7M bins).
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _traj_ok(xg, xo, what):
    dx = np.abs(xg.astype(np.float64) - xo.astype(np.float64)).max(axis=1)
    print(f"{what}: |d| median {np.median(dx):.2e} p99 {np.percentile(dx, 99):.2e} max {dx.max():.2e}",
          flush=True)
    return float(np.median(dx)) <= 1e-5 and float(np.percentile(dx, 99)) <= 1e-3 and float(dx.max()) <= 0.05


def _nvec(p):
    t, f = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)
    return np.stack([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)], axis=1)


def test_sweep_batch_members_equal_the_oracle(gpu, oracle_lib):
    ws = W.sweep(count=16, n_max=20000)
    print(f"case sweep batch of {len(ws)} members, {sum(w.n for w in ws)} cells", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    b.take_step([w.dt for w in ws], 20)
    for k, w in enumerate(ws):
        o = oracle_lib.OracleSim(w.desc)
        o.upload(w.positions, w.polarities, w.types)
        o.take_step(w.dt, 20)
        xg, _, vg = b.download(k)
        xo, _, vo = o.download()
        assert np.array_equal(_bits(xg), _bits(xo)), f"member {k} ({w.n} cells): positions differ"
        assert np.array_equal(_bits(vg), _bits(vo)), f"member {k}: old velocities differ"


def test_served_requests_equal_the_oracle(gpu, oracle_lib):
    from tecs_amd.serving import Server

    reqs = W.serve_requests(8, rate=1000.0, n_max=20000, steps_min=20, steps_max=100)
    print("case served requests", [(r.model, r.workload.n, r.n_steps) for r in reqs], flush=True)
    srv = Server(200 * 10**9, max_chunk_steps=16)
    ids = [srv.submit(r.workload.desc, r.workload.positions, r.workload.polarities, r.workload.types,
                      r.workload.dt, r.n_steps) for r in reqs]
    rounds = 0
    while not all(srv.info(i)["state"] == 2 for i in ids):
        srv.advance()
        rounds += 1
        assert rounds < 10000
    for r, rid in zip(reqs, ids):
        w = r.workload
        xs, ps = srv.result(rid)
        o = oracle_lib.OracleSim(w.desc)
        o.upload(w.positions, w.polarities, w.types)
        o.take_step(w.dt, r.n_steps)
        xo, po, _ = o.download()
        what = f"{r.model} {w.n} cells {r.n_steps} steps"
        if r.model == "epithelium":
            # polarities compared as unit vectors n(theta, phi): phi alone is ill-conditioned
            # near the poles, where a tiny change of direction moves it by O(1)
            assert np.isfinite(xs).all() and _traj_ok(xs, xo, what + " positions"), what
            assert np.isfinite(ps).all() and _traj_ok(_nvec(ps), _nvec(po), what + " polarity vectors"), what
        else:
            assert np.array_equal(_bits(xs), _bits(xo)), what


def test_relaxation_10k_all_1000_steps_bitwise(gpu, oracle_lib):
    w = W.relaxation(10_000)
    print("case relaxation 10000 cells, 1000 Euler steps", flush=True)
    g = GpuSimulation(w.desc)
    o = oracle_lib.OracleSim(w.desc)
    for s in (g, o):
        s.upload(w.positions)
        s.take_step(w.dt, 1000)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert np.array_equal(_bits(xg), _bits(xo))
    assert np.array_equal(_bits(vg), _bits(vo))
    assert g.stats()["pair_interactions"] == o.stats()["pair_interactions"]


def test_epithelium_200k_pairs_and_forces(gpu, oracle_lib):
    w = W.epithelium(200_000)
    print("case epithelium 200000 cells: Gabriel pairs, forces", flush=True)
    g = GpuSimulation(w.desc)
    o = oracle_lib.OracleSim(w.desc)
    for s in (g, o):
        s.upload(w.positions, w.polarities)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    fg, pg = g.evaluate_force(False)
    f64, p64, s64, sp64 = o.forces_f64(ro, co)
    assert oracle_lib.force_excess(fg, f64, s64) <= 0.0
    assert oracle_lib.force_excess(pg, p64, sp64) <= 0.0


def test_epithelium_200k_team_steps_equal_the_thread_sweep(gpu):
    """The 200,000-cell epithelium (a sparse grid) over 3 Heun steps with the 8-lane team
    sweep, whose sums and integrator updates run one force component per lane, against the
    thread sweep: positions, polarities and old velocities bitwise (both add each pair's terms
    in the composition's order and read the same sines and cosines)."""
    w = W.epithelium(200_000)
    print("case epithelium 200000 cells: 3 Heun steps, team8 vs thread", flush=True)
    out = {}
    for v in (abi.SWEEP_THREAD, abi.SWEEP_TEAM8):
        g = GpuSimulation(w.desc)
        g.set_sweep_variant(v)
        g.upload(w.positions, w.polarities)
        g.take_step(w.dt, 3)
        out[v] = g.download()
        g.close()
    (xa, pa, va), (xb, pb, vb) = out[abi.SWEEP_THREAD], out[abi.SWEEP_TEAM8]
    assert np.isfinite(xb).all()
    assert np.array_equal(_bits(xa), _bits(xb)), "positions"
    assert np.array_equal(_bits(pa), _bits(pb)), "polarities"
    assert np.array_equal(_bits(va), _bits(vb)), "old velocities"


def test_sorting_4m_pairs_fresh_and_stale_and_one_step(gpu, oracle_lib):
    w = W.sorting(4_000_000)
    print("case sorting 4000000 cells: pairs, one Heun step, stale pairs", flush=True)
    g = GpuSimulation(w.desc)
    o = oracle_lib.OracleSim(w.desc)
    for s in (g, o):
        s.upload(w.positions, None, w.types)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert rg.shape == ro.shape and np.array_equal(rg, ro), "fresh row pointers"
    assert np.array_equal(cg, co), "fresh pair lists"
    del rg, cg, ro, co
    for s in (g, o):
        s.take_step(w.dt, 1)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert np.array_equal(_bits(xg), _bits(xo)) and np.array_equal(_bits(vg), _bits(vo)), "one Heun step"
    # the positions moved since the step's rebuild: a non-full evaluation sees the stale grid
    rg, cg = g.pairs(False)
    ro, co = o.pairs(False)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co), "stale pair CSR"
