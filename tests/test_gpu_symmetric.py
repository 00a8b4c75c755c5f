"""Symmetric finders (include/pair_finders/*_symmetric.hpp) on the GPU vs the oracle, through
the C ABI.

Bars (stated per test):
  * pair sets (row i = partners j < i, in the reference's visiting order): bit-exact, on fresh
    and on stale grids
  * forces: the reference accumulates with atomics (binned_all_pairs_symmetric.hpp:188-200) and
    so does the GPU, so the summation order is not fixed: every field within 1e-5 of the FP64
    evaluation of the same pair set (oracle.force_excess: per component, relative to the sum of
    absolute pair contributions plus the field's RMS force)
  * the total delta of every field vanishes up to rounding (each pair adds v and -v)
  * trajectories, graph replay and batching: median |dx| <= 1e-5, 99th percentile <= 1e-3,
    max <= 0.05 (a pair at the cutoff or the Gabriel boundary may flip under a different
    summation order and move two cells by up to ~dt*0.2)
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W

pytestmark = pytest.mark.gpu

SYM = [abi.BINNED_ALL_PAIRS_SYMMETRIC, abi.BINNED_GABRIEL_SYMMETRIC, abi.NAIVE_ALL_PAIRS_SYMMETRIC,
       abi.NAIVE_GABRIEL_SYMMETRIC]
BINNED_SYM = [abi.BINNED_ALL_PAIRS_SYMMETRIC, abi.BINNED_GABRIEL_SYMMETRIC]
NAME = lambda f: abi.FINDER_NAMES[f]  # noqa: E731
CASES = [("relaxation", 3000), ("epithelium", 4000), ("sorting", 20000)]


def _workload(name, n, finder):
    if finder in (abi.NAIVE_ALL_PAIRS_SYMMETRIC, abi.NAIVE_GABRIEL_SYMMETRIC):
        n = min(n, 1500)
    if name == "relaxation":
        return W.relaxation(n, finder=finder)
    if name == "epithelium":
        return W.epithelium(n, finder=finder)
    return W.sorting(n, finder=finder)


def _pair(oracle_lib, desc, w, positions=None):
    g = GpuSimulation(desc)
    o = oracle_lib.OracleSim(desc)
    pos = w.positions if positions is None else positions
    for s in (g, o):
        s.upload(pos, w.polarities, w.types)
    return g, o


def _assert_close_states(xa, xb, what):
    d = np.abs(xa.astype(np.float64) - xb.astype(np.float64)).max(axis=1)
    print(f"{what}: median {np.median(d):.2e} p99 {np.quantile(d, 0.99):.2e} max {d.max():.2e}")
    assert np.isfinite(xa).all()
    assert np.median(d) <= 1e-5 and np.quantile(d, 0.99) <= 1e-3 and d.max() <= 0.05, what


@pytest.mark.parametrize("finder", SYM, ids=NAME)
@pytest.mark.parametrize("name,n", CASES)
def test_pairs_bit_exact(gpu, oracle_lib, finder, name, n):
    w = _workload(name, n, finder)
    print(f"case sym pairs {name} n={w.n} finder={NAME(finder)}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert cg.size > 0
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    i = np.repeat(np.arange(w.n), np.diff(rg))
    assert np.all(cg < i)


@pytest.mark.parametrize("finder", BINNED_SYM, ids=NAME)
def test_pairs_stale_grid_bit_exact(gpu, oracle_lib, finder):
    w = _workload("sorting", 20000, finder)
    print(f"case sym stale grid finder={NAME(finder)}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w)
    for s in (g, o):
        s.rebuild()
    moved = (w.positions + W.rng(77).uniform(-0.3, 0.3, w.positions.shape)).astype(np.float32)
    for s in (g, o):
        s.upload(moved, w.polarities, w.types)
    rg, cg = g.pairs(False)
    ro, co = o.pairs(False)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)


@pytest.mark.parametrize("finder", SYM, ids=NAME)
@pytest.mark.parametrize("name,n", CASES)
def test_forces_within_1e5_of_fp64(gpu, oracle_lib, finder, name, n):
    w = _workload(name, n, finder)
    print(f"case sym forces {name} n={w.n} finder={NAME(finder)}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w)
    fg, pg = g.evaluate_force(True)
    rp, cols = o.pairs(True)  # the same positions, so the grid this evaluation built
    f64, p64, s64, sp64 = o.forces_f64(rp, cols)
    assert oracle_lib.force_excess(fg, f64, s64) <= 0.0
    fo, po = o.evaluate_force(True)
    assert oracle_lib.force_excess(fo, f64, s64) <= 0.0
    tot = lambda a: np.abs(a.astype(np.float64).sum(axis=0)).max()  # noqa: E731
    assert tot(fg) <= 1e-5 * np.abs(fg).sum()
    if pg is not None:
        assert oracle_lib.force_excess(pg, p64, sp64) <= 0.0
        assert tot(pg) <= 1e-5 * np.abs(pg).sum()
    assert g.stats()["pair_interactions"] == len(cols)


VARIANTS = [abi.SWEEP_THREAD, abi.SWEEP_TEAM4, abi.SWEEP_TEAM8, abi.SWEEP_TEAM16]
VARIANT_IDS = ["thread", "team4", "team8", "team16"]


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", BINNED_SYM, ids=NAME)
@pytest.mark.parametrize("name,n", [("epithelium", 4000), ("sorting", 20000)])
@pytest.mark.parametrize("stale", [False, True], ids=["fresh", "stale"])
def test_forces_every_variant(gpu, oracle_lib, finder, name, n, variant, stale):
    """Every sweep variant, on the grid built from the positions and on a stale grid whose
    cells moved by up to 0.3 (many change bin: the movers' scatter path)."""
    w = _workload(name, n, finder)
    print(f"case sym variant {name} n={w.n} finder={NAME(finder)} variant={variant} stale={stale}",
          flush=True)
    g = GpuSimulation(w.desc)
    g.set_sweep_variant(variant)
    o = oracle_lib.OracleSim(w.desc)
    for s in (g, o):
        s.upload(w.positions, w.polarities, w.types)
    if stale:
        moved = (w.positions + W.rng(78).uniform(-0.3, 0.3, w.positions.shape)).astype(np.float32)
        for s in (g, o):
            s.rebuild()
            s.upload(moved, w.polarities, w.types)
    fg, pg = g.evaluate_force(not stale)
    rp, cols = o.pairs(not stale)
    f64, p64, s64, sp64 = o.forces_f64(rp, cols)
    assert oracle_lib.force_excess(fg, f64, s64) <= 0.0
    if pg is not None:
        assert oracle_lib.force_excess(pg, p64, sp64) <= 0.0
    assert g.stats()["pair_interactions"] == len(cols)


@pytest.mark.parametrize("radius", [3.0, 1.5], ids=["110-neighbours", "900-neighbours"])
@pytest.mark.parametrize("stale", [False, True], ids=["fresh", "stale"])
def test_dense_cluster_thread_sweep(gpu, oracle_lib, radius, stale):
    """Dense clusters on the all-pairs symmetric thread sweep: ~110 and ~900 neighbours per
    cell, far more accepted candidates than the deferred list holds (the rest are evaluated in
    a second walk), on the grid built from the positions and on a stale grid whose cells moved
    by up to 0.3 (most of them movers)."""
    n = 3000
    pos = W.ball(n, radius, W.rng(22))
    lo, hi = W.bounds(pos)
    d = abi.make_desc(n, pair_finder=abi.BINNED_ALL_PAIRS_SYMMETRIC, min_bounds=lo, max_bounds=hi,
                      integrator=abi.HEUN)
    print(f"case sym dense cluster radius={radius} stale={stale}", flush=True)
    g = GpuSimulation(d)
    g.set_sweep_variant(abi.SWEEP_THREAD)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(pos, None, None)
    if stale:
        moved = (pos + W.rng(23).uniform(-0.3, 0.3, pos.shape)).astype(np.float32)
        for s in (g, o):
            s.rebuild()
            s.upload(moved, None, None)
    fg, _ = g.evaluate_force(not stale)
    rp, cols = o.pairs(not stale)
    f64, _, s64, _ = o.forces_f64(rp, cols)
    assert oracle_lib.force_excess(fg, f64, s64) <= 0.0
    assert g.stats()["pair_interactions"] == len(cols)
    assert abs(fg.astype(np.float64).sum(axis=0)).max() <= 1e-5 * np.abs(fg).sum()


@pytest.mark.parametrize("stale", [False, True], ids=["fresh", "stale"])
def test_search_radius_two(gpu, oracle_lib, stale):
    """Bins of half the cutoff (search radius 2): the thread sweep walks every cell serially,
    the movers of a stale grid included (the wavefront walk of the movers takes radius 1 only
    and leaves these instances alone)."""
    w = W.sorting(20000, finder=abi.BINNED_ALL_PAIRS_SYMMETRIC)
    d = abi.copy_desc(w.desc)
    d.pair_finder_settings.bin_size_scale = 0.5
    print(f"case sym search radius 2 stale={stale}", flush=True)
    g = GpuSimulation(d)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(w.positions, w.polarities, w.types)
    if stale:
        moved = (w.positions + W.rng(79).uniform(-0.3, 0.3, w.positions.shape)).astype(np.float32)
        for s in (g, o):
            s.rebuild()
            s.upload(moved, w.polarities, w.types)
    fg, _ = g.evaluate_force(not stale)
    rp, cols = o.pairs(not stale)
    f64, _, s64, _ = o.forces_f64(rp, cols)
    assert oracle_lib.force_excess(fg, f64, s64) <= 0.0
    assert g.stats()["pair_interactions"] == len(cols)


def test_half_the_functor_calls(gpu):
    """Fresh grid, all-pairs: the symmetric finder calls the functor once per unordered pair,
    the ordered one twice."""
    w = _workload("sorting", 20000, abi.BINNED_ALL_PAIRS)
    print("case sym pair count", flush=True)
    counts = []
    for f in (abi.BINNED_ALL_PAIRS, abi.BINNED_ALL_PAIRS_SYMMETRIC):
        d = abi.copy_desc(w.desc)
        d.pair_finder = f
        s = GpuSimulation(d)
        s.upload(w.positions, w.polarities, w.types)
        s.evaluate_force(True)
        counts.append(s.stats()["pair_interactions"])
    assert counts[0] > 0 and counts[0] == 2 * counts[1]


@pytest.mark.parametrize("finder,name,integrator,com", [
    (abi.BINNED_ALL_PAIRS_SYMMETRIC, "sorting", abi.HEUN, abi.NO_COM_FIXER),
    (abi.BINNED_ALL_PAIRS_SYMMETRIC, "relaxation", abi.EULER, abi.NO_COM_FIXER),
    (abi.BINNED_ALL_PAIRS_SYMMETRIC, "relaxation", abi.EULER, abi.GLOBAL_COM_FIXER),
    (abi.BINNED_ALL_PAIRS_SYMMETRIC, "relaxation", abi.HEUN, abi.GLOBAL_COM_FIXER),
    (abi.BINNED_GABRIEL_SYMMETRIC, "epithelium", abi.HEUN, abi.NO_COM_FIXER),
    (abi.NAIVE_ALL_PAIRS_SYMMETRIC, "relaxation", abi.HEUN, abi.NO_COM_FIXER),
    (abi.NAIVE_GABRIEL_SYMMETRIC, "epithelium", abi.EULER, abi.NO_COM_FIXER),
], ids=lambda v: str(v))
def test_trajectory_tolerance(gpu, oracle_lib, finder, name, integrator, com):
    w = _workload(name, {"sorting": 20000, "relaxation": 3000, "epithelium": 4000}[name], finder)
    d = abi.copy_desc(w.desc)
    d.integrator = integrator
    d.com_fixer = com
    if com == abi.GLOBAL_COM_FIXER:
        d.capacity = d.agent_count + 500
    print(f"case sym trajectory {name} n={w.n} finder={NAME(finder)} integrator={integrator} com={com}",
          flush=True)
    g, o = _pair(oracle_lib, d, w)
    for s in (g, o):
        s.take_step(w.dt, 20)
    xg, pg, vg = g.download()
    xo, po, vo = o.download()
    _assert_close_states(xg, xo, "positions")
    _assert_close_states(vg, vo, "old velocities")
    if pg is not None:
        _assert_close_states(pg, po, "polarities")
    assert g.stats()["steps"] == o.stats()["steps"] == 20
    assert g.stats()["evaluations"] == o.stats()["evaluations"]


@pytest.mark.parametrize("name", ["sorting", "epithelium"])
def test_fused_predictor_thread_sweep(gpu, oracle_lib, name):
    """Heun without the COM fix on the binned all-pairs symmetric thread sweep: the predictor
    runs in the fresh sweep's epilogue (positions, polarities and the flags of the cells it
    takes out of their rebuild bin), the stale sweep then routes the movers' pairs by those
    flags. 20 steps against the oracle at the trajectory bar of this file."""
    w = _workload(name, {"sorting": 20000, "epithelium": 4000}[name], abi.BINNED_ALL_PAIRS_SYMMETRIC)
    d = abi.copy_desc(w.desc)
    d.integrator = abi.HEUN
    print(f"case sym fused predictor {name} n={w.n}", flush=True)
    g, o = _pair(oracle_lib, d, w)
    g.set_sweep_variant(abi.SWEEP_THREAD)
    for s in (g, o):
        s.take_step(w.dt, 20)
    xg, pg, vg = g.download()
    xo, po, vo = o.download()
    _assert_close_states(xg, xo, "positions")
    _assert_close_states(vg, vo, "old velocities")
    if pg is not None:
        _assert_close_states(pg, po, "polarities")
    assert g.stats()["steps"] == o.stats()["steps"] == 20


def test_graph_replay_matches_eager(gpu):
    """Steps replayed from captured graphs end where kernel-by-kernel launches end, to the
    summation-order tolerance; the state that zeroes the deltas before the first step is
    captured apart from the periodic one."""
    w = _workload("sorting", 20000, abi.BINNED_ALL_PAIRS_SYMMETRIC)
    print("case sym graph replay", flush=True)
    sims = []
    for mode in (abi.GRAPHS_ON, abi.GRAPHS_OFF):
        s = GpuSimulation(w.desc)
        s.set_graph_mode(mode)
        s.upload(w.positions, w.polarities, w.types)
        sims.append(s)
    graph, eager = sims
    for call, steps in enumerate([1, 8, 13, 8, 2, 16]):
        graph.take_step(w.dt, steps)
        eager.take_step(w.dt, steps)
        xg, _, _ = graph.download()
        xe, _, _ = eager.download()
        _assert_close_states(xg, xe, f"call {call}")
    assert graph.graph_launches() >= 4
    assert graph.stats()["steps"] == eager.stats()["steps"]


def test_evaluate_between_steps(gpu, oracle_lib):
    """evaluate_force leaves its result in the delta buffer the steps accumulate into; the
    next step starts from zero again."""
    w = _workload("relaxation", 3000, abi.BINNED_ALL_PAIRS_SYMMETRIC)
    print("case sym evaluate between steps", flush=True)
    g, o = _pair(oracle_lib, w.desc, w)
    for s in (g, o):
        s.take_step(w.dt, 3)
        s.evaluate_force(True)
        s.take_step(w.dt, 3)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    _assert_close_states(xg, xo, "positions")


def test_batch_member_matches_alone(gpu):
    ws = [W.sorting(n, finder=abi.BINNED_ALL_PAIRS_SYMMETRIC) for n in (3000, 12000, 500)]
    print("case sym batch of 3", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    b.take_step([w.dt for w in ws], 20)
    for k, w in enumerate(ws):
        s = GpuSimulation(w.desc)
        s.upload(w.positions, w.polarities, w.types)
        s.take_step(w.dt, 20)
        xa, _, _ = s.download()
        xb, _, _ = b.download(k)
        _assert_close_states(xa, xb, f"member {k}")


def test_long_run_ensemble_statistics(gpu, oracle_lib):
    """500 Heun steps of the polarity model with BinnedGabrielSymmetric: mean radius and its
    spread, polarity alignment with the outward normal and the mean number of pairs per cell
    agree within 1% (relative) or 0.01 (absolute)."""
    w = _workload("epithelium", 4000, abi.BINNED_GABRIEL_SYMMETRIC)
    print("case sym long run epithelium n=4000 steps=500", flush=True)
    g, o = _pair(oracle_lib, w.desc, w)
    for s in (g, o):
        s.take_step(w.dt, 500)

    def stats(x, p, rp):
        r = np.linalg.norm(x.astype(np.float64), axis=1)
        u = x / r[:, None]
        t, f = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)
        nrm = np.stack([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)], axis=1)
        return {"mean_radius": r.mean(), "radius_std": r.std(),
                "polarity_along_normal": (nrm * u).sum(axis=1).mean(), "degree": rp[-1] / len(x)}

    xg, pg, _ = g.download()
    xo, po, _ = o.download()
    sg = stats(xg, pg, g.pairs(True)[0])
    so = stats(xo, po, o.pairs(True)[0])
    print("gpu", sg, "oracle", so)
    for k in sg:
        assert abs(sg[k] - so[k]) <= max(0.01 * abs(so[k]), 0.01), k


@pytest.mark.parametrize("finder", SYM, ids=NAME)
@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_populations(gpu, oracle_lib, finder, n):
    """Empty, single-cell and two coincident cells (d = 0: PiecewiseLinear returns zero)."""
    w = W.relaxation(max(n, 1), finder=finder)
    pos = np.zeros((n, 3), np.float32) + np.float32(0.25)
    d = abi.copy_desc(w.desc)
    d.agent_count = d.capacity = n
    print(f"case sym tiny n={n} finder={NAME(finder)}", flush=True)
    g = GpuSimulation(d)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(pos, None, None)
        s.take_step(w.dt, 3)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    assert len(cg) == (1 if n == 2 else 0)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))


@pytest.mark.parametrize("finder", BINNED_SYM, ids=NAME)
def test_cells_outside_the_grid_bounds(gpu, oracle_lib, finder):
    """A 3000-cell ball centred at (16, 16, -17) under the default bounds +-20 (about 1300 cells
    collapse into the edge bins, grid.hpp:222-230): pairs on the fresh grid bit-exact, forces
    within 1e-5 of FP64, 3 Euler steps within the trajectory bar, then the pairs on the grid of
    the last rebuild bit-exact."""
    import oracle

    w = W.relaxation(3000, finder=finder)
    pos = (w.positions + np.float32([16.0, 16.0, -17.0])).astype(np.float32)
    d = abi.copy_desc(w.desc)
    d.pair_finder_settings.min_bounds = (-20.0, -20.0, -20.0)
    d.pair_finder_settings.max_bounds = (20.0, 20.0, 20.0)
    print(f"case sym outside bounds finder={NAME(finder)}", flush=True)
    g, o = _pair(oracle_lib, d, w, positions=pos)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert cg.size > 0 and np.array_equal(rg, ro) and np.array_equal(cg, co)
    fg, _ = g.evaluate_force(True)
    f64, _, s64, _ = o.forces_f64(ro, co)
    assert oracle.force_excess(fg, f64, s64, 1e-5) <= 0.0
    for s in (g, o):
        s.take_step(w.dt, 3)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    _assert_close_states(xg, xo, "outside bounds, 3 steps")
    # the stale grid is the same grid when the positions are the same; compare on the oracle's
    for s in (g, o):
        s.upload(xo, None, None)
        s.rebuild()
    moved = (xo + W.rng(78).uniform(-0.3, 0.3, xo.shape)).astype(np.float32)
    for s in (g, o):
        s.upload(moved, None, None)
    rg, cg = g.pairs(False)
    ro, co = o.pairs(False)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
