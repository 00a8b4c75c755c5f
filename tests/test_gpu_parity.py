"""GPU engine vs the CPU oracle on identical inputs, through the C ABI.

Bars (stated per test):
  * pair sets: bit-exact - same CSR rows, same order (the reference's visiting order)
  * PiecewiseLinear / type-table forces: bit-exact vs the FP32 oracle (same op sequence, same
    summation order), and within 1e-5 of the FP64 evaluation
  * polarity forces (sin/cos/acos/atan2 differ by ulps between GPU and host libm): within
    1e-5 of the FP64 evaluation, per component, relative to the cell's sum of absolute pair
    contributions plus the field's RMS force (oracle.force_excess)
  * trajectories: bit-exact where the forces are; tolerance stated for polarity models
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W

pytestmark = pytest.mark.gpu

FINDERS = [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL, abi.NAIVE_ALL_PAIRS, abi.NAIVE_GABRIEL]


def _with(desc, **kw):
    d = abi.copy_desc(desc)
    for k, v in kw.items():
        if k in ("rebuild_every_n", "bin_size_scale", "gabriel_coefficient"):
            setattr(d.pair_finder_settings, k, v)
        else:
            setattr(d, k, v)
    return d


VARIANTS = [abi.SWEEP_THREAD, abi.SWEEP_TEAM4, abi.SWEEP_TEAM8, abi.SWEEP_TEAM16]
VARIANT_IDS = ["thread", "team4", "team8", "team16"]
TEAM_VARIANTS = [abi.SWEEP_TEAM4, abi.SWEEP_TEAM8, abi.SWEEP_TEAM16]


def _pair(oracle_lib, desc, w, positions=None, variant=abi.SWEEP_AUTO):
    g = GpuSimulation(desc)
    g.set_sweep_variant(variant)
    o = oracle_lib.OracleSim(desc)
    pos = w.positions if positions is None else positions
    for s in (g, o):
        s.upload(pos, w.polarities, w.types)
    return g, o


def _workload(name, n, finder):
    if name == "relaxation":
        return W.relaxation(n, finder=finder)
    if name == "epithelium":
        return W.epithelium(n, finder=finder)
    return W.sorting(n, finder=finder)


def _assert_same_csr(a, b, what):
    ra, ca = a
    rb, cb = b
    assert ra.shape == rb.shape, what
    assert np.array_equal(ra, rb), f"{what}: row pointers differ"
    assert np.array_equal(ca, cb), f"{what}: pair lists differ"


def _force_scale_check(f32, f64, scale, rel=1e-5):
    import oracle

    return oracle.force_excess(f32, f64, scale, rel)


CASES = [
    ("relaxation", 3000),
    ("epithelium", 4000),
    ("sorting", 20000),
]


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", FINDERS, ids=lambda f: abi.FINDER_NAMES[f])
@pytest.mark.parametrize("name,n", CASES)
def test_pairs_full_step_bit_exact(gpu, oracle_lib, finder, name, n, variant):
    if finder in (abi.NAIVE_ALL_PAIRS, abi.NAIVE_GABRIEL):
        if variant in TEAM_VARIANTS:
            pytest.skip("naive finders have one sweep kernel")
        n = min(n, 1500)
    w = _workload(name, n, finder)
    print(f"case {name} n={n} finder={abi.FINDER_NAMES[finder]} variant={variant}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w, variant=variant)
    pg = g.pairs(True)
    po = o.pairs(True)
    assert pg[1].size > 0
    _assert_same_csr(pg, po, f"{name}/{abi.FINDER_NAMES[finder]}")


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
def test_pairs_stale_grid_bit_exact(gpu, oracle_lib, finder, variant):
    """Heun corrector semantics: the grid of the last rebuild with moved positions."""
    w = _workload("sorting", 20000, finder)
    print(f"case stale grid finder={abi.FINDER_NAMES[finder]} variant={variant}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w, variant=variant)
    for s in (g, o):
        s.rebuild()
    moved = (w.positions + W.rng(77).uniform(-0.3, 0.3, w.positions.shape)).astype(np.float32)
    for s in (g, o):
        s.upload(moved, w.polarities, w.types)
    _assert_same_csr(g.pairs(False), o.pairs(False), "stale grid")


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", FINDERS, ids=lambda f: abi.FINDER_NAMES[f])
@pytest.mark.parametrize("name,n", [("relaxation", 3000), ("sorting", 20000)])
def test_forces_bit_exact_piecewise_linear(gpu, oracle_lib, finder, name, n, variant):
    if finder in (abi.NAIVE_ALL_PAIRS, abi.NAIVE_GABRIEL):
        if variant in TEAM_VARIANTS:
            pytest.skip("naive finders have one sweep kernel")
        n = min(n, 1500)
    w = _workload(name, n, finder)
    print(f"case forces {name} n={n} finder={abi.FINDER_NAMES[finder]} variant={variant}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w, variant=variant)
    fg, _ = g.evaluate_force(True)
    fo, _ = o.evaluate_force(True)
    assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32)), "forces differ bitwise"
    rp, cols = o.pairs(False)
    f64, _, s64, _ = o.forces_f64(rp, cols)
    assert _force_scale_check(fg, f64, s64) <= 0.0


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", [abi.BINNED_GABRIEL, abi.NAIVE_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
def test_forces_polarity_within_1e5_of_fp64(gpu, oracle_lib, finder, variant):
    if finder == abi.NAIVE_GABRIEL and variant in TEAM_VARIANTS:
        pytest.skip("naive finders have one sweep kernel")
    n = 4000 if finder == abi.BINNED_GABRIEL else 1500
    w = _workload("epithelium", n, finder)
    print(f"case polarity forces n={n} finder={abi.FINDER_NAMES[finder]} variant={variant}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w, variant=variant)
    fg, pg = g.evaluate_force(True)
    rp, cols = o.pairs(True)  # same positions, so the same freshly built grid
    f64, p64, s64, sp64 = o.forces_f64(rp, cols)
    e1 = _force_scale_check(fg, f64, s64)
    e2 = _force_scale_check(pg, p64, sp64)
    print(f"excess over 1e-5 bound: pos {e1:.3e} pol {e2:.3e}")
    assert e1 <= 0.0 and e2 <= 0.0
    # how often the bar's F_rms floor carries it (the cell's own S_i alone covers the rest)
    import oracle

    for fld, a, b, sc in (("pos", fg, f64, s64), ("pol", pg, p64, sp64)):
        u = oracle.force_bar_usage(a, b, sc)
        print(f"bar usage {fld}: {u}")
        assert u["fraction"] <= 1e-3


def _run_both(oracle_lib, w, desc, steps, variant=abi.SWEEP_AUTO):
    g, o = _pair(oracle_lib, desc, w, variant=variant)
    g.take_step(w.dt, steps)
    o.take_step(w.dt, steps)
    return g, o


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize(
    "name,n,steps",
    [("relaxation", 3000, 100), ("sorting", 20000, 100)],
)
def test_trajectory_bit_exact(gpu, oracle_lib, name, n, steps, variant):
    w = _workload(name, n, abi.BINNED_ALL_PAIRS)
    print(f"case trajectory {name} n={n} steps={steps} variant={variant}", flush=True)
    g, o = _run_both(oracle_lib, w, w.desc, steps, variant)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), (
        f"positions differ after {steps} steps, max {np.abs(xg - xo).max()}"
    )
    assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32)), "old velocities differ"
    assert g.stats()["pair_interactions"] == o.stats()["pair_interactions"]


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
def test_trajectory_gabriel_relaxation_bit_exact(gpu, oracle_lib, variant):
    w = _workload("relaxation", 3000, abi.BINNED_GABRIEL)
    d = _with(w.desc, integrator=abi.HEUN)
    print(f"case trajectory relaxation/BinnedGabriel Heun 50 steps variant={variant}", flush=True)
    g, o = _run_both(oracle_lib, w, d, 50, variant)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))


def test_trajectory_epithelium_tolerance(gpu, oracle_lib):
    """100 Heun steps of the polarity model. Per-step differences are libm ulps; a Gabriel
    pair whose test sits within an ulp of the boundary can flip and move two cells by up to
    ~dt*0.2. Bar: median |dx| <= 1e-5, 99th percentile <= 1e-3, max <= 0.05; polarities alike."""
    w = _workload("epithelium", 4000, abi.BINNED_GABRIEL)
    print("case trajectory epithelium n=4000 steps=100", flush=True)
    g, o = _run_both(oracle_lib, w, w.desc, 100)
    xg, pg, _ = g.download()
    xo, po, _ = o.download()
    dx = np.abs(xg - xo).max(axis=1)
    dp = np.abs(pg - po).max(axis=1)
    print(f"dx median {np.median(dx):.2e} p99 {np.quantile(dx, 0.99):.2e} max {dx.max():.2e}")
    print(f"dpol median {np.median(dp):.2e} p99 {np.quantile(dp, 0.99):.2e} max {dp.max():.2e}")
    assert np.isfinite(xg).all() and np.isfinite(pg).all()
    assert np.median(dx) <= 1e-5 and np.quantile(dx, 0.99) <= 1e-3 and dx.max() <= 0.05
    assert np.median(dp) <= 1e-5 and np.quantile(dp, 0.99) <= 1e-3 and dp.max() <= 0.05


@pytest.mark.parametrize("integrator", [abi.EULER, abi.HEUN], ids=["Euler", "Heun"])
def test_global_com_fixer(gpu, oracle_lib, integrator):
    """GlobalComFixer divides by the capacity (com_fixers.hpp:40-42). Its reduction order is
    backend-defined, so positions agree to a tolerance (1e-5 absolute after 20 steps)."""
    w = _workload("relaxation", 3000, abi.BINNED_ALL_PAIRS)
    d = _with(w.desc, integrator=integrator, com_fixer=abi.GLOBAL_COM_FIXER, capacity=3500)
    print(f"case com fixer integrator={integrator}", flush=True)
    g, o = _run_both(oracle_lib, w, d, 20)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert np.abs(xg - xo).max() <= 1e-5
    assert np.abs(vg - vo).max() <= 1e-5


def test_rebuild_every_n_bit_exact(gpu, oracle_lib):
    w = _workload("sorting", 20000, abi.BINNED_ALL_PAIRS)
    d = _with(w.desc, rebuild_every_n=3)
    print("case rebuild_every_n=3", flush=True)
    g, o = _run_both(oracle_lib, w, d, 10)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))
    assert g.stats()["rebuilds"] == o.stats()["rebuilds"] == 4


@pytest.mark.parametrize("integrator", [abi.EULER, abi.HEUN], ids=["Euler", "Heun"])
def test_old_velocities_across_rebuilds_outside_steps(gpu, oracle_lib, integrator):
    """A step's own rebuild leaves the old velocities to the step's final write; rebuilds
    outside a step (a full evaluation, a pair export) between steps move them with their
    cells. Positions and old velocities stay bitwise the oracle's through both."""
    w = _workload("sorting", 20000, abi.BINNED_ALL_PAIRS)
    d = _with(w.desc, integrator=integrator)
    print(f"case old velocities across rebuilds integrator={integrator}", flush=True)
    g, o = _pair(oracle_lib, d, w)

    def same(what):
        xg, _, vg = g.download()
        xo, _, vo = o.download()
        assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), f"positions differ {what}"
        assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32)), f"old velocities differ {what}"

    for s in (g, o):
        s.take_step(w.dt, 4)
    same("after 4 steps")
    for s in (g, o):
        s.evaluate_force(True)
    same("after an evaluation that rebuilt the grid")
    for s in (g, o):
        s.pairs(True)
    same("after a pair export that rebuilt the grid")
    for s in (g, o):
        s.take_step(w.dt, 3)
    same("after 3 more steps")


def test_batch_member_equals_alone(gpu):
    """Every member of a batch ends bitwise where it ends when run alone."""
    ws = W.sweep(count=6, n_min=500, n_max=20000)
    print("case batch of 6", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    b.take_step([w.dt for w in ws], 20)
    for k, w in enumerate(ws):
        s = GpuSimulation(w.desc)
        s.upload(w.positions, w.polarities, w.types)
        s.take_step(w.dt, 20)
        xa, _, va = s.download()
        xb, _, vb = b.download(k)
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), f"member {k} differs"
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))


@pytest.mark.parametrize("variant", TEAM_VARIANTS, ids=VARIANT_IDS[1:])
@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
def test_dense_cluster_team_overflow_path(gpu, oracle_lib, finder, variant):
    """~110 neighbours per cell: the team sweep's LDS list overflows and the leader walks the
    cell serially; pairs and forces stay bit-exact."""
    g0 = W.rng(21)
    n = 3000
    pos = W.ball(n, 3.0, g0)
    lo, hi = W.bounds(pos)
    d = abi.make_desc(n, pair_finder=finder, min_bounds=lo, max_bounds=hi, integrator=abi.HEUN)
    w = W.Workload("dense", d, pos, None, None, 0.05, 1)
    print(f"case dense cluster finder={abi.FINDER_NAMES[finder]}", flush=True)
    g, o = _pair(oracle_lib, d, w, variant=variant)
    assert g.sweep_variant() == variant
    _assert_same_csr(g.pairs(True), o.pairs(True), "dense")
    fg, _ = g.evaluate_force(True)
    fo, _ = o.evaluate_force(True)
    assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32))


@pytest.mark.parametrize("shape", ["slab", "dense"])
def test_thread_sweep_wide_and_dense_grids(gpu, oracle_lib, shape):
    """The all-pairs thread sweep keeps its row ranges in LDS with 8-bit lengths; a workgroup
    with a range of 256 slots or more ("dense": ~900 neighbours per cell, far more than the
    deferred list holds) takes the row-by-row walk. "slab": a thin sheet of ~100k cells per
    layer of bins, whose neighbour rows lie up to ~100k slots apart. Pairs and forces stay
    bit-exact, and so do a few Heun steps."""
    g0 = W.rng(24)
    if shape == "slab":
        n = 200_000
        pos = np.stack([g0.uniform(-200.0, 200.0, n), g0.uniform(-200.0, 200.0, n),
                        g0.uniform(-0.6, 0.6, n)], axis=1).astype(np.float32)
    else:
        n = 3000
        pos = W.ball(n, 1.5, g0)
    types = (g0.random(n) < 0.5).astype(np.uint8)
    lo, hi = W.bounds(pos)
    d = abi.make_desc(n, pair_finder=abi.BINNED_ALL_PAIRS, min_bounds=lo, max_bounds=hi,
                      integrator=abi.HEUN, types=True, adhesion_table=[[1.0, 0.3], [0.3, 1.0]])
    w = W.Workload(shape, d, pos, None, types, 0.01, 1)
    print(f"case thread sweep row-walk fallback shape={shape} n={n}", flush=True)
    g, o = _pair(oracle_lib, d, w, variant=abi.SWEEP_THREAD)
    assert g.sweep_variant() == abi.SWEEP_THREAD
    _assert_same_csr(g.pairs(True), o.pairs(True), shape)
    fg, _ = g.evaluate_force(True)
    fo, _ = o.evaluate_force(True)
    assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32))
    for sim in (g, o):
        sim.take_step(w.dt, 3)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))
    assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32))


def test_team_variant_requires_search_radius_one(gpu):
    w = _workload("sorting", 5000, abi.BINNED_ALL_PAIRS)
    d = _with(w.desc, bin_size_scale=0.5)
    g = GpuSimulation(d)
    g.set_sweep_variant(abi.SWEEP_TEAM)
    assert g.sweep_variant() == abi.SWEEP_THREAD


@pytest.mark.parametrize("scale", [0.5, 2.0, 1.3])
@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
def test_bin_size_scale_bit_exact(gpu, oracle_lib, finder, scale):
    """Bins of half / twice / 1.3 times the cutoff: search radius 2 (thread sweep) and 1 (team
    sweep); 1.3 is not a power of two, so the Gabriel bin coordinates take the division path."""
    w = _workload("relaxation", 3000, finder)
    d = _with(w.desc, bin_size_scale=scale)
    print(f"case bin_size_scale={scale} finder={abi.FINDER_NAMES[finder]}", flush=True)
    g, o = _pair(oracle_lib, d, w)
    _assert_same_csr(g.pairs(True), o.pairs(True), f"scale {scale}")
    fg, _ = g.evaluate_force(True)
    fo, _ = o.evaluate_force(True)
    assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32))


@pytest.mark.parametrize("name,rebuild_every_n", [("relaxation", 0), ("sorting", 0), ("sorting", 3), ("epithelium", 0)])
def test_graph_replay_matches_eager(gpu, name, rebuild_every_n):
    """Steps replayed from a captured hipGraph end bitwise where kernel-by-kernel launches
    end; checked after every call, with eager work and synchronisations between replays."""
    w = _workload(name, 6000, W.epithelium(10).desc.pair_finder if name == "epithelium" else abi.BINNED_ALL_PAIRS)
    d = _with(w.desc, rebuild_every_n=rebuild_every_n)
    print(f"case graph replay {name} rebuild_every_n={rebuild_every_n}", flush=True)
    sims = []
    for mode in (abi.GRAPHS_ON, abi.GRAPHS_OFF):
        s = GpuSimulation(d)
        s.set_graph_mode(mode)
        s.upload(w.positions, w.polarities, w.types)
        sims.append(s)
    graph, eager = sims
    other = GpuSimulation(W.relaxation(2000).desc)
    other.upload(W.relaxation(2000).positions)
    for call, steps in enumerate([1, 8, 13, 8, 2, 16]):
        graph.take_step(w.dt, steps)
        eager.take_step(w.dt, steps)
        other.evaluate_force(True)  # unrelated eager work between replays
        xg, pg, vg = graph.download()
        xe, pe, ve = eager.download()
        assert np.array_equal(xg.view(np.uint32), xe.view(np.uint32)), f"call {call}"
        assert np.array_equal(vg.view(np.uint32), ve.view(np.uint32)), f"call {call}"
        if pg is not None:
            assert np.array_equal(pg.view(np.uint32), pe.view(np.uint32)), f"call {call}"
    assert graph.graph_launches() >= 4
    assert eager.graph_launches() == 0
    assert graph.stats() == eager.stats()



def test_autotuner_picks_an_exact_variant(gpu, oracle_lib):
    """AUTO times every applicable sweep on the live state at the first take_step; whatever it
    picks, the trajectory equals the oracle's bitwise and the statistics are untouched."""
    w = _workload("sorting", 20000, abi.BINNED_ALL_PAIRS)
    print("case autotune sorting n=20000", flush=True)
    g, o = _run_both(oracle_lib, w, w.desc, 12)
    assert g.sweep_variant() in (abi.SWEEP_THREAD, abi.SWEEP_TEAM4, abi.SWEEP_TEAM8, abi.SWEEP_TEAM16)
    times = g.autotune_ms()
    print("autotune ms", times)
    assert set(times) == {"thread", "team4", "team8", "team16"}
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))
    assert g.stats() == o.stats()


def test_batch_mixed_sparse_and_dense_grids(gpu, oracle_lib):
    """Shell instances of different sizes have different bins per cell, so the counting sort
    uses per-bin counts for some and multi-bin segments for others in one batch; each member
    still matches the oracle's Gabriel pairs after steps and ends where it ends alone."""
    ws = [W.epithelium(n, seed_stream=40 + k) for k, n in enumerate((500, 20000, 3000))]
    print("case batch of shells 500/20000/3000", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    b.take_step([w.dt for w in ws], 6)
    for k, w in enumerate(ws):
        s = GpuSimulation(w.desc)
        s.upload(w.positions, w.polarities, w.types)
        s.take_step(w.dt, 6)
        xa, pa, _ = s.download()
        xb, pb, _ = b.download(k)
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), f"member {k}"
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), f"member {k}"
        o = oracle_lib.OracleSim(w.desc)
        o.upload(xa, pa)
        s.upload(xa, pa)
        _assert_same_csr(s.pairs(True), o.pairs(True), f"member {k} pairs after steps")


def test_snapshots_async_hand_off(gpu, tmp_path):
    """Simulation::write while stepping: every group t<k> holds the state at the k-th write,
    bitwise, although the copies and the file writes overlap the following steps."""
    from tecs_amd import io

    w = _workload("epithelium", 20000, abi.BINNED_GABRIEL)
    s = GpuSimulation(w.desc)
    s.upload(w.positions, w.polarities, w.types)
    ref = GpuSimulation(w.desc)
    ref.upload(w.positions, w.polarities, w.types)
    print("case snapshots epithelium n=20000", flush=True)
    expected = []
    with io.Writer(str(tmp_path / "snap" / "epi")) as writer:
        for k in range(6):
            s.write(writer, k * w.dt * 5)
            x, p, _ = ref.download()
            expected.append((x, p))
            s.take_step(w.dt, 5)
            ref.take_step(w.dt, 5)
    h5 = str(tmp_path / "snap" / "epi.h5")
    for k, (x, p) in enumerate(expected):
        assert np.array_equal(io.read_dataset(h5, f"t{k}/position"), x), k
        assert np.array_equal(io.read_dataset(h5, f"t{k}/polarity"), p), k


def test_server_results_equal_alone(gpu):
    """Continuous batching: requests of three models arrive over several rounds, a tight HBM
    budget makes some wait, members join and leave batches; every result is bitwise the
    result of the same simulation run alone."""
    from tecs_amd.serving import Server

    reqs = [
        (W.relaxation(3000, seed_stream=50), 37, 0),
        (W.epithelium(4000, seed_stream=51), 23, 0),
        (W.sorting(9000, seed_stream=52), 41, 1),
        (W.relaxation(1500, seed_stream=53), 12, 2),
        (W.epithelium(2500, seed_stream=54), 30, 3),
        (W.sorting(5000, seed_stream=55), 9, 3),
    ]
    # a budget that holds only about half of the requests at once
    sizes = []
    for w, _, _ in reqs:
        srv_probe = Server(1 << 40)
        rid = srv_probe.submit(w.desc, w.positions, w.polarities, w.types, w.dt, 1)
        sizes.append(srv_probe.info(rid)["hbm_bytes"])
        srv_probe.close()
    budget = sum(sorted(sizes)[-3:])
    srv = Server(budget, max_chunk_steps=4)
    ids = {}
    rounds = 0
    while True:
        for k, (w, steps, arrive) in enumerate(reqs):
            if arrive == rounds and k not in ids:
                ids[k] = srv.submit(w.desc, w.positions, w.polarities, w.types, w.dt, steps)
        running, pending = srv.advance()
        rounds += 1
        if len(ids) == len(reqs) and all(srv.info(i)["state"] == 2 for i in ids.values()):
            break
        assert rounds < 500
    st = srv.stats()
    print("server stats", st, "rounds", rounds, flush=True)
    assert st["completed"] == len(reqs) and st["regroups"] >= 4
    assert st["hbm_bytes_in_use"] == 0
    for k, (w, steps, _) in enumerate(reqs):
        info = srv.info(ids[k])
        assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
        xs, ps = srv.result(ids[k])
        s = GpuSimulation(w.desc)
        s.upload(w.positions, w.polarities, w.types)
        s.take_step(w.dt, steps)
        xa, pa, _ = s.download()
        assert np.array_equal(xs.view(np.uint32), xa.view(np.uint32)), f"request {k}"
        if pa is not None:
            assert np.array_equal(ps.view(np.uint32), pa.view(np.uint32)), f"request {k}"


def _shell_stats(x, p):
    r = np.linalg.norm(x.astype(np.float64), axis=1)
    u = x / r[:, None]
    t, f = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)
    n = np.stack([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)], axis=1)
    return {
        "mean_radius": r.mean(),
        "radius_std": r.std(),
        "polarity_along_normal": (n * u).sum(axis=1).mean(),
    }


def test_long_run_ensemble_statistics(gpu, oracle_lib):
    """1,000 Heun steps of the polarity model (dt 0.05): GPU and oracle trajectories part at
    the ulp level (libm) and then through flipped Gabriel pairs, so they are compared as
    ensembles: mean radius and its spread, mean polarity alignment with the outward normal, and
    mean Gabriel degree agree within 1% (relative) or 0.01 (absolute)."""
    w = _workload("epithelium", 4000, abi.BINNED_GABRIEL)
    print("case long run epithelium n=4000 steps=1000", flush=True)
    g, o = _run_both(oracle_lib, w, w.desc, 1000)
    xg, pg, _ = g.download()
    xo, po, _ = o.download()
    sg, so = _shell_stats(xg, pg), _shell_stats(xo, po)
    rg, _ = g.pairs(True)
    ro, _ = o.pairs(True)
    sg["degree"] = (rg[-1] / w.n)
    so["degree"] = (ro[-1] / w.n)
    print("gpu", sg, "oracle", so)
    s0 = _shell_stats(w.positions, w.polarities)
    for k in sg:
        assert abs(sg[k] - so[k]) <= max(0.01 * abs(so[k]), 0.01), k
    # the run went somewhere: the shell relaxed away from its initial spread
    assert abs(sg["radius_std"] - s0["radius_std"]) > 1e-3


def test_set_force_between_steps(gpu, oracle_lib):
    """take_step(dt, force) with a new force: 10 steps with PiecewiseLinear(0.8, 2, 1), then 10
    with (0.7, 1.5, 0.5); bitwise the oracle run as two simulations with the two forces."""
    w = _workload("relaxation", 3000, abi.BINNED_ALL_PAIRS)
    print("case set_force relaxation n=3000", flush=True)
    d2 = abi.copy_desc(w.desc)
    d2.force.homeostatic_radius = 0.7
    d2.force.repulsion_scale = 1.5
    d2.force.adhesion_scale = 0.5
    g = GpuSimulation(w.desc)
    g.upload(w.positions, w.polarities, w.types)
    g.take_step(w.dt, 10)
    g.set_force(d2.force)
    g.take_step(w.dt, 10)
    o1 = oracle_lib.OracleSim(w.desc)
    o1.upload(w.positions, w.polarities, w.types)
    o1.take_step(w.dt, 10)
    x1, _, v1 = o1.download()
    o2 = oracle_lib.OracleSim(d2)
    o2.upload(x1, None, None)
    o2.take_step(w.dt, 10)
    xg, _, _ = g.download()
    xo, _, _ = o2.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))
    bad = abi.copy_desc(w.desc).force
    bad.bending_strength = 1.0  # needs the polarity field
    with pytest.raises(Exception):
        g.set_force(bad)


def test_upload_mid_run_restarts_the_grid(gpu, oracle_lib):
    """Steps rebuild incrementally from the last order; an upload of another state in the middle
    of a run makes the next rebuild start from scratch. 30 steps, upload, 30 steps: bitwise the
    oracle doing the same."""
    w = _workload("sorting", 20000, abi.BINNED_ALL_PAIRS)
    w2 = W.sorting(20000, seed_stream=77)
    print("case upload mid-run sorting n=20000", flush=True)
    g, o = _pair(oracle_lib, w.desc, w)
    for s in (g, o):
        s.take_step(w.dt, 30)
        s.upload(w2.positions, w2.polarities, w2.types)
        s.take_step(w.dt, 30)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))
    assert g.stats()["pair_interactions"] == o.stats()["pair_interactions"]


@pytest.mark.parametrize("finder", FINDERS, ids=lambda f: abi.FINDER_NAMES[f])
@pytest.mark.parametrize("n", [0, 1, 2])
def test_tiny_populations(gpu, oracle_lib, finder, n):
    """Empty, single-cell and two-cell simulations (the two cells coincide: d = 0, where
    PiecewiseLinear returns zero, predefined.hpp:125-127): no faults, the oracle's pairs and
    trajectory bitwise."""
    w = W.relaxation(max(n, 1), finder=finder)
    pos = np.zeros((n, 3), np.float32) + np.float32(0.25)
    d = abi.copy_desc(w.desc)
    d.agent_count = d.capacity = n
    print(f"case tiny n={n} finder={abi.FINDER_NAMES[finder]}", flush=True)
    g = GpuSimulation(d)
    o = oracle_lib.OracleSim(d)
    for s in (g, o):
        s.upload(pos, None, None)
        s.take_step(w.dt, 3)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    assert len(cg) == (2 if n == 2 else 0)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))


def _outside_cloud(finder):
    """A 3000-cell ball centred at (16, 16, -17) under the default bounds +-20: about 1300 cells
    lie outside the grid and collapse into its edge bins (grid.hpp:222-230 clamps the bin
    coordinates), whose rows then hold ~20 cells a bin."""
    w = W.relaxation(3000, finder=finder)
    pos = (w.positions + np.float32([16.0, 16.0, -17.0])).astype(np.float32)
    d = abi.copy_desc(w.desc)
    d.pair_finder_settings.min_bounds = (-20.0, -20.0, -20.0)
    d.pair_finder_settings.max_bounds = (20.0, 20.0, 20.0)
    return w, d, pos


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
def test_cells_outside_the_grid_bounds(gpu, oracle_lib, finder, variant):
    """Cells beyond the bounds: pairs on the fresh grid, forces, a 3-step Euler trajectory and
    the pairs on the grid of its last rebuild, all bitwise the oracle's."""
    w, d, pos = _outside_cloud(finder)
    print(f"case outside bounds finder={abi.FINDER_NAMES[finder]} variant={variant}", flush=True)
    g, o = _pair(oracle_lib, d, w, positions=pos, variant=variant)
    pg = g.pairs(True)
    assert pg[1].size > 0
    _assert_same_csr(pg, o.pairs(True), "outside bounds, fresh grid")
    fg, _ = g.evaluate_force(True)
    fo, _ = o.evaluate_force(True)
    assert np.array_equal(fg.view(np.uint32), fo.view(np.uint32)), "forces differ bitwise"
    for s in (g, o):
        s.take_step(w.dt, 3)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), "trajectory differs"
    _assert_same_csr(g.pairs(False), o.pairs(False), "outside bounds, stale grid")


@pytest.mark.parametrize("name", ["sorting", "epithelium"])
def test_snapshots_equal_the_oracle_state(gpu, oracle_lib, tmp_path, name):
    """Simulation::write through the asynchronous hand-off against the oracle's state at the same
    step (simulation.hpp:607-611: t<k> holds the state at the k-th write, agent order, active
    cells): bitwise for the type-table tissue, the trajectory bar for the polarity model:
    median |d| <= 1e-5, p99 <= 1e-3, at most 0.2% of cells beyond 1e-3, and a maximum of 0.05
    for positions and 0.25 for the polarities' unit vectors (a Gabriel test within an ulp of its
    threshold can flip between GPU and host libm and turn a cell's polarity by up to
    ~2 dt per step while it lasts)."""
    from tecs_amd import io

    w = _workload(name, 20000, abi.BINNED_GABRIEL if name == "epithelium" else abi.BINNED_ALL_PAIRS)
    s = GpuSimulation(w.desc)
    o = oracle_lib.OracleSim(w.desc)
    for sim in (s, o):
        sim.upload(w.positions, w.polarities, w.types)
    print(f"case snapshots vs oracle {name} n=20000", flush=True)
    expected = []
    with io.Writer(str(tmp_path / "snap" / name)) as writer:
        for k in range(4):
            s.write(writer, k * w.dt * 7)
            expected.append(o.download()[:2])
            s.take_step(w.dt, 7)
            o.take_step(w.dt, 7)
    h5 = str(tmp_path / "snap" / f"{name}.h5")
    for k, (xo, po) in enumerate(expected):
        xs = io.read_dataset(h5, f"t{k}/position")
        assert xs.shape == (w.n, 3)
        if name == "sorting":
            assert np.array_equal(xs.view(np.uint32), xo.view(np.uint32)), k
            continue
        ps = io.read_dataset(h5, f"t{k}/polarity")
        for what, a, b, bound in (("positions", xs, xo, 0.05), ("polarity vectors", _unit(ps), _unit(po), 0.25)):
            d = np.abs(a.astype(np.float64) - b.astype(np.float64)).max(axis=1)
            stats = (f"t{k} {what}: median {np.median(d):.2e} p99 {np.quantile(d, 0.99):.2e} "
                     f"max {d.max():.2e} beyond 1e-3: {(d > 1e-3).sum()}")
            print(stats, flush=True)
            assert np.median(d) <= 1e-5 and np.quantile(d, 0.99) <= 1e-3, stats
            assert (d > 1e-3).mean() <= 0.002 and d.max() <= bound, stats


def _unit(p):
    t, f = p[:, 0].astype(np.float64), p[:, 1].astype(np.float64)
    return np.stack([np.sin(t) * np.cos(f), np.sin(t) * np.sin(f), np.cos(t)], axis=1)


def _sparse_row_case(kind, finder):
    """Sparse grids (8 or more bins per cell) count per row of bins and keep row-relative bin
    offsets. "long": a slab 1200 bins long in x, so a row holds more entries than the rebuild's
    row histogram (1024) and takes the counting path; "narrow": one bin in x, so every row is a
    single bin (ext[0] = 1)."""
    g = W.rng(7)
    if kind == "long":
        n = 2000
        pos = np.stack([g.uniform(-600.0, 600.0, n), g.uniform(0.0, 2.0, n), g.uniform(0.0, 2.0, n)], 1)
        lo, hi = (-602.0, -2.0, -2.0), (602.0, 4.0, 4.0)
    else:
        n = 1500
        pos = np.stack([g.uniform(0.2, 0.8, n), g.uniform(0.0, 40.0, n), g.uniform(0.0, 40.0, n)], 1)
        lo, hi = (0.0, -60.0, -60.0), (1.0, 100.0, 100.0)
    pos = pos.astype(np.float32)
    d = abi.make_desc(n, pair_finder=finder, min_bounds=lo, max_bounds=hi, integrator=abi.HEUN)
    return W.Workload("sparse-" + kind, d, pos, None, None, 0.05, 30)


@pytest.mark.parametrize("variant", VARIANTS, ids=VARIANT_IDS)
@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
@pytest.mark.parametrize("kind", ["long", "narrow"])
def test_sparse_grid_rows(gpu, oracle_lib, kind, finder, variant):
    """Pairs on the fresh grid, 30 Heun steps (incremental rebuilds rewriting the rows whose
    members moved) and the pairs on the last rebuild's grid, bitwise the oracle's."""
    w = _sparse_row_case(kind, finder)
    e = [int(np.ceil((w.desc.pair_finder_settings.max_bounds[k] - w.desc.pair_finder_settings.min_bounds[k])))
         for k in range(3)]
    assert e[0] * e[1] * e[2] >= 8 * w.n, "the case must be a sparse grid"
    print(f"case sparse rows {kind} finder={abi.FINDER_NAMES[finder]} variant={variant} ext={e}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w, variant=variant)
    pg = g.pairs(True)
    assert pg[1].size > 0
    _assert_same_csr(pg, o.pairs(True), f"sparse rows {kind}, fresh grid")
    for s in (g, o):
        s.take_step(w.dt, w.steps)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), "trajectory differs"
    _assert_same_csr(g.pairs(False), o.pairs(False), f"sparse rows {kind}, stale grid")
