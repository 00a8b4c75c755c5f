"""NaiveDelaunay (include/pair_finders/naive_delaunay.hpp:157-293, 3-D) on the GPU vs the oracle,
through the C ABI.

Bars (stated per test):
  * pair lists (row i = partner j once per qualifying third cell k, j ascending): bit-exact,
    including a lattice full of cospherical quadruples (ties of the <= test, zero
    denominators), all pairs within the cutoff, the largest population (512) and tiny ones
  * forces: the per-item packs are added in (j, k) order on both sides, so PiecewiseLinear
    deltas are bitwise the oracle's; every field within 1e-5 of the FP64 evaluation of the
    same pair list (oracle.force_excess), or within the sequential FP32 summation bound
    m * 2^-24 where a cell sums m > 168 packs; the polarity terms use trig identities on the
    GPU, so those fields meet the FP64 bar only
  * trajectories without the COM fixer: bitwise (same pairs, same sums, same updates); with
    GlobalComFixer (a block reduction against a sequential sum) and polarities: median
    |dx| <= 1e-5, p99 <= 1e-3, max <= 0.05; old velocities (forces) relative to their RMS,
    with the 1e-3 bar at the 90th percentile (items on a sphere-test boundary flip)
  * batches and graph replay run the same kernels in the same order: bitwise
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W
from tecs_amd._lib import TecsError

pytestmark = pytest.mark.gpu


def _lattice(side, spacing):
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    return (g.astype(np.float32) * np.float32(spacing))


def _case(name):
    """(desc, positions, polarities, types) of a named configuration."""
    if name == "spheroid100":
        w = W.spheroid(100)
        return w.desc, w.positions, None, None
    if name == "spheroid160":
        w = W.spheroid(160, seed_stream=5)
        return w.desc, w.positions, None, None
    if name == "allpairs60":
        w = W.spheroid(60, seed_stream=6)
        d = abi.copy_desc(w.desc)
        d.cutoff_distance = 10.0
        return d, w.positions, None, None
    if name == "lattice64":
        pos = _lattice(4, 0.7)
        return abi.make_desc(len(pos), cutoff_distance=1.5, pair_finder=abi.NAIVE_DELAUNAY), pos, None, None
    if name == "max512":
        w = W.spheroid(512, seed_stream=7)
        d = abi.copy_desc(w.desc)
        d.cutoff_distance = 0.45
        return d, w.positions, None, None
    if name == "typed_polar120":
        w = W.spheroid(120, seed_stream=8)
        g = W.rng(9)
        pol = np.stack([g.uniform(0.2, 2.9, 120), g.uniform(-3.1, 3.1, 120)], 1).astype(np.float32)
        typ = (g.random(120) < 0.5).astype(np.uint8)
        d = abi.make_desc(120, pair_finder=abi.NAIVE_DELAUNAY, homeostatic_radius=0.5,
                          polarity=True, types=True, adhesion_table=[[1.0, 0.3], [0.3, 1.0]],
                          bending_strength=0.5, alignment_strength=0.5)
        return d, w.positions, pol, typ
    raise KeyError(name)


def _pair(oracle_lib, desc, pos, pol=None, typ=None):
    g = GpuSimulation(desc)
    o = oracle_lib.OracleSim(desc)
    for s in (g, o):
        s.upload(pos, pol, typ)
    return g, o


def _bitwise(a, b):
    return a is None and b is None or np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _assert_close_states(xa, xb, what, scale=1.0, q=0.99):
    d = np.abs(xa.astype(np.float64) - xb.astype(np.float64)).max(axis=1) / scale
    print(f"{what}: median {np.median(d):.2e} q{q} {np.quantile(d, q):.2e} max {d.max():.2e}")
    assert np.isfinite(xa).all()
    assert np.median(d) <= 1e-5 and np.quantile(d, q) <= 1e-3 and d.max() <= 0.05, what


CASES = ["spheroid100", "spheroid160", "allpairs60", "lattice64", "max512", "typed_polar120"]


@pytest.mark.parametrize("name", CASES)
def test_pairs_bit_exact(gpu, oracle_lib, name):
    d, pos, pol, typ = _case(name)
    print(f"case delaunay pairs {name} n={len(pos)}", flush=True)
    g, o = _pair(oracle_lib, d, pos, pol, typ)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert cg.size > 0
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    for i in range(len(pos)):
        assert np.all(np.diff(cg[rg[i]:rg[i + 1]]) >= 0)


@pytest.mark.parametrize("name", ["spheroid100", "allpairs60", "lattice64", "typed_polar120"])
def test_forces(gpu, oracle_lib, name):
    d, pos, pol, typ = _case(name)
    print(f"case delaunay forces {name} n={len(pos)}", flush=True)
    g, o = _pair(oracle_lib, d, pos, pol, typ)
    fg, pg = g.evaluate_force(True)
    fo, po = o.evaluate_force(True)
    rp, cols = o.pairs(True)
    f64, p64, s64, sp64 = o.forces_f64(rp, cols)
    # a cell with m items sums m packs in FP32 one after the other: error up to about
    # m * 2^-24 of the absolute sum, above 1e-5 once m > 168 (all pairs within the cutoff)
    rel = max(1e-5, float(np.diff(rp).max()) * 2.0 ** -24)
    print(f"items per cell <= {np.diff(rp).max()}, bar {rel:.1e}")
    assert oracle_lib.force_excess(fg, f64, s64, rel) <= 0.0
    if pol is None:
        assert _bitwise(fg, fo)
    else:
        assert oracle_lib.force_excess(pg, p64, sp64, rel) <= 0.0
    assert np.abs(fg).max() > 0.0
    assert g.stats()["pair_interactions"] == len(cols)


@pytest.mark.parametrize("integrator", [abi.HEUN, abi.EULER], ids=["Heun", "Euler"])
def test_trajectory_bitwise(gpu, oracle_lib, integrator):
    w = W.spheroid(100)
    d = abi.copy_desc(w.desc)
    d.integrator = integrator
    print(f"case delaunay trajectory spheroid100 integrator={integrator}", flush=True)
    g, o = _pair(oracle_lib, d, w.positions)
    for s in (g, o):
        s.take_step(0.01, 15)
    xg, _, vg = g.download()
    xo, _, vo = o.download()
    assert _bitwise(xg, xo) and _bitwise(vg, vo)
    assert np.abs(xg - w.positions).max() > 1e-4
    st_g, st_o = g.stats(), o.stats()
    assert st_g["steps"] == st_o["steps"] == 15
    assert st_g["evaluations"] == st_o["evaluations"]
    assert st_g["pair_interactions"] == st_o["pair_interactions"]


@pytest.mark.parametrize("name,com", [("spheroid100", abi.GLOBAL_COM_FIXER),
                                      ("typed_polar120", abi.NO_COM_FIXER)])
def test_trajectory_tolerance(gpu, oracle_lib, name, com):
    d, pos, pol, typ = _case(name)
    d = abi.copy_desc(d)
    d.com_fixer = com
    if com == abi.GLOBAL_COM_FIXER:
        d.capacity = d.agent_count + 40
    print(f"case delaunay trajectory {name} com={com}", flush=True)
    g, o = _pair(oracle_lib, d, pos, pol, typ)
    # the example's dt: at dt = 0.01 a one-ulp change of one coordinate already moves the
    # oracle's own cells by O(1) within 10 steps (pair multiplicities of ~40 make the forces
    # large and the pair set jumps), so no summation-order tolerance could hold there
    for s in (g, o):
        s.take_step(0.001, 10)
    xg, pg, vg = g.download()
    xo, po, vo = o.download()
    _assert_close_states(xg, xo, "positions")
    # old velocities are forces (tens here): compared relative to their RMS magnitude. A cell
    # sums ~100 items, and an item whose sphere test sits on the boundary flips under the
    # 1e-7 differences of the polarity terms, changing that cell's force by ~1%: up to 10% of
    # the cells may be beyond 1e-3
    vrms = float(np.sqrt((vo.astype(np.float64) ** 2).sum(axis=1).mean()))
    _assert_close_states(vg, vo, "old velocities / rms", max(vrms, 1.0), q=0.9)
    if pg is not None:
        _assert_close_states(pg, po, "polarities")


def test_batch_members_match_alone(gpu, oracle_lib):
    """Members of 40, 100 and 300 cells (rows of 256 and 512 counters) in one batch: pairs as
    the oracle's, trajectories bitwise as alone."""
    ws = [W.spheroid(n, seed_stream=10 + n) for n in (40, 100, 300)]
    print("case delaunay batch of 3", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions)
    b.take_step([0.01] * 3, 6)
    for k, w in enumerate(ws):
        s = GpuSimulation(w.desc)
        s.upload(w.positions)
        s.take_step(0.01, 6)
        xa, _, va = s.download()
        xb, _, vb = b.download(k)
        assert _bitwise(xa, xb) and _bitwise(va, vb), k
    w = ws[0]
    o = oracle_lib.OracleSim(w.desc)
    o.upload(w.positions)
    o.take_step(0.01, 6)
    xo, _, _ = o.download()
    xb, _, _ = b.download(0)
    assert _bitwise(xb, xo)


def test_graph_replay_bitwise(gpu):
    w = W.spheroid(100)
    print("case delaunay graph replay", flush=True)
    sims = []
    for mode in (abi.GRAPHS_ON, abi.GRAPHS_OFF):
        s = GpuSimulation(w.desc)
        s.set_graph_mode(mode)
        s.upload(w.positions)
        sims.append(s)
    graph, eager = sims
    for call, steps in enumerate([1, 5, 9, 4]):
        graph.take_step(0.01, steps)
        eager.take_step(0.01, steps)
        xg, _, _ = graph.download()
        xe, _, _ = eager.download()
        assert _bitwise(xg, xe), call
    assert graph.graph_launches() >= 2
    assert graph.stats()["pair_interactions"] == eager.stats()["pair_interactions"]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5])
def test_tiny_populations(gpu, oracle_lib, n):
    """No item has an l below 4 cells; 4 and 5 cells in general position have some."""
    pos = W.rng(11).uniform(-0.4, 0.4, (n, 3)).astype(np.float32)
    d = abi.make_desc(n, pair_finder=abi.NAIVE_DELAUNAY)
    print(f"case delaunay tiny n={n}", flush=True)
    g, o = _pair(oracle_lib, d, pos)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
    assert (len(cg) > 0) == (n >= 4)
    for s in (g, o):
        s.take_step(0.01, 3)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert _bitwise(xg, xo)


def test_population_change(gpu, oracle_lib):
    """Growth within the capacity with new cells uploaded, then shrinking."""
    w = W.spheroid(140, seed_stream=12)
    d = abi.copy_desc(w.desc)
    d.agent_count, d.capacity = 100, 200
    print("case delaunay population 100 -> 140 -> 120", flush=True)
    g, o = _pair(oracle_lib, d, w.positions[:100])
    for s in (g, o):
        s.take_step(0.01, 3)
        s.set_agent_count(140)
        s.upload_range(100, w.positions[100:])
        s.take_step(0.01, 3)
        s.set_agent_count(120)
        s.take_step(0.01, 3)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert _bitwise(xg, xo)
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)


def test_capacity_limit(gpu):
    d = abi.make_desc(100, capacity=513, pair_finder=abi.NAIVE_DELAUNAY)
    with pytest.raises(TecsError, match="512"):
        GpuSimulation(d)


def test_served_spheroids_equal_alone(gpu):
    """Continuous batching of NaiveDelaunay requests (spheroids of 40 to 200 cells arriving
    over several rounds): every result is bitwise the same simulation run alone."""
    from tecs_amd.serving import Server

    reqs = [(W.spheroid(n, seed_stream=60 + n), steps, arrive)
            for n, steps, arrive in ((100, 9, 0), (40, 14, 0), (200, 6, 1), (70, 11, 2))]
    print("case delaunay served", flush=True)
    srv = Server(1 << 34, max_chunk_steps=4)
    ids = {}
    rounds = 0
    while True:
        for k, (w, steps, arrive) in enumerate(reqs):
            if arrive == rounds and k not in ids:
                ids[k] = srv.submit(w.desc, w.positions, None, None, 0.001, steps)
        srv.advance()
        rounds += 1
        if len(ids) == len(reqs) and all(srv.info(i)["state"] == 2 for i in ids.values()):
            break
        assert rounds < 200
    for k, (w, steps, _) in enumerate(reqs):
        xs, _ = srv.result(ids[k])
        s = GpuSimulation(w.desc)
        s.upload(w.positions)
        s.take_step(0.001, steps)
        xa, _, _ = s.download()
        assert _bitwise(xs, xa), k
