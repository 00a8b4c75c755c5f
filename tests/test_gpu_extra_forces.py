"""Generic wall forces, link forces and cell division on the GPU vs the oracle, through the C ABI
(integrators/base.hpp:116-230 with the functors of examples/growth_w_wall.cpp,
growth_w_squeeze.cpp and intercalation.cpp).

Bars: the walls join the ordered finders' deltas in the reference's order (generic before
pairwise), so their trajectories stay bitwise the oracle's (Euler and Heun, naive and binned,
batched, graph-replayed, across a capacity reallocation); division places daughters bitwise
(directions formed on the host from the same draws); each cell adds its link contributions in
link order on both sides (the reference's atomics have no fixed order), so links are bitwise
too, shared endpoints included; the symmetric finders and GlobalComFixer with polarities:
median |dx| <= 1e-5, p99 <= 1e-3, max <= 0.05.
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W

pytestmark = pytest.mark.gpu

WALLS = [[0, 0, 1, 0.0, 10.0], [0, 1, 0, 1.5, 0.5], [1, 1, 0, 2.0, 2.0]]


def _pair(oracle_lib, desc, pos, pol=None, typ=None):
    g = GpuSimulation(desc)
    o = oracle_lib.OracleSim(desc)
    for s in (g, o):
        s.upload(pos, pol, typ)
    return g, o


def _same(g, o, what):
    xg, pg, vg = g.download()
    xo, po, vo = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), what
    assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32)), what
    if pg is not None:
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32)), what


def _close(g, o, what):
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    d = np.abs(xg.astype(np.float64) - xo).max(axis=1)
    print(f"{what}: median {np.median(d):.2e} p99 {np.quantile(d, 0.99):.2e} max {d.max():.2e}")
    assert np.median(d) <= 1e-5 and np.quantile(d, 0.99) <= 1e-3 and d.max() <= 0.05, what


@pytest.mark.parametrize("finder", [abi.BINNED_GABRIEL, abi.BINNED_ALL_PAIRS, abi.NAIVE_ALL_PAIRS],
                         ids=lambda f: abi.FINDER_NAMES[f])
@pytest.mark.parametrize("integrator", [abi.EULER, abi.HEUN], ids=["Euler", "Heun"])
def test_walls_bitwise(gpu, oracle_lib, finder, integrator):
    w = W.relaxation(3000 if finder != abi.NAIVE_ALL_PAIRS else 1200, finder=finder)
    d = abi.copy_desc(w.desc)
    d.integrator = integrator
    print(f"case walls finder={abi.FINDER_NAMES[finder]} integrator={integrator}", flush=True)
    g, o = _pair(oracle_lib, d, w.positions)
    for s in (g, o):
        s.set_walls(WALLS)
        s.take_step(0.05, 12)
    _same(g, o, "walls")
    x, _, _ = g.download()
    assert np.abs(x - w.positions).max() > 0.01


def test_walls_with_com_fixer_and_polarity(gpu, oracle_lib):
    w = W.epithelium(4000)
    d = abi.copy_desc(w.desc)
    d.com_fixer = abi.GLOBAL_COM_FIXER
    print("case walls epithelium com", flush=True)
    g, o = _pair(oracle_lib, d, w.positions, w.polarities)
    for s in (g, o):
        s.set_walls([[0, 0, 1, 5.0, 1.0]])
        s.take_step(w.dt, 10)
    _close(g, o, "walls + com + polarity")


@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS_SYMMETRIC, abi.NAIVE_DELAUNAY],
                         ids=lambda f: abi.FINDER_NAMES[f])
def test_walls_with_delta_only_finders(gpu, oracle_lib, finder):
    w = W.spheroid(100) if finder == abi.NAIVE_DELAUNAY else W.relaxation(3000, finder=finder)
    print(f"case walls finder={abi.FINDER_NAMES[finder]}", flush=True)
    g, o = _pair(oracle_lib, w.desc, w.positions)
    for s in (g, o):
        s.set_walls([[0, 0, 1, 0.5, 3.0]])
        s.take_step(0.001 if finder == abi.NAIVE_DELAUNAY else 0.05, 10)
    if finder == abi.NAIVE_DELAUNAY:
        _same(g, o, "walls + delaunay")
    else:
        _close(g, o, "walls + symmetric")


def test_links_one_per_cell_bitwise(gpu, oracle_lib):
    w = W.relaxation(3000)
    perm = W.rng(21).permutation(w.n)
    links = perm[: 2 * 800].reshape(-1, 2)
    print("case links disjoint", flush=True)
    g, o = _pair(oracle_lib, w.desc, w.positions)
    for s in (g, o):
        s.set_links(links, 0.3)
        s.take_step(0.05, 10)
    _same(g, o, "disjoint links")


@pytest.mark.parametrize("integrator", [abi.EULER, abi.HEUN], ids=["Euler", "Heun"])
def test_links_shared_endpoints(gpu, oracle_lib, integrator):
    w = W.relaxation(3000)
    d = abi.copy_desc(w.desc)
    d.integrator = integrator
    g_ = W.rng(22)
    links = g_.integers(0, w.n, (6000, 2))
    print(f"case links shared integrator={integrator}", flush=True)
    g, o = _pair(oracle_lib, d, w.positions)
    for s in (g, o):
        s.set_links(links, 0.2)
        s.set_walls([[0, 0, 1, 0.0, 2.0]])
        s.take_step(0.05, 10)
    _same(g, o, "shared links")


def test_division_bitwise(gpu, oracle_lib):
    w = W.relaxation(3000, finder=abi.BINNED_GABRIEL)
    d = abi.copy_desc(w.desc)
    d.capacity = 4000
    print("case division", flush=True)
    g, o = _pair(oracle_lib, d, w.positions)
    r = W.rng(23)
    for it in range(4):
        for s in (g, o):
            s.take_step(0.05, 2)
        u = r.random((g.n, 3))
        ng = g.divide(0.05, 0.1875, u)
        no = o.divide(0.05, 0.1875, u)
        assert ng == no > g.n
        for s in (g, o):
            s.set_agent_count(ng)
        _same(g, o, f"division {it}")
    with pytest.raises(Exception, match="capacity"):
        g.divide(1.0, 0.1, np.zeros((g.n, 3)))


def test_growth_with_wall_example(gpu, oracle_lib):
    """examples/growth_w_wall.cpp:19-78 in small: 500 cells in a half ball, BinnedGabriel,
    PiecewiseLinear(0.7), a wall at z = 0 (strength 10), divisions at rate 0.02, Heun dt 0.1."""
    n0, cap = 500, 2000
    g_ = W.rng(24)
    pos = W.ball(n0, 3.0, g_)
    pos[:, 2] = np.abs(pos[:, 2])
    d = abi.make_desc(n0, capacity=cap, pair_finder=abi.BINNED_GABRIEL, homeostatic_radius=0.7,
                      repulsion_scale=1.0, adhesion_scale=1.0, min_bounds=(-12,) * 3, max_bounds=(12,) * 3)
    print("case growth with wall", flush=True)
    g, o = _pair(oracle_lib, d, pos)
    for s in (g, o):
        s.set_walls([[0, 0, 1, 0, 10.0]])
    n = n0
    for it in range(25):
        for s in (g, o):
            s.take_step(0.1, 1)
        u = g_.random((n, 3))
        n = g.divide(0.02, 0.1875, u)
        assert o.divide(0.02, 0.1875, u) == n
        for s in (g, o):
            s.set_agent_count(n)
    assert n > n0
    _same(g, o, "growth")
    x, _, _ = g.download()
    assert x[:, 2].min() > -0.2


def test_graph_replay_and_batch(gpu, oracle_lib):
    ws = [W.relaxation(n, seed_stream=30 + n, finder=abi.BINNED_GABRIEL) for n in (1500, 3000)]
    print("case walls graphs + batch", flush=True)
    sims = []
    for mode in (abi.GRAPHS_ON, abi.GRAPHS_OFF):
        s = GpuSimulation(ws[1].desc)
        s.set_graph_mode(mode)
        s.upload(ws[1].positions)
        s.set_walls(WALLS)
        s.set_links(np.arange(1000).reshape(-1, 2), 0.2)  # disjoint: order-free atomics
        sims.append(s)
    for steps in (1, 6, 9):
        for s in sims:
            s.take_step(0.05, steps)
        xa, _, _ = sims[0].download()
        xb, _, _ = sims[1].download()
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
    assert sims[0].graph_launches() >= 1
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions)
    b.set_walls(1, WALLS)
    b.take_step([0.05, 0.05], 8)
    for k, w in enumerate(ws):
        o = oracle_lib.OracleSim(w.desc)
        o.upload(w.positions)
        if k == 1:
            o.set_walls(WALLS)
        o.take_step(0.05, 8)
        xb, _, _ = b.download(k)
        xo, _, _ = o.download()
        assert np.array_equal(xb.view(np.uint32), xo.view(np.uint32)), k


def test_walls_survive_capacity_reallocation(gpu, oracle_lib):
    w = W.relaxation(1000)
    print("case walls + set_capacity", flush=True)
    g, o = _pair(oracle_lib, w.desc, w.positions)
    for s in (g, o):
        s.set_walls([[0, 0, 1, 0.0, 5.0]])
        s.take_step(0.05, 3)
        s.set_capacity(3000)
        s.take_step(0.05, 3)
    _same(g, o, "after reallocation")


def test_snapshot_writes_links(gpu, tmp_path):
    """Simulation::write passes the links view last (simulation.hpp:607-611): a t<k>/links
    dataset of (a, b) pairs and, in the XDMF, a polyline topology in place of the polyvertex
    one (hdf5_writer.hpp:163-197, :241-266)."""
    import xml.etree.ElementTree as ET

    from tecs_amd import io

    w = W.relaxation(1000)
    links = np.arange(200, dtype=np.int32).reshape(-1, 2)
    print("case snapshot with links", flush=True)
    s = GpuSimulation(w.desc)
    s.upload(w.positions)
    base = str(tmp_path / "links" / "sim")
    with io.Writer(base) as writer:
        s.write(writer, 0.0)
        s.set_links(links, 0.1)
        s.take_step(0.05, 2)
        s.write(writer, 0.1)
    h5 = base + ".h5"
    assert np.array_equal(io.read_dataset(h5, "t1/links"), links.astype(np.float32))
    grids = ET.parse(base + ".xmf").getroot().find("Domain").find("Grid").findall("Grid")
    assert grids[0].find("Topology").get("TopologyType") == "Polyvertex"
    topo = grids[1].findall("Topology")
    assert [t.get("TopologyType") for t in topo] == ["Polyline"]
    assert topo[0].get("NumberOfElements") == "100"
    assert topo[0].find("DataItem").text.strip() == "sim.h5:t1/links"


def test_batch_links_and_division(gpu, oracle_lib):
    """tecs_batch_set_links / tecs_batch_divide on one member of a batch: that member matches the
    oracle, the other is untouched by them (bitwise as alone)."""
    ws = [W.relaxation(n, seed_stream=40 + n, finder=abi.BINNED_GABRIEL) for n in (1200, 2000)]
    descs = []
    for w in ws:
        d = abi.copy_desc(w.desc)
        d.capacity = w.n + 400
        descs.append(d)
    print("case batch links + division", flush=True)
    b = GpuBatch(descs)
    for k, w in enumerate(ws):
        b.upload(k, w.positions)
    links = np.arange(600, dtype=np.int32).reshape(-1, 2)
    b.set_links(1, links, 0.25)
    o = oracle_lib.OracleSim(descs[1])
    o.upload(ws[1].positions)
    o.set_links(links, 0.25)
    alone = GpuSimulation(descs[0])
    alone.upload(ws[0].positions)
    r = W.rng(41)
    n1 = ws[1].n
    for it in range(3):
        b.take_step([0.05, 0.05], 2)
        o.take_step(0.05, 2)
        alone.take_step(0.05, 2)
        u = r.random((n1, 3))
        nb = b.divide(1, 0.05, 0.1875, u)
        assert o.divide(0.05, 0.1875, u) == nb
        b.set_agent_count(1, nb)
        o.set_agent_count(nb)
        n1 = nb
    xb, _, _ = b.download(1)
    xo, _, _ = o.download()
    assert np.array_equal(xb.view(np.uint32), xo.view(np.uint32))
    xa, _, _ = alone.download()
    x0, _, _ = b.download(0)
    assert np.array_equal(xa.view(np.uint32), x0.view(np.uint32))
