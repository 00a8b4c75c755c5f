"""Population changes on the GPU vs the oracle, through the C ABI (Simulation::set_agent_count /
set_capacity, simulation.hpp:266-301).

Bars: bit-exact positions and old velocities for PiecewiseLinear models (growth with new cells
uploaded, shrinking and growing back with the inactive cells' state kept, a capacity grown past
the allocation); the polarity model and the symmetric finder to the trajectory tolerance
(median |dx| <= 1e-5, p99 <= 1e-3, max <= 0.05); batch members bitwise as alone.
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, GpuSimulation, abi, workloads as W
from tecs_amd._lib import TecsError

pytestmark = pytest.mark.gpu


def _both(oracle_lib, desc, w, n):
    g = GpuSimulation(desc)
    o = oracle_lib.OracleSim(desc)
    pol = None if w.polarities is None else w.polarities[:n]
    typ = None if w.types is None else w.types[:n]
    for s in (g, o):
        s.upload(w.positions[:n], pol, typ)
    return g, o


def _same(g, o, what):
    xg, pg, vg = g.download()
    xo, po, vo = o.download()
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32)), what
    assert np.array_equal(vg.view(np.uint32), vo.view(np.uint32)), what
    if pg is not None:
        assert np.array_equal(pg.view(np.uint32), po.view(np.uint32)), what


@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.NAIVE_ALL_PAIRS, abi.BINNED_GABRIEL],
                         ids=lambda f: abi.FINDER_NAMES[f])
@pytest.mark.parametrize("integrator", [abi.EULER, abi.HEUN], ids=["Euler", "Heun"])
def test_grow_with_new_cells(gpu, oracle_lib, finder, integrator):
    w = W.relaxation(3500 if finder != abi.NAIVE_ALL_PAIRS else 1500, finder=finder)
    n0 = w.n - 500
    d = abi.copy_desc(w.desc)
    d.agent_count, d.capacity, d.integrator = n0, w.n + 100, integrator
    print(f"case grow finder={abi.FINDER_NAMES[finder]} integrator={integrator}", flush=True)
    g, o = _both(oracle_lib, d, w, n0)
    for s in (g, o):
        s.take_step(w.dt, 5)
        s.set_agent_count(w.n)
        s.upload_range(n0, w.positions[n0:])
        s.take_step(w.dt, 5)
    _same(g, o, "after growth")
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)


def test_shrink_and_grow_back_keeps_inactive_state(gpu, oracle_lib):
    w = W.sorting(20000)
    print("case shrink / grow sorting 20000", flush=True)
    g, o = _both(oracle_lib, w.desc, w, w.n)
    for s in (g, o):
        s.take_step(w.dt, 3)
        s.set_agent_count(15000)
        s.take_step(w.dt, 3)
        s.set_agent_count(20000)
        s.take_step(w.dt, 3)
    _same(g, o, "shrink / grow")
    assert g.stats()["steps"] == o.stats()["steps"] == 9


def test_capacity_past_the_allocation(gpu, oracle_lib):
    """GlobalComFixer divides by the capacity; the reduction order differs (1e-5 bar as for the
    COM fixer elsewhere). The capacity grows past the allocation (new device buffers), then
    shrinks within it, dropping inactive cells."""
    w = W.relaxation(2000)
    d = abi.copy_desc(w.desc)
    d.com_fixer, d.integrator = abi.GLOBAL_COM_FIXER, abi.HEUN
    print("case capacity 2000 -> 5000 -> 3000", flush=True)
    g, o = _both(oracle_lib, d, w, w.n)
    for s in (g, o):
        s.take_step(w.dt, 3)
        s.set_capacity(5000)
        s.upload_range(4000, np.full((500, 3), 7.0, np.float32))
        s.take_step(w.dt, 3)
        s.set_capacity(4200)
    assert g.population() == (2000, 4200)
    for s in (g, o):
        s.set_capacity(4600)
        s.set_agent_count(4600)
    xg, _, _ = g.download()
    xo, _, _ = o.download()
    assert np.all(xg[4000:4200] == 7.0) and np.all(xg[4200:] == 0.0)
    assert np.abs(xg - xo).max() <= 1e-5
    assert g.stats()["steps"] == 6 and g.stats()["pair_interactions"] > 0


def test_polarity_and_symmetric_growth_tolerance(gpu, oracle_lib):
    for finder in (abi.BINNED_GABRIEL, abi.BINNED_ALL_PAIRS_SYMMETRIC):
        w = W.epithelium(4000, finder=finder)
        n0 = 3500
        d = abi.copy_desc(w.desc)
        d.agent_count = n0
        print(f"case growth tolerance finder={abi.FINDER_NAMES[finder]}", flush=True)
        g, o = _both(oracle_lib, d, w, n0)
        for s in (g, o):
            s.take_step(w.dt, 5)
            s.set_agent_count(w.n)
            s.upload_range(n0, w.positions[n0:], w.polarities[n0:])
            s.take_step(w.dt, 5)
        xg, pg, _ = g.download()
        xo, po, _ = o.download()
        for a, b in ((xg, xo), (pg, po)):
            dd = np.abs(a.astype(np.float64) - b).max(axis=1)
            assert np.median(dd) <= 1e-5 and np.quantile(dd, 0.99) <= 1e-3 and dd.max() <= 0.05


def test_batch_member_population(gpu):
    ws = [W.relaxation(n) for n in (1500, 3000, 800)]
    descs = []
    for w in ws:
        d = abi.copy_desc(w.desc)
        d.capacity = w.n + 400
        descs.append(d)
    print("case batch population", flush=True)
    b = GpuBatch(descs)
    alone = [GpuSimulation(d) for d in descs]
    for k, w in enumerate(ws):
        b.upload(k, w.positions)
        alone[k].upload(w.positions)
    extra = (W.rng(9).uniform(-3, 3, (300, 3))).astype(np.float32)
    b.take_step([w.dt for w in ws], 4)
    b.set_agent_count(1, ws[1].n + 300)
    b.upload_range(1, ws[1].n, extra)
    b.take_step([w.dt for w in ws], 4)
    for k, s in enumerate(alone):
        s.take_step(ws[k].dt, 4)
        if k == 1:
            s.set_agent_count(ws[1].n + 300)
            s.upload_range(ws[1].n, extra)
        s.take_step(ws[k].dt, 4)
        xa, _, va = s.download()
        xb, _, vb = b.download(k)
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32)), k
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), k


def test_batch_refuses_shared_rebuild_period(gpu):
    ws = [W.relaxation(n) for n in (600, 700)]
    descs = []
    for w in ws:
        d = abi.copy_desc(w.desc)
        d.pair_finder_settings.rebuild_every_n = 3
        descs.append(d)
    b = GpuBatch(descs)
    with pytest.raises(TecsError) as e:
        b.set_agent_count(0, 500)
    assert e.value.code == abi.ERR_UNSUPPORTED


def test_capacity_within_the_allocation_uses_every_slot(gpu, oracle_lib):
    """A capacity raised to the member's whole slot range (2000 -> 2048, no reallocation) and
    an agent count at that capacity: uploads, steps, downloads and the pair export cover all
    2048 cells (the host staging buffers hold the slot range, not the creation capacity)."""
    w = W.relaxation(2048)
    d = abi.copy_desc(w.desc)
    d.agent_count, d.capacity = 2000, 2000
    print("case capacity 2000 -> 2048 within the allocation", flush=True)
    g, o = _both(oracle_lib, d, w, 2000)
    for s in (g, o):
        s.take_step(w.dt, 3)
        s.set_capacity(2048)
        s.set_agent_count(2048)
        s.upload_range(2000, w.positions[2000:])
        s.take_step(w.dt, 3)
    _same(g, o, "capacity 2048")
    rg, cg = g.pairs(True)
    ro, co = o.pairs(True)
    assert rg.shape == (2049,)
    assert np.array_equal(rg, ro) and np.array_equal(cg, co)
