"""Pins the oracle's population changes (Simulation::set_agent_count / set_capacity,
simulation.hpp:266-301).

  * growing and uploading the new cells is the same as a fresh simulation of the combined
    state: same pairs and the same step, bitwise (set_agent_count restarts the rebuild counter,
    binned_all_pairs.hpp:60-63, so both build the grid at the next full step)
  * a cell keeps its state while inactive: shrink, step, grow back, and cells that were cut off
    return with the state they had (the views are not touched beyond the active count)
  * set_capacity drops cells beyond a smaller capacity (Kokkos::resize): grown again, they
    are zero
  * a non-full evaluation before the next full step has no grid and fails
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import abi, workloads as W


def _state(w, n):
    return w.positions[:n], None if w.polarities is None else w.polarities[:n], None if w.types is None else w.types[:n]


@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.NAIVE_ALL_PAIRS, abi.BINNED_GABRIEL])
def test_grow_equals_fresh_simulation(oracle_lib, finder):
    w = W.relaxation(1200, finder=finder)
    d = abi.copy_desc(w.desc)
    d.agent_count, d.capacity = 1000, 1500
    a = oracle_lib.OracleSim(d)
    a.upload(*_state(w, 1000))
    a.take_step(w.dt, 3)
    xa, _, _ = a.download()
    a.set_agent_count(1200)
    a.upload_range(1000, w.positions[1000:1200])
    d2 = abi.copy_desc(w.desc)
    d2.capacity = 1500
    b = oracle_lib.OracleSim(d2)
    b.upload(np.concatenate([xa, w.positions[1000:1200]]))
    ra, ca = a.pairs(True)
    rb, cb = b.pairs(True)
    assert np.array_equal(ra, rb) and np.array_equal(ca, cb)
    a.take_step(w.dt, 2)
    b.take_step(w.dt, 2)
    assert np.array_equal(a.download()[0].view(np.uint32), b.download()[0].view(np.uint32))


def test_inactive_cells_keep_their_state(oracle_lib):
    w = W.sorting(6000)
    s = oracle_lib.OracleSim(w.desc)
    s.upload(w.positions, None, w.types)
    s.take_step(w.dt, 2)
    x0, _, v0 = s.download()
    s.set_agent_count(4000)
    s.take_step(w.dt, 2)
    x1, _, _ = s.download()
    s.set_agent_count(6000)
    x2, _, v2 = s.download()
    assert np.array_equal(x2[4000:], x0[4000:]) and np.array_equal(v2[4000:], v0[4000:])
    assert np.array_equal(x2[:4000], x1)
    assert not np.array_equal(x1, x0[:4000])


def test_capacity_drops_cells_and_divides_com(oracle_lib):
    w = W.relaxation(500)
    d = abi.copy_desc(w.desc)
    d.capacity = 800
    s = oracle_lib.OracleSim(d)
    s.upload(w.positions)
    s.upload_range(600, np.ones((100, 3), np.float32))
    s.set_capacity(650)
    s.set_capacity(800)
    s.set_agent_count(800)
    x, _, _ = s.download()
    assert np.all(x[600:650] == 1.0) and np.all(x[650:] == 0.0)


def test_stale_evaluation_after_population_change_fails(oracle_lib):
    w = W.relaxation(800)
    s = oracle_lib.OracleSim(w.desc)
    s.upload(w.positions)
    s.take_step(w.dt, 1)
    s.set_agent_count(700)
    with pytest.raises(RuntimeError):
        s.pairs(False)
    s.pairs(True)
