"""Pins the oracle's generic wall forces, link forces and cell division (examples/growth_w_wall.cpp,
growth_w_squeeze.cpp, intercalation.cpp; integrators/base.hpp:116-230; types/plane.hpp) against
closed forms in float64.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from tecs_amd import abi


def _sim(oracle_lib, pos, finder=abi.NAIVE_ALL_PAIRS, cutoff=1.0, capacity=None, integrator=abi.EULER):
    d = abi.make_desc(len(pos), capacity=capacity, cutoff_distance=cutoff, pair_finder=finder,
                      integrator=integrator)
    s = oracle_lib.OracleSim(d)
    s.upload(np.asarray(pos, np.float32))
    return s


def test_wall_pushes_cells_behind_it(oracle_lib):
    # cells far apart (no pair within the cutoff); plane 2z + 1 = 0 normalises to z + 0.5 = 0
    pos = [[0.0, 0.0, -1.0], [5.0, 0.0, -0.5], [10.0, 0.0, 0.3], [15.0, 0.0, -0.75]]
    s = _sim(oracle_lib, pos)
    s.set_walls([[0.0, 0.0, 2.0, 1.0, 10.0]])
    s.take_step(0.01, 1)
    _, _, v = s.download()
    want = np.array([[0, 0, 5.0], [0, 0, 0], [0, 0, 0], [0, 0, 2.5]])
    assert np.allclose(v, want, rtol=1e-6, atol=0)
    x, _, _ = s.download()
    assert np.allclose(x[:, 2], np.array(pos)[:, 2] + want[:, 2] * 0.01, rtol=0, atol=1e-6)


def test_walls_add_up_and_tilted_normal(oracle_lib):
    """growth_w_squeeze's three walls; a cell outside two of them gets both pushes; a tilted plane
    (1, 1, 0, 0) pushes along (1, 1, 0) / sqrt(2)."""
    pos = [[0.0, -0.7, -0.5], [5.0, 0.0, 0.4], [-3.0, -1.0, 0.0]]
    s = _sim(oracle_lib, pos)
    s.set_walls([[0, 0, 1, 0.25, 0.5], [0, 0, -1, 0.25, 0.5], [0, 1, 0, 0.5, 0.5]])
    s.take_step(0.001, 1)
    _, _, v = s.download()
    want0 = np.array([0, 0.5 * 0.2, 0.5 * 0.25])   # y: sd = -0.2, z: sd = -0.25
    want1 = np.array([0, 0, -0.5 * 0.15])           # upper wall: sd = 0.25 - 0.4
    want2 = np.array([0, 0.5 * 0.5, 0])             # y: sd = -0.5
    assert np.allclose(v, [want0, want1, want2], rtol=1e-6, atol=1e-7)
    t = _sim(oracle_lib, [[-1.0, -2.0, 0.0]])
    t.set_walls([[1, 1, 0, 0, 2.0]])
    t.take_step(0.001, 1)
    _, _, v = t.download()
    sd = (-1.0 - 2.0) / math.sqrt(2)
    assert np.allclose(v[0], np.array([1, 1, 0]) / math.sqrt(2) * (-sd * 2.0), rtol=1e-6)


def test_wall_validation(oracle_lib):
    s = _sim(oracle_lib, [[0.0, 0.0, 0.0]])
    with pytest.raises(RuntimeError, match="zero"):
        s.set_walls([[0, 0, 0, 1, 1]])
    with pytest.raises(RuntimeError, match="count"):
        s.set_walls(np.zeros((abi.MAX_WALLS + 1, 5)) + [0, 0, 1, 0, 1])


def test_link_pulls_endpoints_together(oracle_lib):
    pos = [[0.0, 0.0, 0.0], [3.0, 4.0, 0.0], [20.0, 0.0, 0.0], [20.0, 0.0, 2.0]]
    s = _sim(oracle_lib, pos, cutoff=0.5)
    s.set_links([[0, 1], [2, 3], [3, 2], [1, 1]], 0.5)
    s.take_step(0.1, 1)
    _, _, v = s.download()
    u01 = np.array([3.0, 4.0, 0.0]) / 5.0
    assert np.allclose(v[0], 0.5 * u01, rtol=1e-6) and np.allclose(v[1], -0.5 * u01, rtol=1e-6)
    # two links between 2 and 3, one each way: both pull the pair together
    assert np.allclose(v[2], [0, 0, 1.0], rtol=1e-6) and np.allclose(v[3], [0, 0, -1.0], rtol=1e-6)
    # a single link: Newton's third law holds exactly
    t = _sim(oracle_lib, [[0.1, 0.2, 0.3], [1.3, -0.7, 2.9]], cutoff=0.5)
    t.set_links([[1, 0]], 1.7)
    t.take_step(0.1, 1)
    _, _, w = t.download()
    assert np.array_equal(w[0], -w[1]) and np.abs(w).max() > 0


def test_links_beyond_the_active_count_are_skipped(oracle_lib):
    s = _sim(oracle_lib, [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0]], cutoff=0.5, capacity=4)
    s.set_agent_count(2)
    s.set_links([[0, 2], [1, 3]], 1.0)
    s.take_step(0.1, 1)
    _, _, v = s.download()
    assert np.all(v == 0)
    with pytest.raises(RuntimeError, match="capacity"):
        s.set_links([[0, 4]], 1.0)


def test_divide_places_daughters_in_parent_order(oracle_lib):
    rng = np.random.default_rng(5)
    pos = rng.uniform(-3, 3, (40, 3)).astype(np.float32)
    s = _sim(oracle_lib, pos, capacity=60)
    u = rng.random((40, 3))
    u[:, 0] = np.where(np.arange(40) % 7 == 3, 0.001, 0.9)  # cells 3, 10, 17, 24, 31, 38 divide
    n_new = s.divide(0.005, 0.1875, u)
    parents = np.flatnonzero(u[:, 0] <= 0.005)
    assert n_new == 40 + len(parents)
    s.set_agent_count(n_new)
    x, _, _ = s.download()
    assert np.array_equal(x[:40], pos)
    theta = np.arccos(2.0 * u[parents, 1] - 1.0)
    phi = u[parents, 2] * 2.0 * math.pi
    dirs = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], 1)
    assert np.allclose(x[40:], pos[parents] + 0.1875 * dirs, rtol=0, atol=2e-6)
    assert np.allclose(np.linalg.norm(x[40:] - pos[parents], axis=1), 0.1875, atol=2e-6)


def test_divide_refuses_beyond_capacity(oracle_lib):
    pos = np.zeros((10, 3), np.float32) + np.arange(10, dtype=np.float32)[:, None]
    s = _sim(oracle_lib, pos, capacity=12)
    u = np.zeros((10, 3))
    with pytest.raises(RuntimeError, match="capacity"):
        s.divide(0.5, 0.2, u)
    s.set_agent_count(12)
    x, _, _ = s.download()
    assert np.all(x[10:] == 0)


def test_growth_with_wall_runs(oracle_lib):
    """examples/growth_w_wall.cpp in small: BinnedGabriel, a wall at z = 0 (strength 10),
    divisions, set_agent_count. Cells end on the positive side and the count grows."""
    rng = np.random.default_rng(11)
    n0, cap = 200, 800
    r = 3.0 * np.cbrt(rng.random(n0))
    v = rng.standard_normal((n0, 3))
    pos = (v / np.linalg.norm(v, axis=1, keepdims=True) * r[:, None]).astype(np.float32)
    pos[:, 2] = np.abs(pos[:, 2])
    d = abi.make_desc(n0, capacity=cap, pair_finder=abi.BINNED_GABRIEL, homeostatic_radius=0.7,
                      repulsion_scale=1.0, adhesion_scale=1.0, min_bounds=(-10,) * 3, max_bounds=(10,) * 3)
    s = oracle_lib.OracleSim(d)
    s.upload(pos)
    s.set_walls([[0, 0, 1, 0, 10.0]])
    n = n0
    for _ in range(30):
        s.take_step(0.1, 1)
        n = s.divide(0.02, 0.1875, rng.random((n, 3)))
        s.set_agent_count(n)
    x, _, _ = s.download()
    assert n > n0 and len(x) == n
    assert np.isfinite(x).all() and x[:, 2].min() > -0.2
