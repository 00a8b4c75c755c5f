"""Grid queries on the GPU vs the oracle, through the C ABI (grid.hpp:413-564): the returned agent
indices bitwise, on the grid built from the positions and on a stale grid after steps (the
reference's Grid reads its positions view at query time), for a dense and a sparse grid."""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuSimulation, abi, workloads as W
from tecs_amd._lib import TecsError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["relaxation", "epithelium", "sorting"])
@pytest.mark.parametrize("stale", [False, True], ids=["fresh", "stale"])
def test_queries_match_oracle(gpu, oracle_lib, name, stale):
    w = {"relaxation": W.relaxation(3000), "epithelium": W.epithelium(4000),
         "sorting": W.sorting(20000)}[name]
    print(f"case queries {name} stale={stale}", flush=True)
    g = GpuSimulation(w.desc)
    o = oracle_lib.OracleSim(w.desc)
    for s in (g, o):
        s.upload(w.positions, w.polarities, w.types)
        s.take_step(w.dt, 3 if stale else 1)
        if not stale:
            s.rebuild()
    rng = W.rng(21)
    q = 2000
    pts = (w.positions[rng.integers(0, w.n, q)] + rng.normal(0, 0.5, (q, 3))).astype(np.float32)
    u = rng.uniform(0, 1, q).astype(np.float32)
    for radius_bins in (0, 1, 2):
        assert np.array_equal(g.nearest(pts, radius_bins), o.nearest(pts, radius_bins))
        assert np.array_equal(g.random_in_bins(pts, radius_bins, u), o.random_in_bins(pts, radius_bins, u))
    for radius in (0.5, 1.0, 1.7):
        got = g.random_neighbour(pts, radius, u)
        assert np.array_equal(got, o.random_neighbour(pts, radius, u))
    assert (g.nearest(pts, 1) >= 0).mean() > 0.5


def test_query_errors(gpu):
    w = W.relaxation(500, finder=abi.NAIVE_ALL_PAIRS)
    g = GpuSimulation(w.desc)
    g.upload(w.positions)
    with pytest.raises(TecsError) as e:
        g.nearest(w.positions[:3])
    assert e.value.code == abi.ERR_UNSUPPORTED
    w = W.relaxation(500)
    g = GpuSimulation(w.desc)
    g.upload(w.positions)
    with pytest.raises(TecsError) as e:
        g.nearest(w.positions[:3])
    assert e.value.code == abi.ERR_STATE
