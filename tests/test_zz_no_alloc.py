"""The allocation-free nearest query (tecs_sim_nearest_noalloc) against tecs_sim_nearest.

Bar: identical agent indices (bit-exact, integer results) on two calls in a row, the second
with fewer points than the first (the fixed staging array keeps the first call's contents; the
second must not see them), on the epithelium's sparse grid and the sorting tissue's dense one;
a call above the array's 65536 points is refused with TECS_ERR_CAPACITY and changes nothing.
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuSimulation, abi, workloads as W
from tecs_amd._lib import TecsError

pytestmark = pytest.mark.gpu


def _queries(w, count, seed, spread=1.5):
    g = np.random.default_rng(seed)
    base = w.positions[g.integers(0, w.n, count)]
    return (base + g.uniform(-spread, spread, base.shape)).astype(np.float32)


@pytest.mark.parametrize("name", ["epithelium", "sorting"])
def test_noalloc_nearest_equals_nearest(gpu, name):
    w = W.epithelium(20000) if name == "epithelium" else W.sorting(40000)
    print(f"case nearest_noalloc {name} {w.n} cells", flush=True)
    g = GpuSimulation(w.desc)
    g.upload(w.positions, w.polarities, w.types)
    g.take_step(w.dt, 3)
    first = _queries(w, 50000, 1)
    second = _queries(w, 7000, 2, spread=3.0)
    for radius in (1, 2):
        a1 = g.nearest(first, radius)
        b1 = g.nearest_noalloc(first, radius)
        assert np.array_equal(a1, b1)
        a2 = g.nearest(second, radius)
        b2 = g.nearest_noalloc(second, radius)
        assert np.array_equal(a2, b2)
        assert (b1 >= 0).any() and (b2 >= 0).any()
    with pytest.raises(TecsError) as e:
        g.nearest_noalloc(_queries(w, 65537, 3))
    assert e.value.code == abi.ERR_CAPACITY
    # a refused call changes nothing: the next call gives the same result
    assert np.array_equal(g.nearest_noalloc(second, 2), b2)
    g.close()
