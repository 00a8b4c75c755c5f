"""Sweep variants per size class (tecs_batch_set_size_class_variants and the autotuner): the
members of a batch grouped by cell count in powers of four, each group of classes swept by its
own launch over its chunks. Every variant is exact, so each member must end bitwise where the
oracle ends, whatever variant its class uses."""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuBatch, abi, workloads as W

pytestmark = pytest.mark.gpu

# cell counts in size classes 4, 5, 6 and 7 (4^c <= n < 4^(c+1))
SIZES = [600, 2500, 9000, 30000]
STEPS = 20


def _members(finder):
    return [W.sorting(n, seed_stream=20 + i, finder=finder) for i, n in enumerate(SIZES)]


def _check_against_oracle(oracle_lib, b, ws):
    for k, w in enumerate(ws):
        o = oracle_lib.OracleSim(w.desc)
        o.upload(w.positions, w.polarities, w.types)
        o.take_step(w.dt, STEPS)
        xo, _, vo = o.download()
        xb, _, vb = b.download(k)
        assert np.array_equal(xb.view(np.uint32), xo.view(np.uint32)), f"member {k} ({w.n} cells) differs"
        assert np.array_equal(vb.view(np.uint32), vo.view(np.uint32)), f"member {k}: old velocities differ"


@pytest.mark.parametrize("finder", [abi.BINNED_ALL_PAIRS, abi.BINNED_GABRIEL], ids=lambda f: abi.FINDER_NAMES[f])
def test_chosen_class_variants_bitwise(gpu, oracle_lib, finder):
    """Three launch groups (team4, team16, thread) in every sweep of 20 Heun steps."""
    ws = _members(finder)
    print(f"case size classes chosen finder={abi.FINDER_NAMES[finder]}", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    chosen = {4: abi.SWEEP_TEAM4, 5: abi.SWEEP_THREAD, 6: abi.SWEEP_TEAM16, 7: abi.SWEEP_THREAD}
    b.set_size_class_variants(chosen)
    b.take_step([w.dt for w in ws], STEPS)
    classes, variants = b.size_classes()
    assert classes == [4, 5, 6, 7]
    assert variants == {c: abi.SWEEP_NAMES[v] for c, v in chosen.items()}
    _check_against_oracle(oracle_lib, b, ws)


def test_autotuner_per_class_bitwise(gpu, oracle_lib):
    """The autotuner times every variant on each class alone and reports a variant per class
    with members; the results stay bitwise the oracle's."""
    ws = _members(abi.BINNED_ALL_PAIRS)
    print("case size classes autotuned", flush=True)
    b = GpuBatch([w.desc for w in ws])
    for k, w in enumerate(ws):
        b.upload(k, w.positions, w.polarities, w.types)
    b.take_step([w.dt for w in ws], STEPS)
    classes, variants = b.size_classes()
    assert classes == [4, 5, 6, 7]
    assert sorted(variants) == [4, 5, 6, 7]
    assert set(variants.values()) <= set(abi.SWEEP_NAMES.values())
    assert b.autotune_ms(), "the autotuner reports its times"
    _check_against_oracle(oracle_lib, b, ws)


def test_class_variants_rejected_with_com_fixer(gpu):
    w = W.sorting(2000, finder=abi.BINNED_ALL_PAIRS)
    d = abi.copy_desc(w.desc)
    d.com_fixer = abi.GLOBAL_COM_FIXER
    b = GpuBatch([d, d])
    with pytest.raises(RuntimeError):
        b.set_size_class_variants({5: abi.SWEEP_THREAD})
