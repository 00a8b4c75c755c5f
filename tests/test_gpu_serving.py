"""Serving-loop robustness on the GPU: requests with a rebuild period, a request whose state
becomes non-finite, cancellation; and the non-finite flag of a single simulation.

Bars: every finished request's result bitwise equals the same simulation run alone (and, for
PiecewiseLinear models, the oracle's); a request that overflows fails (state FAILED, no
result) and the other requests in its batch are bitwise unaffected; a cancelled running
request leaves and the others are bitwise unaffected.

The failure is forced from the host with an infinite repulsion scale: any pair within the
cutoff gives an infinite or NaN force through ordinary float arithmetic, so the cells' positions
become non-finite; the kernels clamp non-finite coordinates into the grid's edge bins like any
position outside the bounds.
"""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import GpuSimulation, TecsError, abi, workloads as W
from tecs_amd.serving import Server

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _alone(w, desc, steps):
    s = GpuSimulation(desc)
    s.upload(w.positions, w.polarities, w.types)
    s.take_step(w.dt, steps)
    x, p, _ = s.download()
    s.close()
    return x, p


def _drive(srv, reqs, ids, limit=5000):
    rounds = 0
    while True:
        for k, (w, d, steps, arrive) in enumerate(reqs):
            if arrive == rounds and k not in ids:
                ids[k] = srv.submit(d, w.positions, w.polarities, w.types, w.dt, steps)
        srv.advance()
        rounds += 1
        if len(ids) == len(reqs) and all(srv.info(i)["state"] >= abi.REQUEST_DONE for i in ids.values()):
            return rounds
        assert rounds < limit


def _with_period(w, period):
    d = abi.copy_desc(w.desc)
    d.pair_finder_settings.rebuild_every_n = period
    return d


def test_stale_grid_requests_equal_alone_and_the_oracle(gpu, oracle_lib):
    """Rebuild periods 3 and 2 (two groups), step counts that are not multiples of the period,
    arrivals over several rounds and a budget that holds about half the requests."""
    base = [
        (W.relaxation(3000, seed_stream=60), 3, 37, 0),
        (W.sorting(8000, seed_stream=61), 3, 23, 0),
        (W.relaxation(1500, seed_stream=62), 3, 14, 2),
        (W.sorting(5000, seed_stream=63), 2, 41, 1),
        (W.epithelium(3000, seed_stream=64), 2, 19, 1),
        (W.relaxation(2500, seed_stream=65), 3, 9, 4),
        (W.sorting(4000, seed_stream=66), 3, 30, 6),
    ]
    reqs = [(w, _with_period(w, p), steps, arrive) for w, p, steps, arrive in base]
    sizes = []
    for w, d, _, _ in reqs:
        probe = Server(1 << 40)
        sizes.append(probe.info(probe.submit(d, w.positions, w.polarities, w.types, w.dt, 1))["hbm_bytes"])
        probe.close()
    srv = Server(sum(sorted(sizes)[-4:]), max_chunk_steps=4)
    ids = {}
    print("case served stale-grid requests", [(w.name, w.n, s, p) for w, p, s, _ in base], flush=True)
    rounds = _drive(srv, reqs, ids)
    st = srv.stats()
    print("server stats", st, "rounds", rounds, flush=True)
    assert st["completed"] == len(reqs) and st["regroups"] >= 4 and st["hbm_bytes_in_use"] == 0
    for k, (w, d, steps, _) in enumerate(reqs):
        assert srv.info(ids[k])["state"] == abi.REQUEST_DONE
        xs, ps = srv.result(ids[k])
        xa, pa = _alone(w, d, steps)
        assert np.array_equal(_bits(xs), _bits(xa)), f"request {k} vs alone"
        if pa is not None:
            assert np.array_equal(_bits(ps), _bits(pa)), f"request {k} vs alone (polarity)"
        if w.polarities is None:
            o = oracle_lib.OracleSim(d)
            o.upload(w.positions, None, w.types)
            o.take_step(w.dt, steps)
            assert np.array_equal(_bits(xs), _bits(o.download()[0])), f"request {k} vs oracle"


def test_overflowing_request_fails_alone(gpu):
    good = [(W.relaxation(3000, seed_stream=70), 25), (W.sorting(6000, seed_stream=71), 31)]
    wb = W.relaxation(2000, seed_stream=72)
    bad = abi.copy_desc(wb.desc)
    bad.force.repulsion_scale = float("inf")
    reqs = [(good[0][0], good[0][0].desc, good[0][1], 0), (wb, bad, 40, 0),
            (good[1][0], good[1][0].desc, good[1][1], 0)]
    srv = Server(200 * 10**9, max_chunk_steps=8)
    ids = {}
    print("case served request that overflows", flush=True)
    _drive(srv, reqs, ids)
    assert srv.info(ids[1])["state"] == abi.REQUEST_FAILED
    with pytest.raises(TecsError) as e:
        srv.result(ids[1])
    assert e.value.code == abi.ERR_STATE
    st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == 2 and st["hbm_bytes_in_use"] == 0
    for k in (0, 2):
        w, d, steps, _ = reqs[k]
        assert srv.info(ids[k])["state"] == abi.REQUEST_DONE
        xs, _ = srv.result(ids[k])
        xa, _ = _alone(w, d, steps)
        assert np.array_equal(_bits(xs), _bits(xa)), f"request {k}"


def test_cancel_a_running_request(gpu):
    reqs = [(W.relaxation(3000, seed_stream=80), 60), (W.sorting(6000, seed_stream=81), 60),
            (W.relaxation(2000, seed_stream=82), 60)]
    srv = Server(200 * 10**9, max_chunk_steps=4)
    ids = [srv.submit(w.desc, w.positions, None, w.types, w.dt, n) for w, n in reqs]
    print("case cancel a running request", flush=True)
    for _ in range(3):
        srv.advance()
    assert srv.info(ids[1])["state"] == abi.REQUEST_RUNNING
    srv.cancel(ids[1])
    rounds = 0
    while not all(srv.info(i)["state"] >= abi.REQUEST_DONE for i in ids):
        srv.advance()
        rounds += 1
        assert rounds < 1000
    assert srv.info(ids[1])["state"] == abi.REQUEST_CANCELLED
    assert srv.info(ids[1])["steps_done"] < 60
    for k in (0, 2):
        w, n = reqs[k]
        xs, _ = srv.result(ids[k])
        xa, _ = _alone(w, w.desc, n)
        assert np.array_equal(_bits(xs), _bits(xa)), f"request {k}"
    assert srv.stats()["cancelled"] == 1


@pytest.mark.parametrize("name,variant", [("relaxation", abi.SWEEP_THREAD), ("relaxation", abi.SWEEP_TEAM8),
                                          ("sorting", abi.SWEEP_TEAM8),
                                          ("epithelium", abi.SWEEP_TEAM8), ("epithelium", abi.SWEEP_THREAD)])
def test_simulation_nonfinite_flag(gpu, name, variant):
    """The flag stays clear on a normal run and is set once an overflow reaches the state, on
    the fused epilogues of both sweep kinds (Euler and Heun corrector)."""
    w = {"relaxation": W.relaxation(3000), "sorting": W.sorting(6000), "epithelium": W.epithelium(3000)}[name]
    print(f"case non-finite flag {name} variant {variant}", flush=True)
    g = GpuSimulation(w.desc)
    g.set_sweep_variant(variant)
    g.upload(w.positions, w.polarities, w.types)
    g.take_step(w.dt, 5)
    assert not g.nonfinite()
    f = abi.ForceDesc()
    abi.C.memmove(abi.C.addressof(f), abi.C.addressof(w.desc.force), abi.C.sizeof(abi.ForceDesc))
    f.repulsion_scale = float("inf")
    g.set_force(f)
    g.take_step(w.dt, 5)
    assert g.nonfinite()
    g.close()
