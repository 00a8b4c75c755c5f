"""Serving-loop admission checks that need no GPU (validation happens at submit)."""
from __future__ import annotations

import numpy as np
import pytest

from tecs_amd import TecsError, abi, workloads as W
from tecs_amd.serving import Server


def test_submit_accepts_stale_grid_requests():
    """Requests with a rebuild period are served (grouped by period, re-formed at rebuild
    points), so submit accepts them."""
    w = W.relaxation(500)
    d = abi.copy_desc(w.desc)
    d.pair_finder_settings.rebuild_every_n = 3
    s = Server(1 << 30)
    rid = s.submit(d, w.positions, dt=w.dt, n_steps=10)
    assert s.info(rid)["state"] == abi.REQUEST_PENDING


def test_cancel_pending_requests():
    """A pending request leaves at once: it is CANCELLED, has no result, cannot be cancelled
    again, and the others stay pending (no GPU work: nothing is admitted before advance)."""
    w = W.relaxation(500)
    s = Server(1 << 30)
    a = s.submit(w.desc, w.positions, dt=w.dt, n_steps=10)
    b = s.submit(w.desc, w.positions, dt=w.dt, n_steps=10)
    s.cancel(a)
    assert s.info(a)["state"] == abi.REQUEST_CANCELLED
    assert s.info(b)["state"] == abi.REQUEST_PENDING
    with pytest.raises(TecsError) as e:
        s.result(a)
    assert e.value.code == abi.ERR_STATE
    with pytest.raises(TecsError) as e:
        s.cancel(a)
    assert e.value.code == abi.ERR_STATE
    st = s.stats()
    assert st["cancelled"] == 1 and st["pending"] == 1 and st["hbm_bytes_in_use"] == 0
    with pytest.raises(TecsError):
        s.cancel(12345)


def test_submit_rejects_requests_beyond_the_budget():
    w = W.relaxation(5000)
    s = Server(1 << 16)  # 64 KiB
    with pytest.raises(TecsError) as e:
        s.submit(w.desc, w.positions, dt=w.dt, n_steps=10)
    assert e.value.code == abi.ERR_CAPACITY


def test_submitted_request_waits_pending():
    w = W.relaxation(500)
    s = Server(1 << 30)
    rid = s.submit(w.desc, w.positions, dt=w.dt, n_steps=10)
    info = s.info(rid)
    assert info["state"] == abi.REQUEST_PENDING and info["steps_done"] == 0
    assert info["hbm_bytes"] > 500 * 16
    with pytest.raises(TecsError):
        s.result(rid)
