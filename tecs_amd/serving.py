"""numpy-facing handle over the serving loop (tecs_server_*, include/tecs_amd.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from ._lib import check, load


class Server:
    def __init__(self, hbm_budget_bytes: int, max_chunk_steps: int = 32):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.tecs_server_create(int(hbm_budget_bytes), int(max_chunk_steps), C.byref(h)))
        self._h = h
        self._descs = {}

    def submit(self, desc, positions, polarities=None, types=None, dt=0.05, n_steps=100) -> int:
        p = np.ascontiguousarray(positions, np.float32)
        q = None if polarities is None else np.ascontiguousarray(polarities, np.float32)
        t = None if types is None else np.ascontiguousarray(types, np.uint8)
        rid = C.c_int64(0)
        d = abi.copy_desc(desc)
        check(self._lib.tecs_server_submit(self._h, C.byref(d), p.ctypes.data,
                                           None if q is None else q.ctypes.data,
                                           None if t is None else t.ctypes.data,
                                           float(dt), int(n_steps), C.byref(rid)))
        self._descs[rid.value] = d
        return int(rid.value)

    def advance(self):
        run = C.c_int32(0)
        wait = C.c_int32(0)
        check(self._lib.tecs_server_advance(self._h, C.byref(run), C.byref(wait)))
        return int(run.value), int(wait.value)

    def info(self, rid: int) -> dict:
        i = abi.RequestInfo()
        check(self._lib.tecs_server_request_info(self._h, int(rid), C.byref(i)))
        return {k: getattr(i, k) for k, _ in abi.RequestInfo._fields_}

    def cancel(self, rid: int):
        """A pending request leaves at once, a running one at the end of its group's chunk."""
        check(self._lib.tecs_server_cancel(self._h, int(rid)))

    def result(self, rid: int):
        d = self._descs[rid]
        n = int(d.agent_count)
        pos = np.empty((n, 3), np.float32)
        pol = np.empty((n, 2), np.float32) if d.fields & abi.FIELD_POLARITY else None
        check(self._lib.tecs_server_result(self._h, int(rid), pos.ctypes.data,
                                           None if pol is None else pol.ctypes.data))
        return pos, pol

    def stats(self) -> dict:
        s = abi.ServerStats()
        check(self._lib.tecs_server_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in abi.ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tecs_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
