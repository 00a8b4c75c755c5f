"""ctypes wrapper of the CPU oracle (oracle/tecs_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the benchmark's CPU
baseline leg as the checker. The engine never imports it.

OracleSim mirrors tecs_amd.engine.GpuSimulation method for method, so a parity test runs the
same calls against both and compares the results.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from tecs_amd import abi

HERE = Path(__file__).resolve().parent
SRC = HERE / "tecs_oracle.cpp"
LIB_PATH = HERE / "_build" / "libtecs_oracle.so"

# -ffp-contract=off: every operation of the reference expressions rounds once (no FMA);
# no -ffast-math, no -march=native (the x86-64 baseline has no FMA instruction anyway).
CXXFLAGS = ["-std=c++17", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared"]

_lib = None


def build() -> Path:
    LIB_PATH.parent.mkdir(exist_ok=True)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [os.environ.get("CXX", "g++"), *CXXFLAGS, "-o", str(tmp), str(SRC)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_set_threads": (C.c_int, [C.c_int]),
        "orc_create": (C.c_int, [C.POINTER(abi.SimDesc), C.POINTER(P)]),
        "orc_destroy": (C.c_int, [P]),
        "orc_upload": (C.c_int, [P, P, P, P]),
        "orc_download": (C.c_int, [P, P, P, P]),
        "orc_take_step": (C.c_int, [P, C.c_double, C.c_int]),
        "orc_rebuild": (C.c_int, [P]),
        "orc_evaluate_force": (C.c_int, [P, C.c_int, P, P]),
        "orc_pairs": (C.c_int, [P, C.c_int, P, P, C.c_int64, C.POINTER(C.c_int64)]),
        "orc_get_stats": (C.c_int, [P, C.POINTER(abi.Stats)]),
        "orc_forces_f64": (C.c_int, [P, C.c_int, P, P, P, P, P, P]),
        "orc_vector_op": (C.c_int, [C.c_int, P, P]),
        "orc_polarity_op": (C.c_int, [C.c_int, P, P]),
        "orc_com_fix": (C.c_int, [P, C.c_uint, P]),
        "orc_upload_range": (C.c_int, [P, C.c_uint, C.c_uint, P, P, P]),
        "orc_set_agent_count": (C.c_int, [P, C.c_uint]),
        "orc_set_capacity": (C.c_int, [P, C.c_uint]),
        "orc_query": (C.c_int, [P, C.c_int, C.c_int, P, C.c_int, C.c_float, P, P]),
        "orc_set_walls": (C.c_int, [P, P, C.c_int]),
        "orc_set_links": (C.c_int, [P, P, C.c_int64, C.c_float]),
        "orc_divide": (C.c_int, [P, C.c_float, C.c_float, P, C.POINTER(C.c_uint)]),
        "orc_set_force": (C.c_int, [P, C.POINTER(abi.ForceDesc)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {load().orc_last_error().decode()}")


def set_threads(n: int) -> int:
    return int(load().orc_set_threads(int(n)))


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleSim:
    def __init__(self, desc: abi.SimDesc):
        self._lib = load()
        self.desc = abi.copy_desc(desc)
        self.n = int(desc.agent_count)
        self.has_pol = bool(desc.fields & abi.FIELD_POLARITY)
        self.has_type = bool(desc.fields & abi.FIELD_TYPE)
        h = C.c_void_p()
        _check(self._lib.orc_create(C.byref(self.desc), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, positions, polarities=None, types=None):
        p = np.ascontiguousarray(positions, np.float32).reshape(self.n, 3)
        q = np.ascontiguousarray(polarities, np.float32).reshape(self.n, 2) if polarities is not None else None
        t = np.ascontiguousarray(types, np.uint8).reshape(self.n) if types is not None else None
        _check(self._lib.orc_upload(self._h, _ptr(p), _ptr(q), _ptr(t)))

    def download(self):
        pos = np.empty((self.n, 3), np.float32)
        pol = np.empty((self.n, 2), np.float32) if self.has_pol else None
        ov = np.empty((self.n, 3), np.float32)
        _check(self._lib.orc_download(self._h, _ptr(pos), _ptr(pol), _ptr(ov)))
        return pos, pol, ov

    def take_step(self, dt: float, n_steps: int = 1):
        _check(self._lib.orc_take_step(self._h, float(dt), int(n_steps)))

    def upload_range(self, first: int, positions, polarities=None, types=None):
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        k = p.shape[0]
        q = np.ascontiguousarray(polarities, np.float32).reshape(k, 2) if polarities is not None else None
        t = np.ascontiguousarray(types, np.uint8).reshape(k) if types is not None else None
        _check(self._lib.orc_upload_range(self._h, int(first), k, _ptr(p), _ptr(q), _ptr(t)))

    def set_agent_count(self, n: int):
        _check(self._lib.orc_set_agent_count(self._h, int(n)))
        self.n = int(n)

    def set_capacity(self, capacity: int):
        _check(self._lib.orc_set_capacity(self._h, int(capacity)))

    def set_force(self, force):
        """The tecs_force_desc (abi.ForceDesc) of the following steps."""
        _check(self._lib.orc_set_force(self._h, C.byref(force)))
        C.memmove(C.addressof(self.desc.force), C.addressof(force), C.sizeof(abi.ForceDesc))

    def set_walls(self, walls):
        """walls: rows (a, b, c, d, strength)."""
        w = np.ascontiguousarray(walls, np.float32).reshape(-1, 5)
        _check(self._lib.orc_set_walls(self._h, _ptr(w) if w.size else None, w.shape[0]))

    def set_links(self, links, strength: float):
        lk = np.ascontiguousarray(links, np.int32).reshape(-1, 2)
        _check(self._lib.orc_set_links(self._h, _ptr(lk) if lk.size else None, lk.shape[0], float(strength)))

    def divide(self, rate: float, offset: float, uniforms) -> int:
        u = np.ascontiguousarray(uniforms, np.float64).reshape(self.n, 3)
        out = C.c_uint(0)
        _check(self._lib.orc_divide(self._h, float(rate), float(offset), _ptr(u), C.byref(out)))
        return int(out.value)

    def _query(self, kind, points, radius_bins, radius, uniforms):
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float32).reshape(p.shape[0])
        out = np.empty(p.shape[0], np.int32)
        _check(self._lib.orc_query(self._h, int(kind), p.shape[0], _ptr(p), int(radius_bins), float(radius), _ptr(u), _ptr(out)))
        return out

    def nearest(self, points, search_radius: int = 1):
        return self._query(0, points, search_radius, 0.0, None)

    def random_in_bins(self, points, radius_bins: int, uniforms):
        return self._query(1, points, radius_bins, 0.0, uniforms)

    def random_neighbour(self, points, radius: float, uniforms):
        return self._query(2, points, 0, radius, uniforms)

    def rebuild(self):
        _check(self._lib.orc_rebuild(self._h))

    def evaluate_force(self, is_full_step: bool = True):
        dpos = np.empty((self.n, 3), np.float32)
        dpol = np.empty((self.n, 2), np.float32) if self.has_pol else None
        _check(self._lib.orc_evaluate_force(self._h, 1 if is_full_step else 0, _ptr(dpos), _ptr(dpol)))
        return dpos, dpol

    def pairs(self, is_full_step: bool = True):
        count = C.c_int64(0)
        _check(self._lib.orc_pairs(self._h, 1 if is_full_step else 0, None, None, 0, C.byref(count)))
        row_ptr = np.empty(self.n + 1, np.int64)
        cols = np.empty(max(count.value, 1), np.int32)
        _check(
            self._lib.orc_pairs(self._h, 1 if is_full_step else 0, _ptr(row_ptr), _ptr(cols), cols.size, C.byref(count))
        )
        return row_ptr, cols[: count.value]

    def stats(self) -> dict:
        s = abi.Stats()
        _check(self._lib.orc_get_stats(self._h, C.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in abi.Stats._fields_}

    def forces_f64(self, row_ptr, cols, symmetric=None):
        """FP64 forces of the current stage-0 state over the given pair set, and per component
        the sum of absolute pair contributions. A symmetric pair set (the symmetric finders)
        applies each pair's value v to i and -v to j; by default the finder decides."""
        if symmetric is None:
            symmetric = int(self.desc.pair_finder) in abi.SYMMETRIC_FINDERS
        rp = np.ascontiguousarray(row_ptr, np.int64)
        cc = np.ascontiguousarray(cols, np.int32)
        if cc.size == 0:
            cc = np.zeros(1, np.int32)
        dpos = np.empty((self.n, 3), np.float64)
        spos = np.empty((self.n, 3), np.float64)
        dpol = np.empty((self.n, 2), np.float64)
        spol = np.empty((self.n, 2), np.float64)
        _check(self._lib.orc_forces_f64(self._h, 1 if symmetric else 0, _ptr(rp), _ptr(cc), _ptr(dpos),
                                        _ptr(dpol), _ptr(spos), _ptr(spol)))
        return dpos, dpol, spos, spol


def vector_op(op: int, a, b=None) -> float:
    """0: dot(a, b), 1: length_squared(a), 2: distance_to_squared(a, b) (float32)."""
    buf = np.zeros(6, np.float32)
    buf[:3] = a
    if b is not None:
        buf[3:] = b
    out = np.zeros(1, np.float32)
    _check(load().orc_vector_op(int(op), _ptr(buf), _ptr(out)))
    return out[0]


def polarity_op(op: int, *args) -> np.ndarray:
    """Polarity_ arithmetic of the oracle; see orc_polarity_op in tecs_oracle.cpp."""
    buf = np.concatenate([np.atleast_1d(np.asarray(a, np.float32)) for a in args]).astype(np.float32)
    buf = np.concatenate([buf, np.zeros(8, np.float32)])
    out = np.zeros(8, np.float32)
    _check(load().orc_polarity_op(int(op), _ptr(buf), _ptr(out)))
    return out


def com_fix(deltas, capacity: int) -> np.ndarray:
    d = np.zeros((capacity, 3), np.float32)
    d[: len(deltas)] = deltas
    out = np.zeros(3, np.float32)
    _check(load().orc_com_fix(_ptr(d), int(capacity), _ptr(out)))
    return out


def force_bar_usage(f32, f64, scale, rel: float = 1e-5) -> dict:
    """How much of force_excess's bar each term carries: `floor_needed` counts the components
    whose error exceeds rel * S_i alone (the cell's own contributions), which only the F_rms
    term admits; `max_own` is the largest |F32 - F64| / (rel * S_i) and `max_floor` the largest
    |F32 - F64| / (rel * F_rms)."""
    f64 = np.asarray(f64, np.float64)
    d = np.abs(np.asarray(f32, np.float64) - f64)
    own = rel * np.asarray(scale, np.float64)
    rms = np.sqrt(np.mean(f64 * f64, axis=0, keepdims=True))
    with np.errstate(divide="ignore", invalid="ignore"):
        r_own = np.where(own > 0, d / own, np.where(d > 0, np.inf, 0.0))
        r_floor = np.where(rms > 0, d / (rel * rms), np.where(d > 0, np.inf, 0.0))
    need = d > own
    return {"components": int(d.size), "floor_needed": int(need.sum()),
            "fraction": float(need.mean()) if d.size else 0.0,
            "max_own": float(r_own.max()) if d.size else 0.0,
            "max_floor": float(r_floor.max()) if d.size else 0.0}


def force_excess(f32, f64, scale, rel: float = 1e-5) -> float:
    """How far FP32 forces exceed the bar |F32 - F64| <= rel * (S_i + F_rms) per component.

    S_i is the cell's sum of absolute pair contributions (bounds the FP32 summation error) and
    F_rms the root-mean-square FP64 force of the field (the force scale: a cell whose pair
    forces cancel, or whose pairs sit at the rest length, still carries the rounding error of
    its distances, which is a fraction of the field's force scale, not of its own tiny force).
    Returns max(|F32 - F64| - bound): <= 0 means the bar holds."""
    f64 = np.asarray(f64, np.float64)
    rms = np.sqrt(np.mean(f64 * f64, axis=0, keepdims=True)) if f64.size else 0.0
    bound = rel * (np.asarray(scale, np.float64) + rms)
    return float((np.abs(np.asarray(f32, np.float64) - f64) - bound).max()) if f64.size else 0.0
