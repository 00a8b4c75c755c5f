"""numpy-facing handles over the C ABI: GpuSimulation (tecs_sim) and GpuBatch (tecs_batch).

Arrays are in the reference's agent order and layout: positions (n, 3) float32, polarities
(n, 2) float32 (theta, phi), types (n,) uint8.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from ._lib import check, load


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _u8(a, n):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.shape != (n,):
        raise ValueError(f"expected shape ({n},), got {a.shape}")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data


class GpuSimulation:
    """One simulation on the GPU (the drop-in for Simulation<Config>::take_step)."""

    def __init__(self, desc: abi.SimDesc):
        self._lib = load()
        self.desc = abi.copy_desc(desc)
        self.n = int(desc.agent_count)
        self.has_pol = bool(desc.fields & abi.FIELD_POLARITY)
        self.has_type = bool(desc.fields & abi.FIELD_TYPE)
        h = C.c_void_p()
        check(self._lib.tecs_sim_create(C.byref(self.desc), None, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tecs_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, positions, polarities=None, types=None):
        p = _f32(positions, (self.n, 3))
        q = _f32(polarities, (self.n, 2)) if (polarities is not None and self.has_pol) else None
        t = _u8(types, self.n) if (types is not None and self.has_type) else None
        check(self._lib.tecs_sim_upload(self._h, _ptr(p), _ptr(q), _ptr(t)))

    def download(self):
        pos = np.empty((self.n, 3), np.float32)
        pol = np.empty((self.n, 2), np.float32) if self.has_pol else None
        ov = np.empty((self.n, 3), np.float32)
        check(self._lib.tecs_sim_download(self._h, _ptr(pos), _ptr(pol), _ptr(ov)))
        return pos, pol, ov

    def take_step(self, dt: float, n_steps: int = 1):
        check(self._lib.tecs_sim_take_step(self._h, float(dt), int(n_steps)))

    def upload_range(self, first: int, positions, polarities=None, types=None):
        """Cells first .. first + len(positions) - 1 (active or not)."""
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        k = p.shape[0]
        q = _f32(polarities, (k, 2)) if (polarities is not None and self.has_pol) else None
        t = _u8(types, k) if (types is not None and self.has_type) else None
        check(self._lib.tecs_sim_upload_range(self._h, int(first), k, _ptr(p), _ptr(q), _ptr(t)))

    def set_agent_count(self, n: int):
        """Simulation::set_agent_count (simulation.hpp:285-301)."""
        check(self._lib.tecs_sim_set_agent_count(self._h, int(n)))
        self.n = int(n)
        self.desc.agent_count = self.n

    def set_capacity(self, capacity: int):
        """Simulation::set_capacity (simulation.hpp:266-274)."""
        check(self._lib.tecs_sim_set_capacity(self._h, int(capacity)))
        self.desc.capacity = int(capacity)

    def set_walls(self, walls):
        """Generic wall forces: rows (a, b, c, d, strength) of planes a x + b y + c z + d = 0
        (examples/growth_w_wall.cpp:44-48)."""
        arr, count = abi.make_walls(walls)
        check(self._lib.tecs_sim_set_walls(self._h, C.addressof(arr), count))

    def set_links(self, links, strength: float):
        """Links (rows a, b of agent indices) and the link force strength
        (examples/intercalation.cpp)."""
        lk = np.ascontiguousarray(links, np.int32).reshape(-1, 2)
        check(self._lib.tecs_sim_set_links(self._h, _ptr(lk) if lk.size else None, lk.shape[0], float(strength)))

    def divide(self, rate: float, offset: float, uniforms) -> int:
        """Cell division (examples/growth_w_wall.cpp:52-64) with three draws per active cell;
        returns the count to pass to set_agent_count."""
        u = np.ascontiguousarray(uniforms, np.float64).reshape(self.n, 3)
        out = C.c_uint32(0)
        check(self._lib.tecs_sim_divide(self._h, float(rate), float(offset), _ptr(u), C.byref(out)))
        return int(out.value)

    def nearest(self, points, search_radius: int = 1):
        """Grid::get_nearest_point_index for every row of points (grid.hpp:413-441)."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.empty(p.shape[0], np.int32)
        check(self._lib.tecs_sim_nearest(self._h, p.shape[0], _ptr(p), int(search_radius), _ptr(out)))
        return out

    def nearest_noalloc(self, points, search_radius: int = 1):
        """nearest() without any device allocation (tecs_sim_nearest_noalloc): at most 65536
        points per call, staged in a fixed device array."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.empty(p.shape[0], np.int32)
        check(self._lib.tecs_sim_nearest_noalloc(self._h, p.shape[0], _ptr(p), int(search_radius), _ptr(out)))
        return out

    def random_in_bins(self, points, radius_bins: int, uniforms):
        """Grid::get_random_point_index_in_neighbourhood_bins with the caller's draws."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        u = np.ascontiguousarray(uniforms, np.float32).reshape(p.shape[0])
        out = np.empty(p.shape[0], np.int32)
        check(self._lib.tecs_sim_random_in_bins(self._h, p.shape[0], _ptr(p), int(radius_bins), _ptr(u), _ptr(out)))
        return out

    def random_neighbour(self, points, radius: float, uniforms):
        """Grid::get_random_point_index_in_neighbourhood with the caller's draws."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        u = np.ascontiguousarray(uniforms, np.float32).reshape(p.shape[0])
        out = np.empty(p.shape[0], np.int32)
        check(self._lib.tecs_sim_random_neighbour(self._h, p.shape[0], _ptr(p), float(radius), _ptr(u), _ptr(out)))
        return out

    def population(self):
        n, cap = C.c_uint32(0), C.c_uint32(0)
        check(self._lib.tecs_sim_get_population(self._h, C.byref(n), C.byref(cap)))
        return int(n.value), int(cap.value)

    def set_force(self, force):
        """The tecs_force_desc (abi.ForceDesc) of the following steps."""
        check(self._lib.tecs_sim_set_force(self._h, C.byref(force)))

    def write(self, writer, time: float):
        """Simulation::write(time): asynchronous snapshot into an io.Writer."""
        check(self._lib.tecs_sim_write(self._h, writer.handle, float(time)))

    def synchronize(self):
        check(self._lib.tecs_sim_synchronize(self._h))

    def evaluate_force(self, is_full_step: bool = True):
        dpos = np.empty((self.n, 3), np.float32)
        dpol = np.empty((self.n, 2), np.float32) if self.has_pol else None
        check(self._lib.tecs_sim_evaluate_force(self._h, 1 if is_full_step else 0, _ptr(dpos), _ptr(dpol)))
        return dpos, dpol

    def pairs(self, is_full_step: bool = True):
        """CSR (row_ptr[n+1] int64, cols int32) of the ordered pairs in visiting order."""
        count = C.c_int64(0)
        check(self._lib.tecs_sim_pairs(self._h, 1 if is_full_step else 0, None, None, 0, C.byref(count)))
        row_ptr = np.empty(self.n + 1, np.int64)
        cols = np.empty(max(count.value, 1), np.int32)
        check(
            self._lib.tecs_sim_pairs(
                self._h, 1 if is_full_step else 0, _ptr(row_ptr), _ptr(cols), cols.size, C.byref(count)
            )
        )
        return row_ptr, cols[: count.value]

    def rebuild(self):
        check(self._lib.tecs_sim_rebuild(self._h))

    def nonfinite(self) -> bool:
        """A step ended with a non-finite position or polarity since creation."""
        f = C.c_int32(0)
        check(self._lib.tecs_sim_get_nonfinite(self._h, C.byref(f)))
        return bool(f.value)

    def stats(self) -> dict:
        s = abi.Stats()
        check(self._lib.tecs_sim_get_stats(self._h, C.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in abi.Stats._fields_}

    def reset_stats(self):
        check(self._lib.tecs_sim_reset_stats(self._h))

    def set_sweep_variant(self, variant: int):
        check(self._lib.tecs_sim_set_sweep_variant(self._h, int(variant)))

    def sweep_variant(self) -> int:
        v = C.c_int32(0)
        check(self._lib.tecs_sim_get_sweep_variant(self._h, C.byref(v)))
        return int(v.value)

    def autotune_ms(self) -> dict:
        arr = (C.c_double * 6)()
        check(self._lib.tecs_sim_get_autotune(self._h, arr))
        return {abi.SWEEP_NAMES[v]: arr[v] for v in abi.SWEEP_NAMES if arr[v] > 0}

    def set_graph_mode(self, mode: int):
        check(self._lib.tecs_sim_set_graph_mode(self._h, int(mode)))

    def graph_launches(self) -> int:
        n = C.c_int64(0)
        check(self._lib.tecs_sim_get_graph_launches(self._h, C.byref(n)))
        return int(n.value)

    def set_kernel_timing(self, enabled: bool):
        check(self._lib.tecs_sim_set_kernel_timing(self._h, 1 if enabled else 0))

    def kernel_timing(self, kernel_class: int):
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        check(self._lib.tecs_sim_get_kernel_timing(self._h, int(kernel_class), C.byref(ms), C.byref(n)))
        return float(ms.value), int(n.value)


class GpuBatch:
    """Many independent simulations advanced together (one launch per kernel per substep)."""

    def __init__(self, descs):
        self._lib = load()
        self.descs = [abi.copy_desc(d) for d in descs]
        arr = (abi.SimDesc * len(self.descs))(*self.descs)
        self._arr = arr
        h = C.c_void_p()
        check(self._lib.tecs_batch_create(arr, len(self.descs), None, C.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.descs)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.tecs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, k, positions, polarities=None, types=None):
        d = self.descs[k]
        n = int(d.agent_count)
        p = _f32(positions, (n, 3))
        q = _f32(polarities, (n, 2)) if (polarities is not None and d.fields & abi.FIELD_POLARITY) else None
        t = _u8(types, n) if (types is not None and d.fields & abi.FIELD_TYPE) else None
        check(self._lib.tecs_batch_upload(self._h, int(k), _ptr(p), _ptr(q), _ptr(t)))

    def upload_range(self, k, first: int, positions, polarities=None, types=None):
        d = self.descs[k]
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        m = p.shape[0]
        q = _f32(polarities, (m, 2)) if (polarities is not None and d.fields & abi.FIELD_POLARITY) else None
        t = _u8(types, m) if (types is not None and d.fields & abi.FIELD_TYPE) else None
        check(self._lib.tecs_batch_upload_range(self._h, int(k), int(first), m, _ptr(p), _ptr(q), _ptr(t)))

    def set_agent_count(self, k, n: int):
        check(self._lib.tecs_batch_set_agent_count(self._h, int(k), int(n)))
        self.descs[k].agent_count = int(n)

    def set_capacity(self, k, capacity: int):
        check(self._lib.tecs_batch_set_capacity(self._h, int(k), int(capacity)))
        self.descs[k].capacity = int(capacity)

    def set_walls(self, k, walls):
        arr, count = abi.make_walls(walls)
        check(self._lib.tecs_batch_set_walls(self._h, int(k), C.addressof(arr), count))

    def set_links(self, k, links, strength: float):
        lk = np.ascontiguousarray(links, np.int32).reshape(-1, 2)
        check(self._lib.tecs_batch_set_links(self._h, int(k), _ptr(lk) if lk.size else None, lk.shape[0],
                                             float(strength)))

    def divide(self, k, rate: float, offset: float, uniforms) -> int:
        n = int(self.descs[k].agent_count)
        u = np.ascontiguousarray(uniforms, np.float64).reshape(n, 3)
        out = C.c_uint32(0)
        check(self._lib.tecs_batch_divide(self._h, int(k), float(rate), float(offset), _ptr(u), C.byref(out)))
        return int(out.value)

    def download(self, k):
        d = self.descs[k]
        n = int(d.agent_count)
        pos = np.empty((n, 3), np.float32)
        pol = np.empty((n, 2), np.float32) if d.fields & abi.FIELD_POLARITY else None
        ov = np.empty((n, 3), np.float32)
        check(self._lib.tecs_batch_download(self._h, int(k), _ptr(pos), _ptr(pol), _ptr(ov)))
        return pos, pol, ov

    def write(self, k, writer, time: float):
        check(self._lib.tecs_batch_write(self._h, int(k), writer.handle, float(time)))

    def take_step(self, dts, n_steps: int = 1):
        dts = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, np.float64), (len(self.descs),)))
        check(self._lib.tecs_batch_take_step(self._h, _ptr(dts), int(n_steps)))

    def set_force(self, index: int, force):
        check(self._lib.tecs_batch_set_force(self._h, int(index), C.byref(force)))

    def synchronize(self):
        check(self._lib.tecs_batch_synchronize(self._h))

    def nonfinite(self):
        """Per member: a step ended with a non-finite position or polarity."""
        f = np.zeros(len(self.descs), np.int32)
        check(self._lib.tecs_batch_get_nonfinite(self._h, _ptr(f)))
        return f.astype(bool)

    def stats(self) -> dict:
        s = abi.Stats()
        check(self._lib.tecs_batch_get_stats(self._h, C.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in abi.Stats._fields_}

    def reset_stats(self):
        check(self._lib.tecs_batch_reset_stats(self._h))

    def set_sweep_variant(self, variant: int):
        check(self._lib.tecs_batch_set_sweep_variant(self._h, int(variant)))

    def sweep_variant(self) -> int:
        v = C.c_int32(0)
        check(self._lib.tecs_batch_get_sweep_variant(self._h, C.byref(v)))
        return int(v.value)

    def autotune_ms(self) -> dict:
        arr = (C.c_double * 6)()
        check(self._lib.tecs_batch_get_autotune(self._h, arr))
        return {abi.SWEEP_NAMES[v]: arr[v] for v in abi.SWEEP_NAMES if arr[v] > 0}

    def size_classes(self):
        """(size class of every member, {class: variant name}) as the autotuner chose them."""
        mc = (C.c_int32 * max(1, len(self.descs)))()
        cv = (C.c_int32 * abi.MAX_SIZE_CLASSES)()
        check(self._lib.tecs_batch_get_size_classes(self._h, mc, cv))
        return [int(mc[k]) for k in range(len(self.descs))], {
            c: abi.SWEEP_NAMES[cv[c]] for c in range(abi.MAX_SIZE_CLASSES) if cv[c] >= 0}

    def set_size_class_variants(self, class_variant: dict):
        """{size class: variant} for the classes with members (tecs_batch_set_size_class_variants)."""
        cv = (C.c_int32 * abi.MAX_SIZE_CLASSES)(*([-1] * abi.MAX_SIZE_CLASSES))
        for c, v in class_variant.items():
            cv[int(c)] = int(v)
        check(self._lib.tecs_batch_set_size_class_variants(self._h, cv))

    def set_graph_mode(self, mode: int):
        check(self._lib.tecs_batch_set_graph_mode(self._h, int(mode)))

    def graph_launches(self) -> int:
        n = C.c_int64(0)
        check(self._lib.tecs_batch_get_graph_launches(self._h, C.byref(n)))
        return int(n.value)

    def set_kernel_timing(self, enabled: bool):
        check(self._lib.tecs_batch_set_kernel_timing(self._h, 1 if enabled else 0))

    def kernel_timing(self, kernel_class: int):
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        check(self._lib.tecs_batch_get_kernel_timing(self._h, int(kernel_class), C.byref(ms), C.byref(n)))
        return float(ms.value), int(n.value)
