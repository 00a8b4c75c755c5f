"""HDF5 snapshots in the reference's layout (io::HDF5_Writer, include/io/hdf5_writer.hpp)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load


class Writer:
    """HDF5_Writer(path, settings): <path>.h5 (+ <path>.xmf) with groups t<step>/position etc."""

    def __init__(self, path: str, write_xmf: bool = True):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.tecs_writer_open(str(path).encode(), 1 if write_xmf else 0, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def write(self, time: float, step: int, positions, polarities=None):
        p = np.ascontiguousarray(positions, np.float32)
        q = None if polarities is None else np.ascontiguousarray(polarities, np.float32)
        check(self._lib.tecs_writer_write(self._h, float(time), int(step), p.shape[0], p.ctypes.data,
                                          None if q is None else q.ctypes.data))

    def write_static(self, label: str, data):
        d = np.ascontiguousarray(data, np.uint8)
        check(self._lib.tecs_writer_write_static(self._h, label.encode(), d.size, d.ctypes.data))

    def flush(self):
        check(self._lib.tecs_writer_flush(self._h))

    def close(self):
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            check(self._lib.tecs_writer_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DTYPES = {
    np.dtype(np.float32): 0,
    np.dtype(np.float64): 1,
    np.dtype(np.int32): 2,
    np.dtype(np.uint32): 3,
    np.dtype(np.uint8): 4,
    np.dtype(np.int64): 5,
}


def dataset_shape(h5_path: str, name: str):
    """HDF5_Reader::get_dataset_dimensions (include/io/hdf5_reader.hpp:43-53): (rows, columns),
    columns 1 for a rank-1 dataset."""
    lib = load()
    dims = (C.c_int64 * 2)()
    check(lib.tecs_read_dataset_typed(str(h5_path).encode(), name.encode(), 0, None, 0, dims))
    return int(dims[0]), int(dims[1])


def read_dataset(h5_path: str, name: str, dtype=np.float32) -> np.ndarray:
    """HDF5_Reader::read_dataset (include/io/hdf5_reader.hpp:66-71) into an array of `dtype`
    (float32, float64, int32, uint32, uint8 or int64; HDF5 converts from the file's type), shape
    (rows, columns)."""
    lib = load()
    dt = np.dtype(dtype)
    if dt not in _DTYPES:
        raise ValueError(f"unsupported element type {dt}")
    rows, cols = dataset_shape(h5_path, name)
    out = np.empty((rows, cols), dt)
    dims = (C.c_int64 * 2)()
    check(lib.tecs_read_dataset_typed(str(h5_path).encode(), name.encode(), _DTYPES[dt], out.ctypes.data,
                                      out.size, dims))
    return out
