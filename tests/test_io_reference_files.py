"""The engine's HDF5 reader on the reference's own input files (HDF5_Reader, used by
examples/intercalation_w_gradient.cpp:35-56 and examples/tribolium/main.cpp:53-74).

The files are copied as data fixtures into tests/golden/reference_h5/. The expected contents
come from HDF5's h5dump (tests/golden/make_reference_h5_expected.py): shapes, element types,
SHA-256 of the values, first and last rows. Checked here: every dataset read in its file type
hashes to h5dump's values; dataset shapes (get_dataset_dimensions); the conversions HighFive
would make for a View of another element type (float64 and int64 equal the file values
exactly, cell types read as uint8 equal their int32 values); the tribolium example's
invariant that serosa.h5 has one cell type per point; errors for a short buffer, a missing
dataset and an unknown element type. Needs libhdf5 (the conda one), no GPU.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from tecs_amd import TecsError, abi
from tecs_amd import io as tio
from tecs_amd._lib import load

GOLDEN = Path(__file__).resolve().parent / "golden"
EXPECTED = json.loads((GOLDEN / "reference_h5_expected.json").read_text())
FILE_TYPES = {"H5T_IEEE_F32LE": np.float32, "H5T_STD_I32LE": np.int32}


def _hdf5_or_skip():
    try:
        tio.dataset_shape(GOLDEN / "reference_h5" / "serosa.h5", "POINTS")
    except TecsError as e:
        if e.code == abi.ERR_UNSUPPORTED:
            pytest.skip("libhdf5 not available")
        raise


CASES = [(f, d) for f, ds in EXPECTED.items() for d in ds]


@pytest.mark.parametrize("fname,dataset", CASES, ids=[f"{f}:{d}" for f, d in CASES])
def test_reference_datasets_match_h5dump(fname, dataset):
    _hdf5_or_skip()
    exp = EXPECTED[fname][dataset]
    path = GOLDEN / "reference_h5" / fname
    rows, cols = tio.dataset_shape(path, dataset)
    shape = exp["shape"]
    assert rows == shape[0] and cols == (shape[1] if len(shape) > 1 else 1)
    t = FILE_TYPES[exp["type"]]
    arr = tio.read_dataset(path, dataset, t)
    assert hashlib.sha256(arr.astype(np.dtype(t).newbyteorder("<")).tobytes()).hexdigest() == exp["sha256"]
    assert arr[0].tolist() == exp["first"] and arr[-1].tolist() == exp["last"]
    # conversions on read, as for a View of another element type
    wide = tio.read_dataset(path, dataset, np.float64 if t is np.float32 else np.int64)
    assert np.array_equal(wide, arr.astype(wide.dtype))
    if t is np.int32:
        assert arr.min() >= 0 and arr.max() < 256
        assert np.array_equal(tio.read_dataset(path, dataset, np.uint8), arr.astype(np.uint8))
        assert np.array_equal(tio.read_dataset(path, dataset, np.uint32), arr.astype(np.uint32))


def test_tribolium_inputs_are_consistent():
    """tribolium/main.cpp:53-74 sizes the simulation by serosa's POINTS and reads one cell type
    per point; surface.h5 holds the normals it scales by 15."""
    _hdf5_or_skip()
    ser = GOLDEN / "reference_h5" / "serosa.h5"
    pts = tio.read_dataset(ser, "POINTS")
    types = tio.read_dataset(ser, "cell_type", np.int32)
    assert pts.shape == (3500, 3) and types.shape == (3500, 1)
    normals = tio.read_dataset(GOLDEN / "reference_h5" / "surface.h5", "POINTS")
    assert normals.shape[1] == 3 and np.isfinite(normals).all()


def test_reader_errors():
    _hdf5_or_skip()
    lib = load()
    path = str(GOLDEN / "reference_h5" / "serosa.h5").encode()
    dims = (C.c_int64 * 2)()
    small = np.empty(10, np.float32)
    rc = lib.tecs_read_dataset_typed(path, b"POINTS", 0, small.ctypes.data, small.size, dims)
    assert rc == abi.ERR_CAPACITY and (dims[0], dims[1]) == (3500, 3)
    assert lib.tecs_read_dataset_typed(path, b"no_such", 0, None, 0, dims) == abi.ERR_INVALID
    assert lib.tecs_read_dataset_typed(path, b"POINTS", 99, None, 0, dims) == abi.ERR_INVALID
