"""HDF5 snapshots: the reference's file layout (include/io/hdf5_writer.hpp) on the host path,
read back through the HDF5_Reader counterpart; no GPU needed."""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tecs_amd import TecsError, abi, io


def test_writer_layout_and_round_trip(tmp_path):
    g = np.random.default_rng(3)
    base = tmp_path / "out" / "sim"
    pos = [g.standard_normal((50, 3)).astype(np.float32) for _ in range(3)]
    pol = [g.standard_normal((50, 2)).astype(np.float32) for _ in range(3)]
    with io.Writer(str(base)) as w:
        w.write_static("types", (np.arange(50) % 2).astype(np.uint8))
        for k in range(3):
            w.write(0.05 * k, k, pos[k], pol[k])
    h5 = str(base) + ".h5"
    for k in range(3):
        assert np.array_equal(io.read_dataset(h5, f"t{k}/position"), pos[k])
        assert np.array_equal(io.read_dataset(h5, f"t{k}/polarity"), pol[k])
    # XDMF companion (hdf5_writer.hpp:131-304): a temporal collection of uniform grids whose
    # geometry points at t<k>/position and whose attributes include polarity and the static data
    root = ET.parse(str(base) + ".xmf").getroot()
    grids = root.find("Domain").find("Grid").findall("Grid")
    assert [gr.get("Name") for gr in grids] == ["t0", "t1", "t2"]
    assert grids[1].find("Time").get("Value") == "0.050000"
    geo = grids[2].find("Geometry")
    assert geo.get("GeometryType") == "XYZ"
    assert geo.find("DataItem").text.strip() == "sim.h5:t2/position"
    names = [a.get("Name") for a in grids[2].findall("Attribute")]
    assert names == ["polarity", "types"]
    assert grids[0].find("Topology").get("NumberOfElements") == "50"


def test_write_static_only_once(tmp_path):
    with io.Writer(str(tmp_path / "s")) as w:
        w.write_static("types", np.zeros(4, np.uint8))
        with pytest.raises(TecsError) as e:
            w.write_static("types", np.zeros(4, np.uint8))
        assert "only write to static data once" in str(e.value)


def test_read_missing_dataset_fails(tmp_path):
    with io.Writer(str(tmp_path / "m"), write_xmf=False) as w:
        w.write(0.0, 0, np.zeros((3, 3), np.float32))
    with pytest.raises(TecsError):
        io.read_dataset(str(tmp_path / "m.h5"), "t0/velocity")
    assert not (tmp_path / "m.xmf").exists()
