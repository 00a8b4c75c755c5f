"""Writes reference_h5_expected.json: the contents of the reference's example input files
(copied here as data fixtures from examples/data/sphere_ic.h5 and
examples/tribolium/initial_data/{serosa,surface}.h5), as HDF5's own h5dump tool prints them.

For every dataset: its file type, shape, and the SHA-256 of its values in the file's element
type (float32 / int32, little endian), parsed from `h5dump -m %.9g` (nine significant digits
round-trip every float32), plus the first and last rows. tests/test_io_reference_files.py reads
the files with the engine's reader and compares. Run from the repository root:

    python tests/golden/make_reference_h5_expected.py
"""
from __future__ import annotations

import hashlib
import json
import re
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
FILES = ["sphere_ic.h5", "serosa.h5", "surface.h5"]
H5DUMP = "/opt/conda/bin/h5dump"
TYPES = {"H5T_IEEE_F32LE": np.float32, "H5T_STD_I32LE": np.int32}


def datasets(path: Path):
    head = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True).stdout
    return re.findall(r'DATASET "([^"]+)"', head)


def dump(path: Path, name: str):
    text = subprocess.run([H5DUMP, "-m", "%.9g", "-y", "-w", "0", "-d", name, str(path)],
                          capture_output=True, text=True, check=True).stdout
    ftype = re.search(r"DATATYPE\s+(\S+)", text).group(1)
    shape = [int(v) for v in re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", text).group(1).split(",")]
    body = text[text.index("DATA {") + len("DATA {"):text.rindex("}", 0, text.rindex("}"))]
    values = np.array([float(v) for v in re.findall(r"[-+0-9.eEinfa]+", body)])
    arr = values.astype(TYPES[ftype]).reshape(shape)
    return ftype, shape, arr


def main():
    out = {}
    for f in FILES:
        path = HERE / "reference_h5" / f
        entry = {}
        for name in datasets(path):
            ftype, shape, arr = dump(path, name)
            rows = arr.reshape(shape[0], -1)
            entry[name] = {
                "type": ftype,
                "shape": shape,
                "sha256": hashlib.sha256(np.ascontiguousarray(arr).astype(arr.dtype.newbyteorder("<")).tobytes()).hexdigest(),
                "first": rows[0].tolist(),
                "last": rows[-1].tolist(),
            }
        out[f] = entry
    (HERE / "reference_h5_expected.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
