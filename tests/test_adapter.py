"""The C++ adapter (include/tecs_amd.hpp): the reference's Simulation<Config> / CONFIG_* /
pair_finders / integrators / forces names over the C ABI, driven by tests/cpp/adapter_check.

CPU: Settings defaults, force composition, error behaviour (Simulation without a device throws
tecs_amd::Error). GPU: the adapter's results are bitwise those of the same simulation driven
through the C ABI directly (GpuSimulation), for a polarity model and a type-table model."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from tecs_amd import GpuSimulation, _build, io, workloads as W

BIN = _build.ADAPTER_BIN


def _driver():
    if not BIN.exists():
        _build.build_adapter_check()
    return str(BIN)


def test_adapter_host_checks():
    r = subprocess.run([_driver(), "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks passed" in r.stdout


def test_adapter_rejects_bad_usage():
    r = subprocess.run([_driver(), "epithelium"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def _bounds(desc):
    s = desc.pair_finder_settings
    return np.array(list(s.min_bounds) + list(s.max_bounds), np.float32)


@pytest.mark.gpu
def test_adapter_epithelium_matches_c_abi(tmp_path):
    w = W.epithelium(3000, seed_stream=61)
    steps = 20
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.uint32(w.n).tobytes())
        f.write(_bounds(w.desc).tobytes())
        f.write(np.ascontiguousarray(w.positions, np.float32).tobytes())
        f.write(np.ascontiguousarray(w.polarities, np.float32).tobytes())
    print(f"case adapter epithelium n={w.n} steps={steps}", flush=True)
    snap = tmp_path / "snap"
    r = subprocess.run([_driver(), "epithelium", str(inp), str(out), str(steps), str(w.dt), str(snap)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    x = np.frombuffer(raw[: 12 * w.n], np.float32).reshape(w.n, 3)
    p = np.frombuffer(raw[12 * w.n: 20 * w.n], np.float32).reshape(w.n, 2)
    pairs = int(np.frombuffer(raw[20 * w.n:], np.uint64)[0])
    s = GpuSimulation(w.desc)
    s.upload(w.positions, w.polarities, w.types)
    for _ in range(steps):
        s.take_step(w.dt, 1)
    xa, pa, _ = s.download()
    assert np.array_equal(x.view(np.uint32), xa.view(np.uint32))
    assert np.array_equal(p.view(np.uint32), pa.view(np.uint32))
    assert pairs == s.stats()["pair_interactions"]
    # the adapter's Simulation::write: t0 = initial state, t1 = final state, agent order
    h5 = str(snap) + ".h5"
    assert np.array_equal(io.read_dataset(h5, "t0/position"), w.positions)
    assert np.array_equal(io.read_dataset(h5, "t1/position"), x)
    assert np.array_equal(io.read_dataset(h5, "t1/polarity"), p)


@pytest.mark.gpu
def test_adapter_sorting_matches_c_abi(tmp_path):
    w = W.sorting(8000, seed_stream=62, tau_ab=0.3)
    steps = 20
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.uint32(w.n).tobytes())
        f.write(_bounds(w.desc).tobytes())
        f.write(np.ascontiguousarray(w.positions, np.float32).tobytes())
        f.write(np.ascontiguousarray(w.types, np.uint8).tobytes())
    print(f"case adapter sorting n={w.n} steps={steps}", flush=True)
    r = subprocess.run([_driver(), "sorting", str(inp), str(out), str(steps), str(w.dt), "0.3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    x = np.frombuffer(raw[: 12 * w.n], np.float32).reshape(w.n, 3)
    pairs = int(np.frombuffer(raw[12 * w.n:], np.uint64)[0])
    s = GpuSimulation(w.desc)
    s.upload(w.positions, w.polarities, w.types)
    s.take_step(w.dt, steps)
    xa, _, _ = s.download()
    assert np.array_equal(x.view(np.uint32), xa.view(np.uint32))
    assert pairs == s.stats()["pair_interactions"]
