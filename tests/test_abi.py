"""The C ABI library without a GPU: it loads, exports every declared entry point, its ctypes
mirror has the C layout, and descriptor validation rejects bad input before touching a device.
"""
from __future__ import annotations

import ctypes as C
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

from tecs_amd import _lib, abi

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "tecs_amd.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tecs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(n for n, _, _ in _lib.PROTOTYPES) == names


def test_engine_does_not_contain_the_oracle():
    out = subprocess.run(["nm", "-D", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    assert "orc_" not in out


def test_struct_layout_matches_c():
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "tecs_amd.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("sizeof.tecs_sim_desc %zu\n", sizeof(tecs_sim_desc));
  printf("sizeof.tecs_force_desc %zu\n", sizeof(tecs_force_desc));
  printf("sizeof.tecs_pair_finder_settings %zu\n", sizeof(tecs_pair_finder_settings));
  printf("sizeof.tecs_stats %zu\n", sizeof(tecs_stats));
  P(tecs_sim_desc, agent_count) P(tecs_sim_desc, capacity) P(tecs_sim_desc, fields)
  P(tecs_sim_desc, cutoff_distance) P(tecs_sim_desc, pair_finder)
  P(tecs_sim_desc, pair_finder_settings) P(tecs_sim_desc, integrator) P(tecs_sim_desc, com_fixer)
  P(tecs_sim_desc, force) P(tecs_sim_desc, track_old_velocities)
  P(tecs_force_desc, n_types) P(tecs_force_desc, adhesion_table) P(tecs_force_desc, bending_strength)
  P(tecs_force_desc, alignment_strength)
  P(tecs_pair_finder_settings, bin_size_scale) P(tecs_pair_finder_settings, rebuild_every_n)
  P(tecs_pair_finder_settings, gabriel_coefficient)
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        src = Path(d) / "layout.c"
        src.write_text(prog)
        exe = Path(d) / "layout"
        subprocess.run(["gcc", "-I", str(HEADER.parent), "-o", str(exe), str(src)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    types = {
        "tecs_sim_desc": abi.SimDesc,
        "tecs_force_desc": abi.ForceDesc,
        "tecs_pair_finder_settings": abi.PairFinderSettings,
        "tecs_stats": abi.Stats,
    }
    for key, val in got.items():
        a, b = key.split(".")
        if a == "sizeof":
            assert C.sizeof(types[b]) == int(val), key
        else:
            assert getattr(types[a], b).offset == int(val), key


def test_defaults_match_reference_settings():
    lib = _lib.load()
    d = abi.SimDesc()
    lib.tecs_default_sim_desc(C.byref(d))
    ps = d.pair_finder_settings
    # binned_all_pairs.hpp:20-25, binned_gabriel.hpp:20-26
    assert list(ps.min_bounds) == [-20.0] * 3 and list(ps.max_bounds) == [20.0] * 3
    assert ps.bin_size_scale == 1.0 and ps.rebuild_every_n == 0
    assert abs(ps.gabriel_coefficient - 0.8) < 1e-7
    # RelaxForce = PiecewiseLinear(0.8, 2, 1), relax_force.hpp:27-29
    f = d.force
    assert (abs(f.homeostatic_radius - 0.8) < 1e-7, f.repulsion_scale, f.adhesion_scale) == (True, 2.0, 1.0)
    assert d.track_old_velocities == 1
    py = abi.make_desc(0)
    assert C.string_at(C.addressof(py.pair_finder_settings), C.sizeof(abi.PairFinderSettings)) == \
        C.string_at(C.addressof(ps), C.sizeof(abi.PairFinderSettings))


def _create(desc):
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.tecs_sim_create(C.byref(desc), None, C.byref(h))
    if rc == 0:
        lib.tecs_sim_destroy(h)
    return rc, lib.tecs_last_error().decode()


@pytest.mark.parametrize("mutate,needle", [
    (lambda d: setattr(d, "capacity", 1), "capacity"),
    (lambda d: setattr(d, "pair_finder", 9), "pair_finder"),
    (lambda d: (setattr(d, "pair_finder", 8), setattr(d, "capacity", 513)), "512"),
    (lambda d: setattr(d, "pair_finder", -1), "pair_finder"),
    (lambda d: setattr(d, "integrator", 5), "integrator"),
    (lambda d: setattr(d, "com_fixer", 3), "com_fixer"),
    (lambda d: setattr(d, "cutoff_distance", 0.0), "cutoff"),
    (lambda d: setattr(d.force, "n_types", 2), "TECS_FIELD_TYPE"),
    (lambda d: setattr(d.force, "n_types", 9), "n_types"),
    (lambda d: setattr(d.force, "bending_strength", 1.0), "TECS_FIELD_POLARITY"),
    (lambda d: setattr(d.pair_finder_settings, "bin_size_scale", 0.0), "bin_size_scale"),
    (lambda d: d.pair_finder_settings.max_bounds.__setitem__(0, -30.0), "max_bounds"),
    (lambda d: setattr(d.pair_finder_settings, "rebuild_every_n", -1), "rebuild_every_n"),
])
def test_invalid_descriptors_rejected(mutate, needle):
    d = abi.make_desc(10)
    mutate(d)
    rc, msg = _create(d)
    assert rc == abi.ERR_INVALID
    assert needle in msg


def test_create_without_device_fails_loudly():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    rc, msg = _create(abi.make_desc(10))
    assert rc in (abi.ERR_DEVICE, abi.ERR_OOM)
    assert msg


def test_batch_rejects_mixed_configurations():
    lib = _lib.load()
    a = abi.make_desc(10)
    b = abi.make_desc(10, integrator=abi.EULER)
    arr = (abi.SimDesc * 2)(a, b)
    h = C.c_void_p()
    assert lib.tecs_batch_create(arr, 2, None, C.byref(h)) == abi.ERR_INVALID
    assert b"share" in lib.tecs_last_error()
