"""The engine's correctly rounded division (tecs_amd/csrc/tecs_device.hpp: rcp_refined,
div_by_rcp, recip_rn, div3_rn; Markstein's method) against IEEE division on the GPU, bit for
bit (tests/cpp/div_check.hip):

  * the refined reciprocal for every d with |d| in [2^-16, 2^16] (2^29 values);
  * the quotient for a sample of d significands (the first, a middle and the last 4096) times
    every x significand in [1, 2); TECS_DIV_EXHAUSTIVE=1 runs all 2^23 d significands (2^46
    pairs, about 30 s);
  * pl_div3_rn, the quotients of PiecewiseLinear, on random displacements and scales;
  * recip_rn and div3_rn, which fall back to IEEE division outside their domains, on random
    bit patterns of every kind (zeros, subnormals, infinities, NaNs) and on operands near the
    domains' edges (signed zeros, both signs of d; the quotient's domain has d > 0, as a signed
    zero numerator over a negative d would get the wrong sign).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent / "cpp" / "_build" / "libdiv_check.so"


@pytest.fixture(scope="module")
def lib(gpu):
    if not LIB.exists():
        pytest.fail(f"{LIB} is missing: run __graft_entry__.build()")
    L = C.CDLL(str(LIB))
    u64p = C.POINTER(C.c_ulonglong)
    L.div_check_rcp.argtypes = [C.c_uint32, C.c_uint32, u64p, u64p]
    L.div_check_div.argtypes = [C.c_uint32, C.c_uint32, u64p, u64p]
    L.div_check_random.argtypes = [C.c_uint64, C.c_uint32, C.c_int, u64p, u64p]
    L.div_check_random_pl.argtypes = [C.c_uint64, C.c_uint32, u64p, u64p]
    return L


def _run(fn, *args):
    bad, first = C.c_ulonglong(0), C.c_ulonglong(0)
    rc = fn(*args, C.byref(bad), C.byref(first))
    assert rc == 0, f"check failed to run (rc {rc})"
    return bad.value, first.value


def test_reciprocal_every_d(lib):
    lo, hi = 0x37800000, 0x47800000  # 2^-16, 2^16
    for sign in (0, 0x80000000):
        print(f"case reciprocal, sign bit {sign >> 31}", flush=True)
        bad, first = _run(lib.div_check_rcp, lo | sign, hi - lo + 1)
        assert bad == 0, f"{bad} reciprocals differ, first d bits {first:#010x}"


def test_quotient_significands(lib):
    n = 1 << 23
    if os.environ.get("TECS_DIV_EXHAUSTIVE") == "1":
        chunks = [(k, 1 << 17) for k in range(0, n, 1 << 17)]
    else:
        chunks = [(0, 4096), (n // 2, 4096), (n - 4096, 4096)]
    total = 0
    for d0, cnt in chunks:
        bad, first = _run(lib.div_check_div, d0, cnt)
        total += cnt
        print(f"d significands [{d0}, {d0 + cnt}): {bad} bad", flush=True)
        assert bad == 0, f"{bad} quotients differ, first (d bits << 32 | x bits) {first:#018x}"
    print(f"{total} d significands x {n} x significands checked", flush=True)


@pytest.mark.parametrize("mode,count", [(0, 1 << 24), (1, 1 << 26)])
def test_random_operands(lib, mode, count):
    bad, first = _run(lib.div_check_random, 2807 + mode, count, mode)
    assert bad == 0, f"{bad} operand sets differ, first index {first}"


def test_piecewise_linear_quotients(lib):
    """pl_div3_rn (its cheaper domain test relies on |v| <= d) on random displacements and
    scales of every magnitude and kind, d computed as the force computes it."""
    bad, first = _run(lib.div_check_random_pl, 4242, 1 << 26)
    assert bad == 0, f"{bad} sets differ, first index {first}"
