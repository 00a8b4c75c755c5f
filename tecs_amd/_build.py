"""Build the gfx950 engine library ``tecs_amd/_lib/libtecs_amd.so`` with hipcc.

Flags that matter for results:
  -ffp-contract=off    no fused multiply-add, every operation of the reference expressions
                       rounds once (the pair predicates must match bit for bit)
  -fhip-fp32-correctly-rounded-divide-sqrt
                       IEEE division and square root in single precision
  -Xclang -target-feature -Xclang -packed-fp32-ops
                       no packed-FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libtecs_amd.so"
SOURCES = ["tecs_api.hip", "kernels_sweep.hip", "kernels_team.hip", "kernels_grid.hip", "kernels_query.hip",
           "kernels_delaunay.hip", "kernels_extra.hip", "io_snapshot.hip", "server.hip"]
ARCH = os.environ.get("TECS_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or Path(cand).exists()):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 3) -> Path:
    """Compile every source to an object in parallel, then link the shared library."""
    OUT_DIR.mkdir(exist_ok=True)
    obj_dir = OUT_DIR / "obj"
    obj_dir.mkdir(exist_ok=True)
    hipcc = _hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = obj_dir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, "-I", str(PKG.parent / "include"), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[[p.poll() is None for _, p in procs].index(True)][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append(f"--- {src} ---\n{text}")
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-ldl", "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))


ADAPTER_SRC = PKG.parent / "tests" / "cpp" / "adapter_check.cpp"
ADAPTER_BIN = PKG.parent / "tests" / "cpp" / "_build" / "adapter_check"


DIV_CHECK_SRC = PKG.parent / "tests" / "cpp" / "div_check.hip"
DIV_CHECK_LIB = PKG.parent / "tests" / "cpp" / "_build" / "libdiv_check.so"


def build_div_check() -> Path:
    """The GPU check of the engine's division against IEEE division (tests/test_gpu_fast_div.py),
    compiled with the engine's flags."""
    DIV_CHECK_LIB.parent.mkdir(parents=True, exist_ok=True)
    tmp = DIV_CHECK_LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-shared", str(DIV_CHECK_SRC), "-o", str(tmp)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("division check build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, DIV_CHECK_LIB)
    return DIV_CHECK_LIB


def build_adapter_check() -> Path:
    """The C++ adapter test driver (include/tecs_amd.hpp over the C ABI), linked against the
    engine library with a relative run path so the tree can move."""
    ADAPTER_BIN.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", str(PKG.parent / "include"),
           str(ADAPTER_SRC), "-o", str(ADAPTER_BIN), "-L", str(OUT_DIR), "-ltecs_amd",
           "-Wl,-rpath,$ORIGIN/../../../tecs_amd/_lib"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("adapter check build failed:\n" + r.stdout.decode(errors="replace"))
    return ADAPTER_BIN
