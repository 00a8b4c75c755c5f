"""Loader for the gfx950 engine library (tecs_amd/_lib/libtecs_amd.so).

The engine has no CPU fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from . import abi

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libtecs_amd.so"

_lib = None

# (name, restype, argtypes) for every entry point declared in include/tecs_amd.h
_P = C.c_void_p
_I32 = C.c_int32
_I64P = C.POINTER(C.c_int64)
_F32P = C.c_void_p  # numpy buffers are passed as raw addresses
PROTOTYPES = [
    ("tecs_default_pair_finder_settings", None, [C.POINTER(abi.PairFinderSettings)]),
    ("tecs_default_sim_desc", None, [C.POINTER(abi.SimDesc)]),
    ("tecs_abi_version", C.c_int, []),
    ("tecs_last_error", C.c_char_p, []),
    ("tecs_device_count", C.c_int, [C.POINTER(C.c_int32)]),
    ("tecs_sim_create", C.c_int, [C.POINTER(abi.SimDesc), _P, C.POINTER(_P)]),
    ("tecs_sim_destroy", C.c_int, [_P]),
    ("tecs_sim_upload", C.c_int, [_P, _F32P, _F32P, _P]),
    ("tecs_sim_download", C.c_int, [_P, _F32P, _F32P, _F32P]),
    ("tecs_sim_upload_range", C.c_int, [_P, C.c_uint32, C.c_uint32, _F32P, _F32P, _P]),
    ("tecs_sim_set_agent_count", C.c_int, [_P, C.c_uint32]),
    ("tecs_sim_set_capacity", C.c_int, [_P, C.c_uint32]),
    ("tecs_sim_get_population", C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("tecs_sim_nearest", C.c_int, [_P, _I32, _F32P, _I32, _P]),
    ("tecs_sim_nearest_noalloc", C.c_int, [_P, _I32, _F32P, _I32, _P]),
    ("tecs_sim_random_in_bins", C.c_int, [_P, _I32, _F32P, _I32, _F32P, _P]),
    ("tecs_sim_random_neighbour", C.c_int, [_P, _I32, _F32P, C.c_float, _F32P, _P]),
    ("tecs_sim_set_walls", C.c_int, [_P, _P, _I32]),
    ("tecs_sim_set_links", C.c_int, [_P, _P, C.c_int64, C.c_float]),
    ("tecs_sim_divide", C.c_int, [_P, C.c_float, C.c_float, _P, C.POINTER(C.c_uint32)]),
    ("tecs_sim_take_step", C.c_int, [_P, C.c_double, _I32]),
    ("tecs_sim_set_force", C.c_int, [_P, _P]),
    ("tecs_sim_evaluate_force", C.c_int, [_P, _I32, _F32P, _F32P]),
    ("tecs_sim_pairs", C.c_int, [_P, _I32, _P, _P, C.c_int64, _I64P]),
    ("tecs_sim_rebuild", C.c_int, [_P]),
    ("tecs_sim_synchronize", C.c_int, [_P]),
    ("tecs_sim_get_nonfinite", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("tecs_sim_get_stats", C.c_int, [_P, C.POINTER(abi.Stats)]),
    ("tecs_sim_reset_stats", C.c_int, [_P]),
    ("tecs_sim_set_sweep_variant", C.c_int, [_P, _I32]),
    ("tecs_sim_get_sweep_variant", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("tecs_sim_get_autotune", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("tecs_sim_set_graph_mode", C.c_int, [_P, _I32]),
    ("tecs_sim_get_graph_launches", C.c_int, [_P, _I64P]),
    ("tecs_sim_set_kernel_timing", C.c_int, [_P, _I32]),
    ("tecs_sim_get_kernel_timing", C.c_int, [_P, _I32, C.POINTER(C.c_double), _I64P]),
    ("tecs_batch_create", C.c_int, [C.POINTER(abi.SimDesc), _I32, _P, C.POINTER(_P)]),
    ("tecs_batch_destroy", C.c_int, [_P]),
    ("tecs_batch_upload", C.c_int, [_P, _I32, _F32P, _F32P, _P]),
    ("tecs_batch_download", C.c_int, [_P, _I32, _F32P, _F32P, _F32P]),
    ("tecs_batch_take_step", C.c_int, [_P, _P, _I32]),
    ("tecs_batch_set_force", C.c_int, [_P, _I32, _P]),
    ("tecs_batch_upload_range", C.c_int, [_P, _I32, C.c_uint32, C.c_uint32, _F32P, _F32P, _P]),
    ("tecs_batch_set_agent_count", C.c_int, [_P, _I32, C.c_uint32]),
    ("tecs_batch_set_walls", C.c_int, [_P, _I32, _P, _I32]),
    ("tecs_batch_set_links", C.c_int, [_P, _I32, _P, C.c_int64, C.c_float]),
    ("tecs_batch_divide", C.c_int, [_P, _I32, C.c_float, C.c_float, _P, C.POINTER(C.c_uint32)]),
    ("tecs_batch_set_capacity", C.c_int, [_P, _I32, C.c_uint32]),
    ("tecs_batch_synchronize", C.c_int, [_P]),
    ("tecs_batch_get_nonfinite", C.c_int, [_P, _P]),
    ("tecs_batch_get_stats", C.c_int, [_P, C.POINTER(abi.Stats)]),
    ("tecs_batch_reset_stats", C.c_int, [_P]),
    ("tecs_batch_set_sweep_variant", C.c_int, [_P, _I32]),
    ("tecs_batch_get_sweep_variant", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("tecs_batch_get_autotune", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("tecs_batch_get_size_classes", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("tecs_batch_set_size_class_variants", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("tecs_batch_set_graph_mode", C.c_int, [_P, _I32]),
    ("tecs_batch_get_graph_launches", C.c_int, [_P, _I64P]),
    ("tecs_batch_set_kernel_timing", C.c_int, [_P, _I32]),
    ("tecs_batch_get_kernel_timing", C.c_int, [_P, _I32, C.POINTER(C.c_double), _I64P]),
    ("tecs_writer_open", C.c_int, [C.c_char_p, _I32, C.POINTER(_P)]),
    ("tecs_writer_close", C.c_int, [_P]),
    ("tecs_writer_flush", C.c_int, [_P]),
    ("tecs_writer_write", C.c_int, [_P, C.c_double, C.c_uint32, C.c_uint32, _F32P, _F32P]),
    ("tecs_writer_write_static", C.c_int, [_P, C.c_char_p, C.c_uint32, _P]),
    ("tecs_sim_write", C.c_int, [_P, _P, C.c_double]),
    ("tecs_batch_write", C.c_int, [_P, _I32, _P, C.c_double]),
    ("tecs_read_dataset", C.c_int, [C.c_char_p, C.c_char_p, _F32P, C.c_int64, _I64P]),
    ("tecs_read_dataset_typed", C.c_int, [C.c_char_p, C.c_char_p, _I32, _P, C.c_int64, _I64P]),
    ("tecs_server_create", C.c_int, [C.c_uint64, _I32, C.POINTER(_P)]),
    ("tecs_server_destroy", C.c_int, [_P]),
    ("tecs_server_submit", C.c_int, [_P, C.POINTER(abi.SimDesc), _F32P, _F32P, _P, C.c_double, _I32, _I64P]),
    ("tecs_server_advance", C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("tecs_server_request_info", C.c_int, [_P, C.c_int64, C.POINTER(abi.RequestInfo)]),
    ("tecs_server_cancel", C.c_int, [_P, C.c_int64]),
    ("tecs_server_result", C.c_int, [_P, C.c_int64, _F32P, _F32P]),
    ("tecs_server_get_stats", C.c_int, [_P, C.POINTER(abi.ServerStats)]),
    ("tecs_copy_probe", C.c_int, [C.c_int64, _I32, C.POINTER(C.c_double)]),
]


class TecsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tecs error {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load the engine library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the gfx950 engine first (python -m tecs_amd._build)"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.tecs_abi_version() != abi.ABI_VERSION:
        raise RuntimeError("engine library ABI version mismatch")
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != abi.OK:
        raise TecsError(rc, load().tecs_last_error().decode(errors="replace"))


def device_count() -> int:
    n = C.c_int32(0)
    check(load().tecs_device_count(C.byref(n)))
    return int(n.value)
