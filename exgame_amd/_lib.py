"""ctypes binding of libexgame_hip.so (the C ABI declared in include/exgame.h).

The product path has no fallback: if the HIP library is not built, or no GPU is
visible, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libexgame_hip.so")

EXG_OK = 0
STATUS = {
    1: "NoShader", 2: "Trap", 3: "Fatal", 4: "WorkerGone", 5: "Backend", 6: "Invalid", 7: "Unsupported", 8: "Full",
}


class ExgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i32p = C.POINTER(C.c_int32)


class DeviceInfo(C.Structure):
    _fields_ = [("compute_units", C.c_int), ("wave_size", C.c_int), ("lds_bytes_per_cu", C.c_int64),
                ("hbm_bytes", C.c_int64), ("name", C.c_char * 64)]


# name → (argtypes); every function returns int status
SIGNATURES = {
    "exg_device_info_get": [C.POINTER(DeviceInfo)],
    "exg_world_create": [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp, C.c_int, _ip, _ip, _ip, _dp, C.c_double,
                         C.c_int, C.POINTER(C.c_void_p)],
    "exg_world_destroy": [C.c_void_p],
    "exg_world_set_solver": [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int],
    "exg_world_set_ns_engine": [C.c_void_p, C.c_int],
    "exg_world_set_joint_mapping": [C.c_void_p, C.c_int],
    "exg_world_set_graph": [C.c_void_p, C.c_int],
    "exg_world_step": [C.c_void_p, C.c_int, C.c_double, C.c_double],
    "exg_world_sync": [C.c_void_p],
    "exg_world_fatals": [C.c_void_p, C.POINTER(C.c_int64)],
    "exg_world_download": [C.c_void_p, _dp, _dp, _dp],
    "exg_world_charges": [C.c_void_p, _dp],
    "exg_world_external": [C.c_void_p, _dp],
    "exg_world_stats": [C.c_void_p, _i32p],
    "exg_world_last_timing": [C.c_void_p, _dp],
    "exg_gravity_accumulate": [C.c_int, _dp, C.c_double, _dp],
    "exg_gravity_accumulate_precision": [C.c_int, _dp, C.c_double, C.c_int, _dp],
    "exg_world_set_precision": [C.c_void_p, C.c_int, C.c_int],
    "exg_world_set_gravity_island_reduce": [C.c_void_p, C.c_int],
    "exg_world_set_uniform_field": [C.c_void_p, C.c_double, C.c_double, C.c_double],
}

_lib = None


def lib():
    """Load the in-tree library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.exg_last_error.restype = C.c_char_p
    L.exg_last_error.argtypes = []
    _lib = L
    return L


def check(code: int) -> None:
    if code != EXG_OK:
        raise ExgError(code, lib().exg_last_error().decode(errors="replace"))


def fptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def device_info() -> dict:
    d = DeviceInfo()
    check(lib().exg_device_info_get(C.byref(d)))
    return {"compute_units": d.compute_units, "wave_size": d.wave_size, "lds_bytes_per_cu": d.lds_bytes_per_cu,
            "hbm_bytes": d.hbm_bytes, "name": d.name.decode()}
