"""Typed slot storage + row-driven dispatch: the host mirror of aristotle::World and
of the reference Accelerator's per-kind API over the C ABI (include/exgame.h).

`Store.insert(type, value)` returns a `Key(index, generation)` exactly as
`WriteView::add` does; `Accelerator` methods take lists of ROW-slot keys, one
round at a time, like `Accelerator::submit_rounds`
(crates/newton/src/accelerator/mod.rs:557-965).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

SCALAR, MOTOR, TWIST, WRENCH, WRENCH_PAIR, JACOBIAN, BLOCK, DIAG_INERTIA, FULL_INERTIA, ROW = range(10)
WIDTH = {SCALAR: 1, MOTOR: 8, TWIST: 6, WRENCH: 6, WRENCH_PAIR: 12, JACOBIAN: 288, BLOCK: 36, DIAG_INERTIA: 3,
         FULL_INERTIA: 9, ROW: 128}

# MessageKind numbering (crates/newton/src/accelerator/mod.rs:55-98)
KIND = {
    "EvalSimpleSum": 0,
    "EvalCriticallyDampedWarpedPlain": 1, "EvalCriticallyDampedWarpedJacobian": 2,
    "EvalPerpendicularDamperWarpedPlain": 3, "EvalPerpendicularDamperWarpedJacobian": 4,
    "EvalSimpleSpringDamperPlain": 5, "EvalSimpleSpringDamperJacobian": 6,
    "EvalTorsionalDamperWarpedPlain": 7, "EvalTorsionalDamperWarpedJacobian": 8,
    "Pre": 9, "Gather": 10, "Gemm": 11, "BodyPostDiagonal": 12, "BodyPostFull": 13, "AssembleBlock": 14,
    "BlockMatVec": 15, "BlockReduce": 16, "BlockCopy": 17, "BodyPostGatheredDiagonal": 18,
    "BodyPostGatheredFull": 19,
}
JOINT_SHADER = {0: "SimpleSpringDamper", 1: "CriticallyDampedWarped", 2: "PerpendicularDamperWarped",
                3: "TorsionalDamperWarped"}

# storage scalar (World::builder().usual::<T>())
SCALARS = {"f32": 0, "f64": 1}
# kernels of the joint and Pre rows
CODEGEN = {"handwritten": 0, "traced": 1}

_u32p = C.POINTER(C.c_uint32)
_SIGS = {
    "exg_store_create": [C.c_uint64, C.POINTER(C.c_void_p)],
    "exg_store_create_scalar": [C.c_uint64, C.c_int, C.POINTER(C.c_void_p)],
    "exg_store_scalar": [C.c_void_p, C.POINTER(C.c_int)],
    "exg_store_set_codegen": [C.c_void_p, C.c_int],
    "exg_store_destroy": [C.c_void_p],
    "exg_slot_insert": [C.c_void_p, C.c_int, C.c_void_p, _u32p, _u32p],
    "exg_slot_read": [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(C.c_int)],
    "exg_slot_write": [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(C.c_int)],
    "exg_slot_remove": [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_void_p, C.POINTER(C.c_int)],
    "exg_storage_info": [C.c_void_p, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "exg_storage_device_ptr": [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)],
    "exg_dispatch": [C.c_void_p, C.c_int, _u32p, C.c_uint32],
    "exg_flush": [C.c_void_p],
    "exg_fatals": [C.c_void_p, C.c_int, _u32p],
}


def _bind():
    L = _lib.lib()
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    return L


@dataclass(frozen=True)
class Key:
    index: int
    generation: int


def f32_bits(x: float) -> int:
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


class Store:
    """Slots hold the reference's Pod layouts over the storage scalar: screws as their
    grade-2 store (see exgame.h), motors as 8 even blades, blocks as [[T; 6]; 6]."""

    def __init__(self, capacity: int = 256 * 1024, scalar: str = "f64"):
        self.L = _bind()
        h = C.c_void_p()
        _lib.check(self.L.exg_store_create_scalar(int(capacity), SCALARS[scalar], C.byref(h)))
        self.h = h
        self.dtype = np.float32 if scalar == "f32" else np.float64

    def _buf(self, type_: int, value=None):
        if type_ == ROW:
            a = np.zeros(128, dtype=np.uint32)
            if value is not None:
                v = np.asarray(value, dtype=np.uint32).reshape(-1)
                a[: v.size] = v
        else:
            a = np.zeros(WIDTH[type_], dtype=self.dtype)
            if value is not None:
                a[:] = np.asarray(value, dtype=np.float64).reshape(-1)
        return a

    def insert(self, type_: int, value) -> Key:
        a = self._buf(type_, value)
        i, g = C.c_uint32(), C.c_uint32()
        _lib.check(self.L.exg_slot_insert(self.h, type_, a.ctypes.data_as(C.c_void_p), C.byref(i), C.byref(g)))
        return Key(i.value, g.value)

    def read(self, type_: int, key: Key):
        a = self._buf(type_)
        found = C.c_int()
        _lib.check(self.L.exg_slot_read(self.h, type_, key.index, key.generation, a.ctypes.data_as(C.c_void_p),
                                        C.byref(found)))
        return a if found.value else None

    def write(self, type_: int, key: Key, value) -> bool:
        a = self._buf(type_, value)
        found = C.c_int()
        _lib.check(self.L.exg_slot_write(self.h, type_, key.index, key.generation, a.ctypes.data_as(C.c_void_p),
                                         C.byref(found)))
        return bool(found.value)

    def remove(self, type_: int, key: Key):
        a = self._buf(type_)
        found = C.c_int()
        _lib.check(self.L.exg_slot_remove(self.h, type_, key.index, key.generation, a.ctypes.data_as(C.c_void_p),
                                          C.byref(found)))
        return a if found.value else None

    def info(self, type_: int) -> dict:
        cap, ln, gen = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _lib.check(self.L.exg_storage_info(self.h, type_, C.byref(cap), C.byref(ln), C.byref(gen)))
        return {"capacity": cap.value, "len": ln.value, "generation": gen.value}

    def dispatch(self, kind: int, rows) -> None:
        slots = np.ascontiguousarray([r.index for r in rows], dtype=np.uint32)
        _lib.check(self.L.exg_dispatch(self.h, int(kind), slots.ctypes.data_as(_u32p), len(slots)))

    def set_codegen(self, mode: str) -> None:
        """Joint and Pre rows on the hand-written kernels or on the device functions
        traced from the same laws at build time (exgame_amd/viete)."""
        _lib.check(self.L.exg_store_set_codegen(self.h, CODEGEN[mode]))

    def fatals(self, kind: int) -> int:
        """Rows of the last dispatch of `kind` with non-finite outputs."""
        out = C.c_uint32()
        _lib.check(self.L.exg_fatals(self.h, int(kind), C.byref(out)))
        return out.value

    def flush(self) -> None:
        _lib.check(self.L.exg_flush(self.h))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.exg_store_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Accelerator:
    """Per-kind entry points of newton::Accelerator (accelerator/mod.rs:640-873) over a Store."""

    def __init__(self, store: Store):
        self.store = store

    def pre(self, rows):
        self.store.dispatch(KIND["Pre"], rows)

    def eval_joints(self, joint_kind: int, jacobian: bool, rows):
        name = "Eval" + JOINT_SHADER[joint_kind] + ("Jacobian" if jacobian else "Plain")
        self.store.dispatch(KIND[name], rows)

    def gather(self, rows):
        self.store.dispatch(KIND["Gather"], rows)

    def gemm(self, rounds):
        for r in rounds:
            self.store.dispatch(KIND["Gemm"], r)

    def body_post(self, rows, diagonal: bool):
        self.store.dispatch(KIND["BodyPostDiagonal" if diagonal else "BodyPostFull"], rows)

    def body_post_gathered(self, rows, diagonal: bool):
        self.store.dispatch(KIND["BodyPostGatheredDiagonal" if diagonal else "BodyPostGatheredFull"], rows)

    def assemble(self, rounds):
        for r in rounds:
            self.store.dispatch(KIND["AssembleBlock"], r)

    def block_matvec(self, rounds):
        for r in rounds:
            self.store.dispatch(KIND["BlockMatVec"], r)

    def block_reduce(self, rows):
        self.store.dispatch(KIND["BlockReduce"], rows)

    def block_copy(self, rows):
        self.store.dispatch(KIND["BlockCopy"], rows)

    def simple_sum(self, rows):
        self.store.dispatch(KIND["EvalSimpleSum"], rows)
