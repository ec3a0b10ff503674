"""Many independent worlds stepped together on one device (the batched and served
forms): a host handle over exg_arena_* of the C ABI (include/exgame.h).

Worlds are admitted into fixed-capacity device arrays and retired to free them; a
step advances every admitted world. Each world's trajectory is bit-identical alone
or in any mix of other worlds."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, pga
from .world import NS_ENGINES, PRECISIONS

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_SIGS = {
    "exg_arena_create": [C.c_int, C.c_int, C.c_int, C.c_longlong, C.POINTER(C.c_void_p)],
    "exg_arena_destroy": [C.c_void_p],
    "exg_arena_set_solver": [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int],
    "exg_arena_set_gravity_chunk": [C.c_void_p, C.c_int],
    "exg_arena_set_graph": [C.c_void_p, C.c_int],
    "exg_arena_autotune_gravity": [C.c_void_p, C.POINTER(C.c_int)],
    "exg_arena_set_joint_mapping": [C.c_void_p, C.c_int],
    "exg_arena_autotune_joints": [C.c_void_p, C.POINTER(C.c_int)],
    "exg_arena_set_precision": [C.c_void_p, C.c_int, C.c_int],
    "exg_arena_admit": [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp, C.c_int, _ip, _ip, _ip, _dp,
                        C.c_double, C.c_int, _ip],
    "exg_arena_retire": [C.c_void_p, C.c_int],
    "exg_arena_step": [C.c_void_p, C.c_int, C.c_double, C.c_double],
    "exg_arena_sync": [C.c_void_p],
    "exg_arena_mark": [C.c_void_p, C.POINTER(C.c_uint64)],
    "exg_arena_mark_done": [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)],
    "exg_arena_mark_wait": [C.c_void_p, C.c_uint64],
    "exg_arena_download": [C.c_void_p, C.c_int, _dp, _dp, _dp],
    "exg_arena_world_stats": [C.c_void_p, C.c_int, C.POINTER(C.c_int32)],
    "exg_arena_last_ms": [C.c_void_p, _dp],
    "exg_arena_usage": [C.c_void_p, _i64p, _i64p],
    "exg_arena_footprint": [C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_int, _i64p],
}


def _bind():
    L = _lib.lib()
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    return L


def inverse_doubles(scene) -> int:
    """Doubles of accepted inverses a scene needs: Σ (6m)² over its islands."""
    m = np.diff(np.asarray(scene.island_start, dtype=np.int64))
    return int(np.sum(36 * m * m))


def footprint(scene, gravity_chunk: int = 0) -> int:
    """HBM bytes the scene occupies in an arena (gravity chunk setting 0: the
    single-world rule)."""
    out = C.c_int64()
    _lib.check(_bind().exg_arena_footprint(scene.n_bodies, scene.n_islands, scene.n_joints, inverse_doubles(scene),
                                           int(gravity_chunk), C.byref(out)))
    return out.value


class Arena:
    def __init__(self, bodies: int, islands: int, joints: int, inverses: int, tol2: float = 1e-20,
                 ns_warm: int = 2, ns_seed: int = 24, ns_engine: str = "mfma", gravity_precision: str = "f64",
                 gravity_chunk: int = 0, jacobian_precision: str = "f64", graph: bool = False):
        self.L = _bind()
        h = C.c_void_p()
        _lib.check(self.L.exg_arena_create(int(bodies), int(islands), int(joints), int(inverses), C.byref(h)))
        self.h = h
        self.gravity_chunk = int(gravity_chunk)
        _lib.check(self.L.exg_arena_set_solver(self.h, float(tol2), int(ns_warm), int(ns_seed), NS_ENGINES[ns_engine],
                                               PRECISIONS[gravity_precision]))
        _lib.check(self.L.exg_arena_set_gravity_chunk(self.h, self.gravity_chunk))
        _lib.check(self.L.exg_arena_set_precision(self.h, PRECISIONS[gravity_precision],
                                                  PRECISIONS[jacobian_precision]))
        _lib.check(self.L.exg_arena_set_graph(self.h, int(bool(graph))))
        self.shapes = {}

    def set_gravity_chunk(self, chunk: int) -> None:
        """Sources per gravity chunk of worlds admitted from now on (0: the chunking a
        stand-alone world of the same size uses); admitted worlds keep theirs."""
        _lib.check(self.L.exg_arena_set_gravity_chunk(self.h, int(chunk)))
        self.gravity_chunk = int(chunk)

    def autotune_gravity(self) -> int:
        """Time the gravity pass of the live worlds under the chunk settings 0 (the
        single-world rule) and 128 … 8192, and keep the fastest for worlds admitted from
        now on (admitted worlds keep their chunking, so their results do not change);
        returns it."""
        out = C.c_int(0)
        _lib.check(self.L.exg_arena_autotune_gravity(self.h, C.byref(out)))
        self.gravity_chunk = out.value
        return out.value

    def set_joint_mapping(self, mapping: str) -> None:
        """Joint mapping of the 8-body islands of worlds admitted from now on."""
        from .world import JOINT_MAPPINGS
        _lib.check(self.L.exg_arena_set_joint_mapping(self.h, JOINT_MAPPINGS[mapping]))

    def autotune_joints(self) -> str:
        """Time one step of the live worlds with all of them on each joint mapping (from a
        saved state that is put back; the live worlds keep their mappings) and make the
        fastest the mapping of later admissions; returns its name."""
        from .world import JOINT_MAPPINGS
        out = C.c_int(0)
        _lib.check(self.L.exg_arena_autotune_joints(self.h, C.byref(out)))
        return {v: k for k, v in JOINT_MAPPINGS.items()}[out.value]

    def admit(self, scene, G: float = 1.0, gravity: bool = True):
        """World id, or None when the arena has no room for it now."""
        from .world import kinematic_last_order
        s = scene
        ist = np.ascontiguousarray(s.island_start, dtype=np.int32)
        order, back = kinematic_last_order(s.mass, ist)   # kinematic bodies last in each island
        body = (lambda x: x) if order is None else (lambda x: np.asarray(x)[order])
        ends = (lambda x: x) if order is None else (lambda x: back[np.asarray(x)])
        pose = np.ascontiguousarray(body(s.pose), dtype=np.float64)
        mom = np.ascontiguousarray(body(s.momentum), dtype=np.float64)
        mass = np.ascontiguousarray(body(s.mass), dtype=np.float64)
        tens = np.ascontiguousarray(body(s.inertia), dtype=np.float64)
        size = np.ascontiguousarray(body(s.size), dtype=np.float64)
        origin = np.ascontiguousarray(s.origin, dtype=np.float64)
        if s.n_joints:
            jk = np.ascontiguousarray(s.joint_kind, dtype=np.int32)
            ja = np.ascontiguousarray(ends(s.joint_a), dtype=np.int32)
            jb = np.ascontiguousarray(ends(s.joint_b), dtype=np.int32)
            jp = np.ascontiguousarray(s.joint_params, dtype=np.float64).reshape(-1)
        else:
            jk = ja = jb = np.zeros(1, dtype=np.int32)
            jp = np.zeros(10)
        wid = C.c_int(-1)
        rc = self.L.exg_arena_admit(self.h, s.n_bodies, _lib.fptr(pose), _lib.fptr(mom), _lib.fptr(mass),
                                    _lib.fptr(tens), _lib.fptr(size), s.n_islands, _lib.iptr(ist), _lib.fptr(origin),
                                    s.n_joints, _lib.iptr(jk), _lib.iptr(ja), _lib.iptr(jb), _lib.fptr(jp), float(G),
                                    int(bool(gravity)), C.byref(wid))
        if rc == 8:
            return None
        _lib.check(rc)
        self.shapes[wid.value] = (s.n_bodies, s.n_islands, np.diff(ist).astype(np.int64), back)
        return wid.value

    def retire(self, wid: int) -> None:
        _lib.check(self.L.exg_arena_retire(self.h, int(wid)))
        self.shapes.pop(wid, None)

    def step(self, n: int = 1, dt: float = 1e-2, warp: float = 1.0) -> None:
        _lib.check(self.L.exg_arena_step(self.h, int(n), float(dt), float(warp)))

    def sync(self) -> None:
        _lib.check(self.L.exg_arena_sync(self.h))

    def mark(self) -> int:
        """A completion mark behind the work enqueued so far; returns its ticket."""
        t = C.c_uint64()
        _lib.check(self.L.exg_arena_mark(self.h, C.byref(t)))
        return t.value

    def mark_done(self, ticket: int) -> bool:
        d = C.c_int()
        _lib.check(self.L.exg_arena_mark_done(self.h, int(ticket), C.byref(d)))
        return bool(d.value)

    def mark_wait(self, ticket: int) -> None:
        _lib.check(self.L.exg_arena_mark_wait(self.h, int(ticket)))

    def last_ms(self) -> float:
        v = C.c_double()
        _lib.check(self.L.exg_arena_last_ms(self.h, C.byref(v)))
        return v.value

    def state(self, wid: int) -> dict:
        nb, ni, sizes, back = self.shapes[wid]
        pose = np.zeros((nb, 8))
        mom = np.zeros((nb, 6))
        origin = np.zeros((ni, 3))
        _lib.check(self.L.exg_arena_download(self.h, int(wid), _lib.fptr(pose), _lib.fptr(mom), _lib.fptr(origin)))
        if back is not None:   # back to the scene's body order
            pose, mom = pose[back], mom[back]
        return {"pose": pose, "momentum": mom, "origin": origin,
                "position": pga.position(pose) + np.repeat(origin, sizes, axis=0)}

    def stats(self, wid: int) -> np.ndarray:
        nb, ni, _, _ = self.shapes[wid]
        out = np.zeros((ni, 8), dtype=np.int32)
        _lib.check(self.L.exg_arena_world_stats(self.h, int(wid), out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def usage(self):
        used = (C.c_int64 * 4)()
        cap = (C.c_int64 * 4)()
        _lib.check(self.L.exg_arena_usage(self.h, used, cap))
        return list(used), list(cap)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.exg_arena_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
