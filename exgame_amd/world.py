"""Device-resident world: the host-side mirror of the reference's stepping API.

`DeviceWorld.step(n, dt, warp)` is one `Mechanism::step` per step over every island
followed by the gravity propagator's `advance_epoch`
(crates/newton/src/mechanism/mechanism.rs:526-600, crates/newton/src/gravity.rs:392-397),
with the state resident in HBM between calls. Scenes come from exgame_amd.scenes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, pga

INTEGRATORS = {"explicit_euler": 0, "symplectic_euler": 1, "newton": 2, "lie_euler": 3}
# Newton–Schulz products: FP64 MFMA tiles, or scalar FMAs in the reference's block order
NS_ENGINES = {"mfma": 0, "valu_ordered": 1}
# f64 throughout, or FP32 arithmetic with two-sum compensated accumulation
PRECISIONS = {"f64": 0, "f32c": 1}


def kinematic_last_order(mass, island_start):
    """The library takes an island's kinematic bodies (mass 0) after its dynamic ones.
    Returns (order, back): order[p] is the scene index of the body uploaded at p (each
    island's dynamic bodies in their order, then its kinematic ones), back its inverse;
    (None, None) when the scene already has that order. Dynamic bodies keep their
    relative order, so the gravity pass's higher-slot rule between them is unchanged
    (pairs with a kinematic body carry no force)."""
    kin = np.asarray(mass) == 0.0
    st = np.asarray(island_start)
    order = np.concatenate([np.concatenate([np.flatnonzero(~kin[a:b]) + a, np.flatnonzero(kin[a:b]) + a])
                            for a, b in zip(st[:-1], st[1:])]).astype(np.int64)
    if np.array_equal(order, np.arange(len(kin))):
        return None, None
    back = np.empty_like(order)
    back[order] = np.arange(len(kin))
    return order, back


@dataclass(frozen=True)
class Epoch:
    """dt and time warp of one step (crates/aristotle/src/epoch.rs:238-248)."""
    dt: float
    warp: float = 1.0


# how an 8-body island's joint work is spread over its waves (bit-identical results)
JOINT_MAPPINGS = {"split": 0, "flat": 1}


class DeviceWorld:
    def __init__(self, scene, G: float = 1.0, gravity: bool = True, integrator: str = "newton", tol2: float = 1e-8,
                 ns_warm: int = 2, ns_seed: int = 24, ns_engine: str = "mfma", gravity_precision: str = "f64",
                 jacobian_precision: str = "f64", graph: bool = False, joint_mapping: str = "split",
                 gravity_island_reduce: bool = True, uniform_field=None):
        L = _lib.lib()
        s = scene
        self.nb, self.ni, self.nj = s.n_bodies, s.n_islands, s.n_joints
        self.island_start = np.ascontiguousarray(s.island_start, dtype=np.int32)
        self._order, self._back = kinematic_last_order(s.mass, self.island_start)
        body = (lambda x: x) if self._order is None else (lambda x: np.asarray(x)[self._order])
        ends = (lambda x: x) if self._order is None else (lambda x: self._back[np.asarray(x)])
        arrs = dict(
            pose=np.ascontiguousarray(body(s.pose), dtype=np.float64),
            mom=np.ascontiguousarray(body(s.momentum), dtype=np.float64),
            mass=np.ascontiguousarray(body(s.mass), dtype=np.float64),
            tens=np.ascontiguousarray(body(s.inertia), dtype=np.float64),
            size=np.ascontiguousarray(body(s.size), dtype=np.float64),
            origin=np.ascontiguousarray(s.origin, dtype=np.float64),
            jp=np.ascontiguousarray(s.joint_params, dtype=np.float64).reshape(-1),
        )
        ints = dict(
            jk=np.ascontiguousarray(s.joint_kind, dtype=np.int32),
            ja=np.ascontiguousarray(ends(s.joint_a), dtype=np.int32),
            jb=np.ascontiguousarray(ends(s.joint_b), dtype=np.int32),
        )
        if self.nj == 0:
            arrs["jp"] = np.zeros(10)
            for k in ints:
                ints[k] = np.zeros(1, dtype=np.int32)
        h = C.c_void_p()
        _lib.check(L.exg_world_create(
            self.nb, _lib.fptr(arrs["pose"]), _lib.fptr(arrs["mom"]), _lib.fptr(arrs["mass"]), _lib.fptr(arrs["tens"]),
            _lib.fptr(arrs["size"]), self.ni, _lib.iptr(self.island_start), _lib.fptr(arrs["origin"]), self.nj,
            _lib.iptr(ints["jk"]), _lib.iptr(ints["ja"]), _lib.iptr(ints["jb"]), _lib.fptr(arrs["jp"]), float(G),
            int(bool(gravity)), C.byref(h)))
        self.h = h
        _lib.check(L.exg_world_set_solver(self.h, INTEGRATORS[integrator], float(tol2), int(ns_warm), int(ns_seed)))
        _lib.check(L.exg_world_set_ns_engine(self.h, NS_ENGINES[ns_engine]))
        _lib.check(L.exg_world_set_precision(self.h, PRECISIONS[gravity_precision], PRECISIONS[jacobian_precision]))
        _lib.check(L.exg_world_set_graph(self.h, int(bool(graph))))
        _lib.check(L.exg_world_set_joint_mapping(self.h, JOINT_MAPPINGS[joint_mapping]))
        if not gravity_island_reduce:
            _lib.check(L.exg_world_set_gravity_island_reduce(self.h, 0))
        if uniform_field is not None:   # UniformField: F = m·a on every body (field.rs:16-44)
            ax, ay, az = (float(v) for v in uniform_field)
            _lib.check(L.exg_world_set_uniform_field(self.h, ax, ay, az))

    def step(self, n: int = 1, dt: float = 1e-2, warp: float = 1.0) -> None:
        if isinstance(dt, Epoch):
            dt, warp = dt.dt, dt.warp
        _lib.check(_lib.lib().exg_world_step(self.h, int(n), float(dt), float(warp)))

    def sync(self) -> None:
        _lib.check(_lib.lib().exg_world_sync(self.h))

    def state(self) -> dict:
        pose = np.zeros((self.nb, 8))
        mom = np.zeros((self.nb, 6))
        origin = np.zeros((self.ni, 3))
        _lib.check(_lib.lib().exg_world_download(self.h, _lib.fptr(pose), _lib.fptr(mom), _lib.fptr(origin)))
        pose, mom = self._scene_order(pose), self._scene_order(mom)
        isl = np.repeat(np.arange(self.ni), np.diff(self.island_start))
        position = pga.position(pose) + origin[isl]
        return {"pose": pose, "momentum": mom, "origin": origin, "position": position}

    def charges(self) -> np.ndarray:
        out = np.zeros((self.nb, 8))
        _lib.check(_lib.lib().exg_world_charges(self.h, _lib.fptr(out)))
        return self._scene_order(out)

    def _scene_order(self, x: np.ndarray) -> np.ndarray:
        """Per-body rows from the device's order back to the scene's."""
        return x if self._back is None else x[self._back]

    def external(self) -> np.ndarray:
        out = np.zeros((self.nb, 6))
        _lib.check(_lib.lib().exg_world_external(self.h, _lib.fptr(out)))
        return self._scene_order(out)

    def stats(self) -> np.ndarray:
        out = np.zeros((self.ni, 8), dtype=np.int32)
        _lib.check(_lib.lib().exg_world_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def fatals(self) -> int:
        """Island-steps since creation that ended with a non-finite state."""
        out = C.c_int64()
        _lib.check(_lib.lib().exg_world_fatals(self.h, C.byref(out)))
        return out.value

    def last_timing(self) -> dict:
        out = np.zeros(2)
        _lib.check(_lib.lib().exg_world_last_timing(self.h, _lib.fptr(out)))
        return {"gravity_ms": float(out[0]), "island_ms": float(out[1])}

    def close(self) -> None:
        if getattr(self, "h", None):
            _lib.lib().exg_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gravity_accumulate(charges: np.ndarray, G: float = 1.0, precision: str = "f64") -> np.ndarray:
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    n = ch.size // 8
    out = np.zeros((n, 6))
    _lib.check(_lib.lib().exg_gravity_accumulate_precision(n, _lib.fptr(ch), float(G), PRECISIONS[precision],
                                                           _lib.fptr(out)))
    return out
