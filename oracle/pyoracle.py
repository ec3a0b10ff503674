"""ORACLE — test infrastructure only.

ctypes bindings for the CPU f64 restatement of the reference's per-step
dynamics (built from oracle/*.hpp into oracle/_build/liborc.so). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the
product (exgame_amd) never does.

Conventions at this boundary:
  * screws (twists, wrenches) are 6 doubles: linear|force xyz, angular|torque xyz
    (the coordinate order of Twist::new / Wrench::new, crates/clifford/src/pga3/screw.rs:235-290);
  * motors are 8 doubles: the even blades {1, e12, e13, e23, e01, e02, e03, e0123}
    in ascending blade order (crates/clifford/src/pga3/motor.rs:119-152).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liborc.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


NATIVE_LIB_PATH = os.path.join(HERE, "_build", "liborc_native.so")


def build() -> None:
    """Compile the oracle with its own Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def use_native() -> None:
    """Compile the oracle for this host's CPU (-march=native: the CPU-baseline timing
    build; always rebuilt, a copy made on another machine may not run here) and load
    it. Call before any other use."""
    global LIB_PATH
    if _lib is not None:
        raise RuntimeError("use_native() must come before the oracle is loaded")
    subprocess.run(["make", "-s", "-B", "-C", HERE, "native"], check=True)
    LIB_PATH = NATIVE_LIB_PATH


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if n is not None:
        assert a.size == n, (a.shape, n)
    return a


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    sig = {
        "orc_version": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
        "orc_max_threads": (C.c_int, []),
        "orc_sinc_sq": (C.c_double, [C.c_double]),
        "orc_cos_sq": (C.c_double, [C.c_double]),
        "orc_dsinc_sq": (C.c_double, [C.c_double]),
        "orc_blade_prod": (None, [C.c_int, C.c_int, _ip, _ip]),
        "orc_complement_sign": (C.c_int, [C.c_int, C.c_int]),
        "orc_angular_embed_sign": (C.c_int, []),
        "orc_motor_exp_twist": (None, [_dp, C.c_double, _dp]),
        "orc_motor_exp_raw": (None, [_dp, _dp]),
        "orc_motor_compose": (None, [_dp, _dp, _dp]),
        "orc_motor_inverse": (None, [_dp, _dp]),
        "orc_motor_apply_point": (None, [_dp, _dp, _dp]),
        "orc_motor_point_raw": (None, [_dp, _dp, _dp]),
        "orc_twist_transport": (None, [_dp, _dp, _dp]),
        "orc_wrench_transport": (None, [_dp, _dp, _dp]),
        "orc_power": (C.c_double, [_dp, _dp]),
        "orc_velocity_at": (None, [_dp, _dp, _dp]),
        "orc_twist_bracket": (None, [_dp, _dp, _dp]),
        "orc_coad_bracket": (None, [_dp, _dp, _dp]),
        "orc_join_points": (None, [_dp, _dp, _dp, _dp, _dp]),
        "orc_wrench_from_points": (None, [_dp, _dp, _dp]),
        "orc_screw_raw": (None, [_dp, _dp]),
        "orc_joint_plain": (None, [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp]),
        "orc_joint_jacobian": (None, [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp]),
        "orc_pre": (None, [_dp, _dp, C.c_double, _dp]),
        "orc_body_post": (None, [_dp, _dp, C.c_double, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp]),
        "orc_pair_force": (C.c_int, [_dp, _dp, _dp]),
        "orc_gravity_accumulate": (None, [C.c_int, _dp, C.c_double, _dp]),
        "orc_world_new": (C.c_void_p, [C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _ip, _dp, C.c_int, _ip, _ip, _ip,
                                       _dp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]),
        "orc_world_free": (None, [C.c_void_p]),
        "orc_world_step": (None, [C.c_void_p, C.c_int, C.c_double, C.c_double]),
        "orc_world_get": (None, [C.c_void_p, _dp, _dp, _dp, _dp]),
        "orc_world_front": (None, [C.c_void_p, _dp]),
        "orc_world_stats": (None, [C.c_void_p, _i64p]),
        "orc_world_span_trace": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.c_int]),
        "orc_screw_to_pod": (None, [_dp, _dp]),
        "orc_screw_from_pod": (None, [_dp, _dp]),
        "orc_ns_solve": (None, [C.c_int, _dp, _dp, C.c_int, _dp, _dp, _dp]),
        "orc_world_set_uniform": (None, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
        "orc_rawmap_new": (C.c_void_p, [C.c_uint64]),
        "orc_rawmap_free": (None, [C.c_void_p]),
        "orc_rawmap_insert": (None, [C.c_void_p, C.c_uint64, _u32p, _u32p]),
        "orc_rawmap_remove": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _u64p]),
        "orc_rawmap_get": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _u64p]),
        "orc_rawmap_info": (None, [C.c_void_p, _u64p, _u64p, _u64p, _u32p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# ───────────── scalar / PGA helpers ─────────────
def sinc_sq(u):
    return lib().orc_sinc_sq(float(u))


def cos_sq(u):
    return lib().orc_cos_sq(float(u))


def dsinc_sq(u):
    return lib().orc_dsinc_sq(float(u))


def blade_prod(a, b):
    s, m = C.c_int(), C.c_int()
    lib().orc_blade_prod(a, b, C.byref(s), C.byref(m))
    return s.value, m.value


def twist_exp(twist6, dt):
    out = np.zeros(8)
    lib().orc_motor_exp_twist(_p(_f64(twist6, 6)), float(dt), _p(out))
    return out


def motor_exp_raw(twist6):
    out = np.zeros(8)
    lib().orc_motor_exp_raw(_p(_f64(twist6, 6)), _p(out))
    return out


def compose(self8, rhs8):
    out = np.zeros(8)
    lib().orc_motor_compose(_p(_f64(self8, 8)), _p(_f64(rhs8, 8)), _p(out))
    return out


def motor_inverse(m8):
    out = np.zeros(8)
    lib().orc_motor_inverse(_p(_f64(m8, 8)), _p(out))
    return out


def apply_point(m8, p3):
    out = np.zeros(3)
    lib().orc_motor_apply_point(_p(_f64(m8, 8)), _p(_f64(p3, 3)), _p(out))
    return out


def point_raw(m8, p3):
    out = np.zeros(16)
    lib().orc_motor_point_raw(_p(_f64(m8, 8)), _p(_f64(p3, 3)), _p(out))
    return out


def twist_transport(t6, m8):
    out = np.zeros(6)
    lib().orc_twist_transport(_p(_f64(t6, 6)), _p(_f64(m8, 8)), _p(out))
    return out


def wrench_transport(w6, m8):
    out = np.zeros(6)
    lib().orc_wrench_transport(_p(_f64(w6, 6)), _p(_f64(m8, 8)), _p(out))
    return out


def power(w6, t6):
    return lib().orc_power(_p(_f64(w6, 6)), _p(_f64(t6, 6)))


def velocity_at(t6, p3):
    out = np.zeros(3)
    lib().orc_velocity_at(_p(_f64(t6, 6)), _p(_f64(p3, 3)), _p(out))
    return out


def twist_bracket(a6, b6):
    out = np.zeros(6)
    lib().orc_twist_bracket(_p(_f64(a6, 6)), _p(_f64(b6, 6)), _p(out))
    return out


def coad_bracket(w6, t6):
    out = np.zeros(6)
    lib().orc_coad_bracket(_p(_f64(w6, 6)), _p(_f64(t6, 6)), _p(out))
    return out


def join_points(p3, q3):
    d, m, n = np.zeros(3), np.zeros(3), np.zeros(1)
    lib().orc_join_points(_p(_f64(p3, 3)), _p(_f64(q3, 3)), _p(d), _p(m), _p(n))
    return d, m, float(n[0])


def wrench_from_points(p3, q3):
    out = np.zeros(6)
    lib().orc_wrench_from_points(_p(_f64(p3, 3)), _p(_f64(q3, 3)), _p(out))
    return out


def screw_raw(s6):
    out = np.zeros(16)
    lib().orc_screw_raw(_p(_f64(s6, 6)), _p(out))
    return out


# ───────────── joints and solver stages ─────────────
SIMPLE_SPRING_DAMPER = 0
CRITICALLY_DAMPED_WARPED = 1
PERPENDICULAR_DAMPER_WARPED = 2
TORSIONAL_DAMPER_WARPED = 3


def joint_plain(kind, params, poses16, vels12, dt, warp):
    out = np.zeros(12)
    prm = np.zeros(10)
    prm[: len(params)] = params
    lib().orc_joint_plain(kind, _p(prm), _p(_f64(poses16, 16)), _p(_f64(vels12, 12)), float(dt), float(warp), _p(out))
    return out.reshape(2, 6)


def joint_jacobian(kind, params, poses16, vels12, dt, warp):
    val = np.zeros(12)
    jac = np.zeros(288)
    prm = np.zeros(10)
    prm[: len(params)] = params
    lib().orc_joint_jacobian(kind, _p(prm), _p(_f64(poses16, 16)), _p(_f64(vels12, 12)), float(dt), float(warp),
                             _p(val), _p(jac))
    return val.reshape(2, 6), jac.reshape(2, 24, 6)


def pre(pose8, vel6, retraction):
    out = np.zeros(8)
    lib().orc_pre(_p(_f64(pose8, 8)), _p(_f64(vel6, 6)), float(retraction), _p(out))
    return out


def body_post(pose8, vel6, mass, tensor9, snap6, total6, half, floor2=1e-18):
    block, rhs, scale = np.zeros(36), np.zeros(6), np.zeros(6)
    lib().orc_body_post(_p(_f64(pose8, 8)), _p(_f64(vel6, 6)), float(mass), _p(_f64(tensor9, 9)), _p(_f64(snap6, 6)),
                        _p(_f64(total6, 6)), float(half), float(floor2), _p(block), _p(rhs), _p(scale))
    return block.reshape(6, 6), rhs, scale


def pair_force(lo8, hi8):
    out = np.zeros(3)
    ok = lib().orc_pair_force(_p(_f64(lo8, 8)), _p(_f64(hi8, 8)), _p(out))
    return out if ok else None


def gravity_accumulate(charges, G=1.0):
    ch = _f64(charges)
    n = ch.size // 8
    out = np.zeros((n, 6))
    lib().orc_gravity_accumulate(n, _p(ch), float(G), _p(out))
    return out


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return lib().orc_max_threads()


def ns_solve(a_blocks, rhs, iters):
    """Transpose-seeded Newton–Schulz on block matrix a_blocks (m, m, 6, 6)."""
    a = _f64(a_blocks)
    m = int(round((a.size // 36) ** 0.5))
    x = np.zeros(m * m * 36)
    dv = np.zeros(m * 6)
    rho = np.zeros(max(iters, 1))
    lib().orc_ns_solve(m, _p(a), _p(_f64(rhs, m * 6)), int(iters), _p(x), _p(dv), _p(rho))
    return x.reshape(m, m, 6, 6), dv.reshape(m, 6), rho[:iters]


def screw_to_pod(c):
    """The Pod store of screws with constructor coordinates c (..., 6)."""
    c = np.asarray(c, dtype=np.float64)
    flat = np.ascontiguousarray(c.reshape(-1, 6))
    out = np.zeros_like(flat)
    for i in range(flat.shape[0]):
        lib().orc_screw_to_pod(_p(flat[i]), _p(out[i]))
    return out.reshape(c.shape)


def screw_from_pod(q):
    q = np.asarray(q, dtype=np.float64)
    flat = np.ascontiguousarray(q.reshape(-1, 6))
    out = np.zeros_like(flat)
    for i in range(flat.shape[0]):
        lib().orc_screw_from_pod(_p(flat[i]), _p(out[i]))
    return out.reshape(q.shape)


# ───────────── world ─────────────
INTEGRATOR = {"explicit_euler": 0, "symplectic_euler": 1, "newton": 2, "lie_euler": 3}
NO_SIGNAL = -100000


class OracleWorld:
    """The oracle's world: islands + joints + N-body gravity, stepped on the CPU."""

    def __init__(self, scene, integrator="newton", tol2=1e-8, ns_warm=2, ns_seed=24, gravity=True, G=1.0):
        L = lib()
        s = scene
        self.nb = int(s.n_bodies)
        self.ni = int(s.n_islands)
        self._keep = []

        def f(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return _p(a)

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(a)
            return a.ctypes.data_as(_ip)

        self.h = L.orc_world_new(self.nb, f(s.pose), f(s.momentum), f(s.mass), f(s.inertia), f(s.size), self.ni,
                                 i(s.island_start), f(s.origin), int(s.n_joints), i(s.joint_kind), i(s.joint_a),
                                 i(s.joint_b), f(s.joint_params), float(G), int(bool(gravity)),
                                 INTEGRATOR[integrator], float(tol2), int(ns_warm), int(ns_seed))
        self._keep = []

    def step(self, n=1, dt=1e-2, warp=1.0):
        lib().orc_world_step(self.h, int(n), float(dt), float(warp))

    def set_uniform(self, accel):
        lib().orc_world_set_uniform(self.h, float(accel[0]), float(accel[1]), float(accel[2]))

    def state(self):
        pose = np.zeros((self.nb, 8))
        mom = np.zeros((self.nb, 6))
        origin = np.zeros((self.ni, 3))
        absp = np.zeros((self.nb, 3))
        lib().orc_world_get(self.h, _p(pose), _p(mom), _p(origin), _p(absp))
        return {"pose": pose, "momentum": mom, "origin": origin, "position": absp}

    def front(self):
        out = np.zeros((self.nb, 8))
        lib().orc_world_front(self.h, _p(out))
        return out

    def stats(self):
        out = np.zeros((self.ni, 6), dtype=np.int64)
        lib().orc_world_stats(self.h, out.ctypes.data_as(_i64p))
        return out

    def span_trace(self, island=0, max_spans=1024):
        """The spans of `island`'s last step: (depth, converged, committed, signal) rows;
        signal is NO_SIGNAL when the step read none after that span."""
        out = np.zeros((max_spans, 4), dtype=np.int32)
        n = lib().orc_world_span_trace(self.h, int(island), out.ctypes.data_as(C.POINTER(C.c_int32)), max_spans)
        return out[:min(n, max_spans)].copy()

    def close(self):
        if getattr(self, "h", None):
            lib().orc_world_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleRawMap:
    def __init__(self, capacity):
        self.h = lib().orc_rawmap_new(int(capacity))

    def insert(self, value):
        i, g = C.c_uint32(), C.c_uint32()
        lib().orc_rawmap_insert(self.h, int(value), C.byref(i), C.byref(g))
        return i.value, g.value

    def remove(self, index, generation):
        v = C.c_uint64()
        ok = lib().orc_rawmap_remove(self.h, int(index), int(generation), C.byref(v))
        return v.value if ok else None

    def get(self, index, generation):
        v = C.c_uint64()
        ok = lib().orc_rawmap_get(self.h, int(index), int(generation), C.byref(v))
        return v.value if ok else None

    def info(self):
        cap, gen, ln = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ff = C.c_uint32()
        lib().orc_rawmap_info(self.h, C.byref(cap), C.byref(gen), C.byref(ln), C.byref(ff))
        return {"capacity": cap.value, "generation": gen.value, "len": ln.value, "first_free": ff.value}

    def __del__(self):
        try:
            lib().orc_rawmap_free(self.h)
        except Exception:
            pass
