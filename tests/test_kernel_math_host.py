"""The device kernels' closed-form physics (exgame_amd/csrc/physics_dev.h, the same
`__host__ __device__` code the HIP kernels run) compiled for the host and compared
with the oracle on the CPU: motor exp/compose, sandwiches, transports, every joint
law and its Jacobian lanes, body_post and commit. Catches a wrong generated formula
without a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libkmath.so")
SRC = os.path.join(HERE, "csrc", "kmath.cpp")
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def km():
    deps = [SRC] + [os.path.join(HERE, "..", "exgame_amd", "csrc", f) for f in ("physics_dev.h", "pga3_gen.h", "dual.h")]
    if not os.path.exists(SO) or any(os.path.getmtime(d) > os.path.getmtime(SO) for d in deps):
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        subprocess.run(["hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
                        "-ffp-contract=off", SRC, "-o", SO], check=True)
    L = C.CDLL(SO)
    L.km_pairing.restype = C.c_double
    return L


def p(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


RNG = np.random.default_rng(2024)


def rmotor():
    return O.twist_exp(RNG.normal(size=6), 1.0)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(1.0, np.abs(np.asarray(b)).max())


def test_small_angle_sin_cos_within_one_ulp(km):
    # the kernels' sin/cos: a series for |x| <= 1/4, the library beyond; against the
    # host libm through numpy (correctly rounded to within an ulp) and the exact
    # special values
    x = np.concatenate([RNG.uniform(-0.25, 0.25, 20000), np.ldexp(RNG.uniform(-1, 1, 2000), -RNG.integers(3, 60, 2000)),
                        [0.25, -0.25, np.nextafter(0.25, 1), 0.3, 1.0, 3.0, -7.5, 0.0, -0.0, 1e-300]])
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    km.km_sincos(x.ctypes.data_as(_dp), len(x), s.ctypes.data_as(_dp), c.ctypes.data_as(_dp))
    ulp_s = np.abs(s - np.sin(x)) / np.spacing(np.abs(np.sin(x)) + 1e-320)
    ulp_c = np.abs(c - np.cos(x)) / np.spacing(np.abs(np.cos(x)))
    assert ulp_s.max() <= 1.0 and ulp_c.max() <= 1.0, (ulp_s.max(), ulp_c.max())
    assert s[-3] == 0.0 and np.signbit(s[-2]) and c[-3] == 1.0 and s[-1] == 1e-300


def test_motor_ops(km):
    for _ in range(10):
        t = RNG.normal(size=6)
        dt = RNG.uniform(0.01, 2.0)
        out, op = p(np.zeros(8))
        _, tp = p(t)
        km.km_twist_exp(tp, C.c_double(dt), op)
        assert rel(out, O.twist_exp(t, dt)) < 1e-14
        a, b = rmotor(), rmotor()
        _, ap = p(a)
        _, bp = p(b)
        km.km_compose(ap, bp, op)
        assert rel(out, O.compose(a, b)) < 1e-14
    # the Taylor arm of the seams: tiny angles
    t = np.array([1e-3, -5e-4, 2e-4, 3e-6, -1e-6, 2e-6])
    out, op = p(np.zeros(8))
    km.km_twist_exp(p(t)[1], C.c_double(1.0), op)
    assert rel(out, O.twist_exp(t, 1.0)) < 1e-15


def test_sandwiches_and_transports(km):
    for _ in range(10):
        m = rmotor()
        x = RNG.normal(size=3)
        out, op = p(np.zeros(4))
        km.km_point(p(m)[1], p(x)[1], op)
        assert rel(out[1:] / out[0], O.apply_point(m, x)) < 1e-14
        t = RNG.normal(size=6)
        o6, o6p = p(np.zeros(6))
        km.km_transport_contra(p(t)[1], p(m)[1], o6p)
        assert rel(o6, O.twist_transport(t, m)) < 1e-14
        km.km_transport_co(p(t)[1], p(m)[1], o6p)
        assert rel(o6, O.wrench_transport(t, m)) < 1e-14
        w = RNG.normal(size=6)
        assert abs(km.km_pairing(p(w)[1], p(t)[1]) - O.power(w, t)) < 1e-14 * max(1, abs(O.power(w, t)))
        v3, v3p = p(np.zeros(3))
        km.km_velocity_at(p(t)[1], p(x)[1], v3p)
        assert rel(v3, O.velocity_at(t, x)) < 1e-14
        km.km_coad_bracket(p(w)[1], p(t)[1], o6p)
        assert rel(o6, O.coad_bracket(w, t)) < 1e-14


PARAMS = {
    0: [0.3, -0.2, 0.15, -0.1, 0.25, -0.05, 1.3, 7.0, 2.0, 1e-6],
    1: [0.3, -0.2, 0.15, -0.1, 0.25, -0.05, 1.3, 7.0, 2.0, 1e-6],
    2: [0.7, 1e-3] + [0.0] * 8,
    3: [5.0] + [0.0] * 9,
}


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_joints_and_jacobian_lanes(km, kind):
    prm = PARAMS[kind]
    for _ in range(5):
        poses = np.concatenate([rmotor(), rmotor()])
        vels = RNG.normal(size=12)
        dt, warp = 0.011, 1.3
        out, op = p(np.zeros(12))
        km.km_joint_plain(kind, p(prm)[1], p(poses)[1], p(vels)[1], C.c_double(dt), C.c_double(warp), op)
        want = O.joint_plain(kind, prm, poses, vels, dt, warp)
        assert rel(out.reshape(2, 6), want) < 1e-13
        val, vp = p(np.zeros(12))
        jac, jp = p(np.zeros(288))
        km.km_joint_jacobian(kind, p(prm)[1], p(poses)[1], p(vels)[1], C.c_double(dt), C.c_double(warp), vp, jp)
        wv, wj = O.joint_jacobian(kind, prm, poses, vels, dt, warp)
        assert rel(val.reshape(2, 6), wv) < 1e-13
        assert rel(jac.reshape(2, 24, 6), wj) < 1e-12


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_f32_jacobian_lanes_track_the_f64_jacobian(km, kind):
    # the mixed-precision Jacobian: Dual<float> lanes on f32-narrowed poses and
    # velocities; relative error a few f32 ulps times the conditioning of the law
    prm = PARAMS[kind]
    worst = 0.0
    for _ in range(5):
        poses = np.concatenate([rmotor(), rmotor()])
        vels = RNG.normal(size=12)
        dt, warp = 0.011, 1.3
        jac, jp = p(np.zeros(288))
        km.km_joint_jacobian_f32(kind, p(prm)[1], p(poses)[1], p(vels)[1], C.c_double(dt), C.c_double(warp), jp)
        _, wj = O.joint_jacobian(kind, prm, poses, vels, dt, warp)
        worst = max(worst, rel(jac.reshape(2, 24, 6), wj))
    assert worst < 2e-5, worst


def test_body_post_matches(km):
    for _ in range(5):
        m = rmotor()
        v = RNG.normal(size=6)
        R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
        t9 = (R @ np.diag(RNG.uniform(0.5, 2, 3)) @ R.T).reshape(-1)
        snap, tot = RNG.normal(size=6), RNG.normal(size=6)
        blk, bp = p(np.zeros(36))
        rhs, rp = p(np.zeros(6))
        sc, sp = p(np.zeros(2))
        km.km_body_post(p(m)[1], p(v)[1], C.c_double(1.7), p(t9)[1], p(snap)[1], p(tot)[1], C.c_double(0.005),
                        C.c_double(1e-18), bp, rp, sp)
        wb, wr, ws = O.body_post(m, v, 1.7, t9, snap, tot, 0.005)
        assert rel(blk.reshape(6, 6), wb) < 1e-13
        assert rel(rhs, wr) < 1e-12
        assert rel(sc, ws[[0, 3]]) < 1e-13


def test_commit_matches_oracle_world_pieces(km):
    # commit = the momentum/pose update of newton.rs:708-753; check against a host
    # composition of oracle primitives
    for conv in (0, 1):
        m0 = rmotor()
        R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
        t9 = (R @ np.diag(RNG.uniform(0.5, 2, 3)) @ R.T).reshape(-1)
        vs, snap = RNG.normal(size=6), RNG.normal(size=6)
        mass, span = 1.3, 0.01
        pose, pp = p(m0.copy())
        mom, mp = p(np.zeros(6))
        km.km_commit(pp, mp, C.c_double(mass), p(t9)[1], p(vs)[1], p(snap)[1], C.c_double(span), C.c_double(span / 2),
                     conv)
        pose_mid = O.compose(m0, O.twist_exp(vs, span / 2))
        vb = O.twist_transport(vs, O.motor_inverse(pose_mid))
        pb = np.concatenate([mass * vb[:3], t9.reshape(3, 3) @ vb[3:]])
        pmid = O.wrench_transport(pb, pose_mid)
        pnew = 2 * pmid - snap if conv else pmid
        mnew = O.compose(m0, O.twist_exp(vs, span))
        assert rel(pose, mnew) < 1e-14
        assert rel(mom, O.wrench_transport(pnew, O.motor_inverse(mnew))) < 1e-13
