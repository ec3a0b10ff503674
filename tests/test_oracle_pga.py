"""Pins the oracle's PGA3 restatement to the reference's own convention tests.

Each check cites the reference test it restates; the values are what those tests
assert (geometry fixed by the +I anchor), not the oracle's formulas retyped.
"""
import numpy as np
import pytest

import pyoracle as O

Z3 = [0.0, 0.0, 0.0]


def tw(lin, ang):
    return list(lin) + list(ang)


def close(a, b, eps=1e-9):
    return np.max(np.abs(np.asarray(a, float) - np.asarray(b, float))) < eps


# ── peano/src/scalar.rs:199-330 ──
def test_study_seams_endpoints():
    assert abs(O.sinc_sq(0.0) - 1.0) < 1e-15
    assert abs(O.sinc_sq(np.pi ** 2)) < 1e-14
    assert abs(O.cos_sq(0.0) - 1.0) < 1e-15
    assert abs(O.cos_sq(np.pi ** 2) + 1.0) < 1e-14
    assert abs(O.dsinc_sq(0.0) + 1.0 / 6.0) < 1e-13


@pytest.mark.parametrize("u", [0.0, 1e-12, 1e-10, 9e-9, 1e-8, 1.1e-8, 1e-7, 1e-6, 1e-5])
def test_sinc_branches_agree(u):
    l = np.sqrt(u)
    ref = 1.0 if l == 0 else np.sin(l) / l
    assert abs(O.sinc_sq(u) - ref) < 1e-13


@pytest.mark.parametrize("u", [0.0, 1e-12, 1e-10, 9e-10, 1e-9, 1.1e-9, 1e-8, 1e-6, 1e-4])
def test_cos_branches_agree(u):
    assert abs(O.cos_sq(u) - np.cos(np.sqrt(u))) < 1e-13


def test_sinc_decreasing_and_dsinc_is_derivative():
    a, b, c = O.sinc_sq(0.01), O.sinc_sq(0.5), O.sinc_sq(2.0)
    assert 1.0 > a > b > c
    u, h = 0.5, 1e-7
    fd = (O.sinc_sq(u + h) - O.sinc_sq(u - h)) / (2 * h)
    assert abs(O.dsinc_sq(u) - fd) < 1e-7


# ── clifford/src/algebra/mod.rs tests ──
def test_quaternion_style_blade_signs_and_degeneracy():
    assert O.blade_prod(0b1000, 0b1000) == (0, 0)         # e0² = 0
    s, m = O.blade_prod(0b1000, 0b0001)
    assert m == 0b1001 and s != 0
    assert O.blade_prod(1, 2) == (1, 3)                     # e1e2 = +e12
    assert O.blade_prod(2, 1) == (-1, 3)                    # anticommutation
    for b in range(16):                                     # b ∧ dual(b) = +I
        comp = 15 ^ b
        s, c = O.blade_prod(b, comp)
        assert c == 15 and s * O.lib().orc_complement_sign(b, comp) == 1


# ── clifford/src/pga3/motor.rs:368-560 ──
def test_rigid_motions():
    idm = np.array([1.0, 0, 0, 0, 0, 0, 0, 0])
    assert close(O.apply_point(idm, [1, 2, 3]), [1, 2, 3])
    mz = O.twist_exp(tw(Z3, [0, 0, np.pi / 2]), 1.0)
    assert close(O.apply_point(mz, [1, 0, 0]), [0, 1, 0])
    mt = O.twist_exp(tw([2, 0, 0], Z3), 1.0)
    assert close(O.apply_point(mt, Z3), [2, 0, 0])


def test_translator_half_is_forced():
    # T = 1 + c·B with B the bivector of Twist::new(r, 0): the sandwich shifts by 2c·r
    r = np.array([1.0, 2.0, 3.0])
    raw = O.screw_raw(tw(r, Z3))
    for c, want in ((0.5, r), (1.0 / 3.0, 2.0 / 3.0 * r)):
        m8 = np.zeros(8)
        m8[0] = 1.0
        m8[4:7] = c * raw[[9, 10, 12]]
        assert close(O.apply_point(m8, Z3), want, 1e-12)


@pytest.mark.parametrize("wz,dt", [(np.pi, 1.0), (np.pi / 2, 2.0)])
def test_rotation_pi_and_dt_linearity(wz, dt):
    m = O.twist_exp(tw(Z3, [0, 0, wz]), dt)
    assert close(O.apply_point(m, [1, 0, 0]), [-1, 0, 0])


def test_rotor_unit_and_inverse():
    m = O.twist_exp(tw(Z3, [0.3, -0.2, 0.5]), 1.0)
    assert abs(np.sum(m[:4] ** 2) - 1.0) < 1e-9
    a = O.motor_exp_raw(tw([0.1, -0.2, 0.3], [0.4, 0.1, -0.2]))
    b = O.motor_exp_raw(tw([0.2, 0.1, -0.1], [-0.3, 0.2, 0.1]))
    p = O.compose(a, O.motor_inverse(a))
    assert close(p, [1, 0, 0, 0, 0, 0, 0, 0])
    assert not close(O.compose(a, b), O.compose(b, a), 1e-6)   # the motor group is not commutative
    # compose(a, b) = "a, then b": applying it moves a point as b(a(p))
    q = np.array([0.3, -1.0, 2.0])
    assert close(O.apply_point(O.compose(a, b), q), O.apply_point(b, O.apply_point(a, q)), 1e-12)


# ── clifford/src/pga3/screw.rs:375-685 ──
def test_power_invariance_under_motor_transport():
    w = [2.0, -1.0, 0.5, 0.3, 0.8, -0.4]
    t = [1.0, 2.0, 3.0, 0.5, -1.0, 0.7]
    for m in (O.twist_exp(tw([0.3, -0.2, 0.5], [0.1, 0.4, -0.2]), 1.0), O.motor_exp_raw(tw(Z3, [0.4, -0.7, 0.3]))):
        assert abs(O.power(w, t) - O.power(O.wrench_transport(w, m), O.twist_transport(t, m))) < 1e-9
    # and twist/wrench transport differ
    m = O.twist_exp(tw([0.3, -0.2, 0.5], [0.1, 0.4, -0.2]), 1.0)
    assert not close(O.twist_transport(t, m), O.wrench_transport(t, m), 1e-6)


def test_power_is_the_pairing_f_v_plus_tau_omega():
    w = np.array([1.0, -2.0, 0.5, 0.3, 0.7, -0.1])
    t = np.array([0.4, 2.0, -1.0, 1.0, 0.0, 3.0])
    assert abs(O.power(w, t) - float(w @ t)) < 1e-12


def test_wrench_parallel_axis():
    f, m0, r = np.array([2.0, -1.0, 3.0]), np.array([0.5, 1.0, -0.5]), np.array([1.0, 2.0, 3.0])
    mt = O.twist_exp(tw(r, Z3), 1.0)
    w = O.wrench_transport(tw(f, m0), mt)
    assert close(w[:3], f)
    dtau = w[3:] - m0
    assert abs(dtau @ dtau - np.cross(r, f) @ np.cross(r, f)) < 1e-9


def test_se3_structure_constants_and_jacobi():
    ex, ey = [1.0, 0, 0], [0, 1.0, 0]
    rr = O.twist_bracket(tw(Z3, ex), tw(Z3, ey))
    assert abs(rr[5]) > 0.5 and close(rr[:3], Z3, 1e-12)
    tt = O.twist_bracket(tw(ex, Z3), tw(ey, Z3))
    assert close(tt, np.zeros(6), 1e-12)
    rt = O.twist_bracket(tw(Z3, ex), tw(ey, Z3))
    assert abs(rt[2]) > 0.5 and close(rt[3:], Z3, 1e-12)
    a = tw([1, 2, 3], [0.5, -1, 0.7])
    b = tw([-2, 1, 0.5], [0.3, 0.8, -0.4])
    c = tw([0.2, -0.6, 1.1], [-0.5, 0.9, 0.2])
    br = O.twist_bracket
    j = br(a, br(b, c)) + br(b, br(c, a)) + br(c, br(a, b))
    assert np.abs(j).max() < 1e-12


def test_ad_is_bracket_automorphism_and_coad_duality():
    a = tw([1, 2, 3], [0.5, -1, 0.7])
    b = tw([-2, 1, 0.5], [0.3, 0.8, -0.4])
    m = O.twist_exp(tw([0.3, -0.2, 0.5], [0.1, 0.4, -0.2]), 1.0)
    lhs = O.twist_transport(O.twist_bracket(a, b), m)
    rhs = O.twist_bracket(O.twist_transport(a, m), O.twist_transport(b, m))
    assert close(lhs, rhs)
    w = tw([2.0, -1.0, 0.5], [0.3, 0.8, -0.4])
    t = tw([1.0, 2.0, 3.0], [0.5, -1.0, 0.7])
    s = tw([-0.5, 0.9, 0.2], [0.2, -0.6, 1.1])
    assert abs(O.power(O.coad_bracket(w, t), s) + O.power(w, O.twist_bracket(t, s))) < 1e-9


def test_velocity_at_is_omega_cross_r():
    w = np.array([0.2, -0.5, 1.3])
    r = np.array([1.0, 2.0, -0.5])
    assert close(O.velocity_at(tw(Z3, w), r), np.cross(w, r))
    v = np.array([0.3, 0.1, -0.7])   # spatial twist: v(r) = v + ω×r
    assert close(O.velocity_at(tw(v, w), r), v + np.cross(w, r))


# ── clifford/src/pga3/objects.rs:212-289 ──
def test_line_direction_and_distance():
    d, mom, n = O.join_points([1, 0, 0], [0, 0, 0])
    assert close(np.abs(d), [1, 0, 0]) and abs(n - 1.0) < 1e-12
    d, mom, n = O.join_points([1, 1, 0], [-1, 1, 0])
    assert abs(np.linalg.norm(mom) / n - 1.0) < 1e-12
    d, mom, n = O.join_points([1, 0, 0], [-1, 0, 0])
    assert np.linalg.norm(mom) < 1e-12
    d, mom, n = O.join_points([1, 2, 3], [1, 2, 3])
    assert n == 0.0 and np.abs(d).max() == 0.0


def test_wrench_of_a_line_is_force_along_it_with_moment():
    p, q = np.array([1.0, 2.0, -1.0]), np.array([0.5, -1.0, 2.0])
    w = O.wrench_from_points(p, q)
    # |force| = |p − q|, direction ∥ (p − q), moment p × F
    assert close(np.abs(w[:3]), np.abs(p - q), 1e-12)
    assert close(w[3:], np.cross(p, w[:3]), 1e-12)
