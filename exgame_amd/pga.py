"""Host-side PGA3 in numpy, batched over bodies.

Used to build scenes and to convert between body-frame momenta and world
velocities on the host; the device kernels carry their own closed forms.

Conventions (the reference's, crates/clifford/src/algebra/mod.rs:84-147 and
markup.rs:13-100): blade = bitmask with e1 = bit 0, e2 = bit 1, e3 = bit 2 and the
degenerate e0 = bit 3; products take the permutation parity of merging into
ascending order; dual signs follow b ∧ dual(b) = +I.

  motor  = 8 floats, even blades in ascending order {1, e12, e13, e23, e01, e02, e03, e0123}
  screw  = 6 floats in constructor order (linear|force xyz, angular|torque xyz)
"""
from __future__ import annotations

import numpy as np

ZERO_MASK = 0b1000
EVEN_BLADES = (0, 3, 5, 6, 9, 10, 12, 15)
G2_BLADES = (3, 5, 6, 9, 10, 12)


def _popc(x: int) -> int:
    return bin(x).count("1")


def blade_prod(a: int, b: int):
    if (a & b) & ZERO_MASK:
        return 0, 0
    swaps = 0
    aa = a >> 1
    while aa:
        swaps += _popc(aa & b)
        aa >>= 1
    return (1 if swaps % 2 == 0 else -1), a ^ b


def complement_sign(a: int, b: int) -> int:
    inv = 0
    for bi in range(4):
        if (a >> bi) & 1:
            inv += _popc(b & ((1 << bi) - 1))
    return 1 if inv % 2 == 0 else -1


def _axis_vec(i: int) -> int:
    return 1 << i  # the Euclidean generators are the low three bits


def _euclid(i):
    v = _axis_vec(i)
    comp = 0b0111 ^ v
    return comp, complement_sign(v, comp)


def _radical(i):
    s, m = blade_prod(ZERO_MASK, _axis_vec(i))
    return m, s


def _point(i):
    em, es = _euclid(i)
    s, m = blade_prod(ZERO_MASK, em)
    return m, es * s


def _sigma() -> int:
    s_w, _ = blade_prod(_axis_vec(2), _axis_vec(0))
    _, s_e1 = _euclid(1)
    m_b, s_b = _euclid(2)
    m_p, s_p = _point(0)
    _, s_d = _point(1)
    s1, _ = blade_prod(m_b, m_p)
    s2, _ = blade_prod(m_p, m_b)
    return s_e1 * s_w * s_d * ((s1 - s2) // 2) * s_b * s_p


SIGMA = _sigma()
RADICAL = [_radical(i) for i in range(3)]
ANGULAR = [(_euclid(i)[0], _euclid(i)[1] * SIGMA) for i in range(3)]
POINT = [_point(i) for i in range(3)]

# dense Cayley tensor: out[k] = Σ_ij CAYLEY[i, j, k] a_i b_j
CAYLEY = np.zeros((16, 16, 16))
for _i in range(16):
    for _j in range(16):
        _s, _m = blade_prod(_i, _j)
        if _s:
            CAYLEY[_i, _j, _m] = _s
REV = np.array([-1.0 if ((_popc(b) * (_popc(b) - 1)) // 2) % 2 else 1.0 for b in range(16)])
DUAL = np.zeros((16, 16))
UNDUAL = np.zeros((16, 16))
for _s in range(16):
    DUAL[15 ^ _s, _s] = complement_sign(_s, 15 ^ _s)
    UNDUAL[15 ^ _s, _s] = complement_sign(15 ^ _s, _s)


def gp(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.einsum("...i,...j,ijk->...k", a, b, CAYLEY, optimize=True)


def motor16(m8: np.ndarray) -> np.ndarray:
    m8 = np.asarray(m8, dtype=np.float64)
    out = np.zeros(m8.shape[:-1] + (16,))
    out[..., list(EVEN_BLADES)] = m8
    return out


def motor8(m16: np.ndarray) -> np.ndarray:
    return np.asarray(m16)[..., list(EVEN_BLADES)]


def screw16(s6: np.ndarray) -> np.ndarray:
    s6 = np.asarray(s6, dtype=np.float64)
    out = np.zeros(s6.shape[:-1] + (16,))
    for i in range(3):
        m, s = RADICAL[i]
        out[..., m] += s * s6[..., i]
        m, s = ANGULAR[i]
        out[..., m] += s * s6[..., 3 + i]
    return out


def screw6(m16: np.ndarray) -> np.ndarray:
    m16 = np.asarray(m16)
    out = np.zeros(m16.shape[:-1] + (6,))
    for i in range(3):
        m, s = RADICAL[i]
        out[..., i] = s * m16[..., m]
        m, s = ANGULAR[i]
        out[..., 3 + i] = s * m16[..., m]
    return out


def point16(p3: np.ndarray) -> np.ndarray:
    p3 = np.asarray(p3, dtype=np.float64)
    out = np.zeros(p3.shape[:-1] + (16,))
    out[..., 7] = 1.0
    for i in range(3):
        m, s = POINT[i]
        out[..., m] += s * p3[..., i]
    return out


def point3(p16: np.ndarray) -> np.ndarray:
    p16 = np.asarray(p16)
    w = p16[..., 7]
    return np.stack([POINT[i][1] * p16[..., POINT[i][0]] / w for i in range(3)], axis=-1)


def reverse(a):
    return a * REV


def conjugate(m16, x16):
    return gp(gp(m16, x16), reverse(m16))


def motor_exp(b16: np.ndarray) -> np.ndarray:
    """Motor::exp of a bivector via the Study form (crates/clifford/src/pga3/motor.rs:194-221)."""
    b2 = gp(b16, b16)
    u = -b2[..., 0]
    p = b2[..., 15]
    l = np.sqrt(np.maximum(u, 0.0))
    small = np.abs(u) < 1e-9
    with np.errstate(invalid="ignore", divide="ignore"):
        su = np.where(small, 1 - u * (1 / 6 - u * (1 / 120 - u * (1 / 5040))), np.sin(l) / np.where(small, 1, l))
        cu = np.where(small, 1 - u * (1 / 2 - u * (1 / 24 - u * (1 / 720))), np.cos(l))
        smalld = np.abs(u * 1e-5) < 1e-9
        du = np.where(smalld, -1 / 6 + u * (1 / 60 - u * (1 / 1680 - u * (1 / 90720))),
                      (l * np.cos(l) - np.sin(l)) / np.where(smalld, 1, 2 * l * l * l))
    i16 = np.zeros(16)
    i16[15] = 1.0
    ib = gp(np.broadcast_to(i16, b16.shape), b16)
    out = np.zeros(b16.shape)
    out[..., 0] = cu
    out[..., 15] = p * 0.5 * su
    out += b16 * su[..., None] - ib * (p * du)[..., None]
    return out


def twist_exp(t6: np.ndarray, dt: float) -> np.ndarray:
    """Twist::exp(dt) = Motor::exp(½·dt·T) as a 16-blade motor."""
    return motor_exp(screw16(t6) * (0.5 * dt))


def compose(self16, rhs16):
    """`self.compose(rhs)` = rhs·self: first self, then rhs (motor.rs:149-158)."""
    return gp(rhs16, self16)


def pose_from(position: np.ndarray, rotation_vector: np.ndarray) -> np.ndarray:
    """Motor that rotates by `rotation_vector` (axis·angle) about the origin, then
    translates to `position`, as 8 components."""
    pos = np.asarray(position, dtype=np.float64)
    rv = np.asarray(rotation_vector, dtype=np.float64)
    z = np.zeros_like(pos)
    r = twist_exp(np.concatenate([z, rv], axis=-1), 1.0)
    t = twist_exp(np.concatenate([pos, z], axis=-1), 1.0)
    return motor8(compose(r, t))


def transport_contra(t6, m16):
    return screw6(conjugate(m16, screw16(t6)))


def transport_co(w6, m16):
    d = np.einsum("ij,...j->...i", DUAL, screw16(w6))
    return screw6(np.einsum("ij,...j->...i", UNDUAL, conjugate(m16, d)))


def position(m8: np.ndarray) -> np.ndarray:
    m16 = motor16(m8)
    return point3(conjugate(m16, point16(np.zeros(m16.shape[:-1] + (3,)))))


def inertia_apply(mass, tensor9, v6):
    t = np.asarray(tensor9).reshape(-1, 3, 3)
    lin = v6[..., :3] * np.asarray(mass)[..., None]
    ang = np.einsum("nij,nj->ni", t, v6[..., 3:])
    return np.concatenate([lin, ang], axis=-1)


def inertia_apply_inverse(mass, tensor9, p6):
    t = np.asarray(tensor9).reshape(-1, 3, 3)
    lin = p6[..., :3] * (1.0 / np.asarray(mass))[..., None]
    ang = np.linalg.solve(t, p6[..., 3:][..., None])[..., 0]
    return np.concatenate([lin, ang], axis=-1)


def world_velocity(m8, mass, tensor9, mom6):
    v_body = inertia_apply_inverse(mass, tensor9, mom6)
    return transport_contra(v_body, motor16(m8))


def world_momentum(m8, mom6):
    return transport_co(mom6, motor16(m8))
