"""Seeded synthetic scenes for the workloads the hot path is measured on.

Every random draw comes from one counter-based SplitMix64 stream per purpose, so a
scene is a pure function of its seed and sizes (no numpy distribution code is
involved, whose algorithms may change between versions):

  u01(k)    = (splitmix64(seed_stream + k) >> 11) · 2⁻⁵³
  normal    = Box–Muller on two u01 draws
  lognormal = exp(μ + σ·normal)

A `Scene` is the flat, structure-of-arrays world the C ABI consumes:

  pose      (B, 8)  motor, island frame          momentum (B, 6) body frame (force|torque)
  mass      (B,)    inertia (B, 9) body tensor   size     (B,)   effective size (m/4)^(1/3)
  island_start (I+1,) bodies of island i are [start[i], start[i+1])
  origin    (I, 3)  island anchor in the absolute frame (the centre of mass)
  joint_kind/a/b (J,) with a, b global body indices inside one island; joint_params (J, 10)
  island_joint_start (I+1,) joints of island i, in declaration order
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import pga

SIMPLE_SPRING_DAMPER = 0
CRITICALLY_DAMPED_WARPED = 1
PERPENDICULAR_DAMPER_WARPED = 2
TORSIONAL_DAMPER_WARPED = 3

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based uniform stream: draw k of stream (seed, tag) is splitmix64 of a
    mixed 64-bit counter."""

    def __init__(self, seed: int, tag: int):
        base = splitmix64(np.array([np.uint64(seed) * np.uint64(0x100000001B3) ^ np.uint64(tag)], dtype=np.uint64))
        self.base = base[0]
        self.k = 0

    def u01(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = (np.arange(self.k, self.k + n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15) + self.base) & _M64
        self.k += n
        return (splitmix64(idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform(self, lo, hi, n):
        return lo + (hi - lo) * self.u01(n)

    def normal(self, n: int) -> np.ndarray:
        u1 = self.u01(n)
        u2 = self.u01(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)

    def unit_vectors(self, n: int) -> np.ndarray:
        v = np.stack([self.normal(n), self.normal(n), self.normal(n)], axis=-1)
        return v / np.linalg.norm(v, axis=-1, keepdims=True)

    def in_ball(self, n: int, radius: float) -> np.ndarray:
        r = radius * np.cbrt(self.u01(n))
        return self.unit_vectors(n) * r[:, None]

    def rotations(self, n: int) -> np.ndarray:
        """Uniform random rotations as rotation vectors (axis·angle), from unit quaternions."""
        q = np.stack([self.normal(n) for _ in range(4)], axis=-1)
        q /= np.linalg.norm(q, axis=-1, keepdims=True)
        q *= np.where(q[:, :1] < 0, -1.0, 1.0)
        ang = 2.0 * np.arccos(np.clip(q[:, 0], -1.0, 1.0))
        s = np.sqrt(np.maximum(1.0 - q[:, 0] ** 2, 0.0))
        axis = np.where(s[:, None] > 1e-12, q[:, 1:] / np.maximum(s, 1e-300)[:, None], np.array([1.0, 0, 0]))
        return axis * ang[:, None]


def rotation_matrix(rv: np.ndarray) -> np.ndarray:
    th = np.linalg.norm(rv, axis=-1)
    k = np.where(th[:, None] > 0, rv / np.maximum(th, 1e-300)[:, None], 0.0)
    K = np.zeros((len(rv), 3, 3))
    K[:, 0, 1], K[:, 0, 2], K[:, 1, 0] = -k[:, 2], k[:, 1], k[:, 2]
    K[:, 1, 2], K[:, 2, 0], K[:, 2, 1] = -k[:, 0], -k[:, 1], k[:, 0]
    s, c = np.sin(th)[:, None, None], np.cos(th)[:, None, None]
    return np.eye(3)[None] + s * K + (1 - c) * (K @ K)


@dataclass
class Scene:
    pose: np.ndarray
    momentum: np.ndarray
    mass: np.ndarray
    inertia: np.ndarray
    size: np.ndarray
    island_start: np.ndarray
    origin: np.ndarray
    joint_kind: np.ndarray
    joint_a: np.ndarray
    joint_b: np.ndarray
    joint_params: np.ndarray
    island_joint_start: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def n_bodies(self) -> int:
        return int(self.pose.shape[0])

    @property
    def n_islands(self) -> int:
        return int(self.origin.shape[0])

    @property
    def n_joints(self) -> int:
        return int(self.joint_kind.shape[0])

    def max_island(self) -> int:
        return int(np.max(np.diff(self.island_start))) if self.n_islands else 0

    def absolute_positions(self) -> np.ndarray:
        isl = np.repeat(np.arange(self.n_islands), np.diff(self.island_start))
        return pga.position(self.pose) + self.origin[isl]


def effective_size(mass: np.ndarray) -> np.ndarray:
    """(m/4)^(1/3) — crates/newton/src/body.rs:70-77, with the host's libm pow. A
    kinematic body (mass 0) gets 1: its pairs carry no mass, so its size never enters a
    force."""
    return np.array([float(np.power(m / 4.0, 1.0 / 3.0)) if m != 0.0 else 1.0
                     for m in np.asarray(mass, dtype=np.float64)])


def _assemble(positions, rotvecs, mass, tensors, vbody, island_sizes, joints, meta) -> Scene:
    """Island-local poses around each island's centre of mass (what the reference's
    Island::from_body/merge + recompute_centroid produce)."""
    B = len(mass)
    island_start = np.concatenate([[0], np.cumsum(island_sizes)]).astype(np.int32)
    I = len(island_sizes)
    origin = np.zeros((I, 3))
    local = np.zeros_like(positions)
    for k in range(I):
        s, e = island_start[k], island_start[k + 1]
        mm = mass[s:e]
        # an island without dynamic mass keeps its first body's position as its origin
        com = (positions[s:e] * mm[:, None]).sum(0) / mm.sum() if mm.sum() > 0.0 else positions[s].copy()
        origin[k] = com
        local[s:e] = positions[s:e] - com
    pose = pga.pose_from(local, rotvecs)
    mom = pga.inertia_apply(mass, tensors.reshape(B, 9), vbody)
    kin = np.asarray(mass) == 0.0
    mom[kin] = vbody[kin]   # a kinematic body's momentum slot holds its velocity (inertia.rs:135)
    if joints:
        jk, ja, jb, jp = zip(*joints)
        joint_kind = np.array(jk, dtype=np.int32)
        joint_a = np.array(ja, dtype=np.int32)
        joint_b = np.array(jb, dtype=np.int32)
        joint_params = np.array([list(p) + [0.0] * (10 - len(p)) for p in jp], dtype=np.float64)
    else:
        joint_kind = np.zeros(0, np.int32)
        joint_a = np.zeros(0, np.int32)
        joint_b = np.zeros(0, np.int32)
        joint_params = np.zeros((0, 10))
    # joints must be grouped by island in declaration order
    isl_of_body = np.repeat(np.arange(I), island_sizes)
    jisl = isl_of_body[joint_a] if len(joint_a) else np.zeros(0, np.int64)
    assert np.all(np.diff(jisl) >= 0), "joints must be declared island by island"
    if len(joint_a):
        assert np.all(isl_of_body[joint_b] == jisl), "a joint must stay inside one island"
    island_joint_start = np.searchsorted(jisl, np.arange(I + 1)).astype(np.int32)
    return Scene(pose=np.ascontiguousarray(pose), momentum=np.ascontiguousarray(mom), mass=mass.copy(),
                 inertia=np.ascontiguousarray(tensors.reshape(B, 9)), size=effective_size(mass),
                 island_start=island_start, origin=origin, joint_kind=joint_kind, joint_a=joint_a, joint_b=joint_b,
                 joint_params=joint_params, island_joint_start=island_joint_start, meta=meta)


def plain_gravity(n: int = 1024, seed: int = 0) -> Scene:
    """Point masses in a unit ball: velocities N(0, 0.05²) per axis, masses
    lognormal(0, 0.5) normalised to a total of 1, each its own island; isotropic
    inertia with moment 1 as RigidBody::body_at_with_speed_and_mass builds it."""
    st = Stream(seed, 1)
    pos = st.in_ball(n, 1.0)
    vel = 0.05 * np.stack([st.normal(n), st.normal(n), st.normal(n)], axis=-1)
    mass = np.exp(0.5 * st.normal(n))
    mass = mass / mass.sum()
    tens = np.tile(np.eye(3), (n, 1, 1))
    vb = np.concatenate([vel, np.zeros((n, 3))], axis=-1)
    return _assemble(pos, np.zeros((n, 3)), mass, tens, vb, np.ones(n, dtype=np.int64), [],
                     {"workload": "plain_gravity", "n": n, "seed": seed})


def chain_bodies(st: Stream, n_chains: int, m: int, radius: float, center=(0.0, 0.0, 0.0)):
    """Body draws for `n_chains` chains of `m` bodies: origins uniform in a ball,
    isotropic directions, link length 1 × U[0.95, 1.05]."""
    origins = st.in_ball(n_chains, radius) + np.asarray(center)[None]
    dirs = st.unit_vectors(n_chains)
    links = st.uniform(0.95, 1.05, n_chains * (m - 1)).reshape(n_chains, m - 1)
    s = np.concatenate([np.zeros((n_chains, 1)), np.cumsum(links, axis=1)], axis=1)
    s -= s.mean(axis=1, keepdims=True)
    pos = (origins[:, None, :] + dirs[:, None, :] * s[:, :, None]).reshape(-1, 3)
    return pos, links


def rigid_draws(st: Stream, n: int):
    """Masses U[0.5, 2], principal moments U[0.5, 2]³ under a random rotation (a full
    body-frame tensor), a random initial orientation and a body-frame twist with
    every component N(0, 0.05²)."""
    mass = st.uniform(0.5, 2.0, n)
    moments = st.uniform(0.5, 2.0, 3 * n).reshape(n, 3)
    R = rotation_matrix(st.rotations(n))
    tens = R @ (moments[:, :, None] * np.eye(3)[None]) @ np.transpose(R, (0, 2, 1))
    tens = 0.5 * (tens + np.transpose(tens, (0, 2, 1)))
    rot = st.rotations(n)
    vb = 0.05 * np.stack([st.normal(n) for _ in range(6)], axis=-1)
    return mass, tens, rot, vb


def chain_joints(first_body: int, links: np.ndarray, k: float, c: float, c_tors: float, eps: float):
    """Adjacent bodies: an axial spring-damper (anchors at the centres of mass, rest =
    link length) then a torsional damper, link by link."""
    out = []
    for q, L in enumerate(links):
        a, b = first_body + q, first_body + q + 1
        out.append((SIMPLE_SPRING_DAMPER, a, b, (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, float(L), k, c, eps)))
        out.append((TORSIONAL_DAMPER_WARPED, a, b, (c_tors,)))
    return out


def jointed_chains(n_chains: int = 512, m: int = 8, seed: int = 0, radius: float = 50.0, k: float = 1e3,
                   c: float = 10.0, c_tors: float = 5.0, eps: float = 1e-6, braced: bool = False) -> Scene:
    """Chains of `m` rigid bodies (4,096 bodies = 512 chains of 8 by default), each
    chain one island. With `braced` (m ≥ 4), each chain also gets a critically damped
    warped axial spring from body 0 to body 2 and a perpendicular damper from body 1
    to body 3: 2(m−1) + 2 joints per island and all four force laws in play."""
    st = Stream(seed, 2)
    pos, links = chain_bodies(st, n_chains, m, radius)
    n = n_chains * m
    mass, tens, rot, vb = rigid_draws(st, n)
    joints = []
    for ch in range(n_chains):
        joints += chain_joints(ch * m, links[ch], k, c, c_tors, eps)
        if braced and m >= 4:
            b0 = ch * m
            rest = float(links[ch][0] + links[ch][1])
            joints.append((CRITICALLY_DAMPED_WARPED, b0, b0 + 2, (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, rest, 0.5 * k, c, eps)))
            joints.append((PERPENDICULAR_DAMPER_WARPED, b0 + 1, b0 + 3, (0.5 * c, eps)))
    return _assemble(pos, rot, mass, tens, vb, np.full(n_chains, m, dtype=np.int64), joints,
                     {"workload": "jointed_chains", "n_chains": n_chains, "m": m, "seed": seed, "radius": radius,
                      "braced": braced})


def anchored_chains(n_chains: int = 16, m: int = 8, seed: int = 0, radius: float = 20.0, k: float = 1e3,
                    c: float = 10.0, c_tors: float = 5.0, eps: float = 1e-6, anchor_speed: float = 0.2,
                    kinematic_first: bool = False) -> Scene:
    """Chains of `m` bodies whose first body (along the chain) is kinematic (mass 0,
    Inertia::Kinematic, crates/newton/src/inertia.rs:103): it is a reference that the
    integrators never move, with a prescribed body-frame velocity of components
    N(0, anchor_speed²) that its joints' dampers see. The other bodies and joints are
    the jointed-chain ones. The anchor's slot is the island's last (the layout the
    device kernels take: dynamic bodies first) or, with `kinematic_first`, its first."""
    st = Stream(seed, 7)
    pos, links = chain_bodies(st, n_chains, m, radius)
    n = n_chains * m
    mass, tens, rot, vb = rigid_draws(st, n)
    anchor_v = anchor_speed * np.stack([st.normal(n_chains) for _ in range(6)], axis=-1)
    # along-chain index q → slot: the anchor (q = 0) last, or first
    slot = np.array([(m - 1 if q == 0 else q - 1) if not kinematic_first else q for q in range(m)])
    P, Mv, T, R, V = (np.zeros_like(pos), np.zeros(n), np.zeros_like(tens), np.zeros_like(rot), np.zeros_like(vb))
    joints = []
    for ch in range(n_chains):
        b0 = ch * m
        for q in range(m):
            src, dst = b0 + q, b0 + slot[q]
            P[dst], R[dst] = pos[src], rot[src]
            if q == 0:
                Mv[dst], T[dst], V[dst] = 0.0, np.zeros((3, 3)), anchor_v[ch]
            else:
                Mv[dst], T[dst], V[dst] = mass[src], tens[src], vb[src]
        for (kind, a, b, prm) in chain_joints(0, links[ch], k, c, c_tors, eps):
            joints.append((kind, b0 + int(slot[a]), b0 + int(slot[b]), prm))
    return _assemble(P, R, Mv, T, V, np.full(n_chains, m, dtype=np.int64), joints,
                     {"workload": "anchored_chains", "n_chains": n_chains, "m": m, "seed": seed,
                      "anchor_speed": anchor_speed, "kinematic_first": kinematic_first})


def cloth_grids(shapes=((4, 4), (5, 9)), seed: int = 0, radius: float = 20.0, k: float = 1e3, c: float = 10.0,
                c_tors: float = 5.0, eps: float = 1e-6, chains=()) -> Scene:
    """Mechanism-sized islands: one cloth per (rows, cols) in `shapes`, every body
    dynamic. A cloth is a grid of rigid bodies at spacing 1 × U[0.95, 1.05] in a
    random plane, jittered by U[−0.02, 0.02] per axis; axial spring-dampers join
    grid neighbours along both axes (rest = the initial distance), torsional dampers
    follow the rows, and the first cell carries a critically damped warped diagonal
    and a perpendicular damper across the other diagonal. A 5 × 9 cloth has 45
    bodies and 118 joints, a 4 × 4 cloth 16 and 38. `chains` adds chains of the
    given lengths (the jointed-chain joints: 2(m − 1) per chain)."""
    st = Stream(seed, 5)
    positions, rots, masses, tensors, vbs, sizes, joints = [], [], [], [], [], [], []
    body = 0
    centers = st.in_ball(len(shapes), radius)
    R = rotation_matrix(st.rotations(len(shapes)))
    for q, (rows, cols) in enumerate(shapes):
        m = rows * cols
        gap = st.uniform(0.95, 1.05, 2)
        grid = np.array([[gap[0] * (r - 0.5 * (rows - 1)), gap[1] * (cc - 0.5 * (cols - 1)), 0.0]
                         for r in range(rows) for cc in range(cols)])
        grid += st.uniform(-0.02, 0.02, 3 * m).reshape(m, 3)
        p = centers[q][None] + grid @ R[q].T
        mass, tens, rot, vb = rigid_draws(st, m)
        positions.append(p); rots.append(rot); masses.append(mass); tensors.append(tens); vbs.append(vb)

        def at(r, cc):
            return body + r * cols + cc

        def dist(u, v):
            return float(np.linalg.norm(p[u - body] - p[v - body]))

        for r in range(rows):
            for cc in range(cols):
                if cc + 1 < cols:
                    u, v = at(r, cc), at(r, cc + 1)
                    joints.append((SIMPLE_SPRING_DAMPER, u, v, (0, 0, 0, 0, 0, 0, dist(u, v), k, c, eps)))
                    joints.append((TORSIONAL_DAMPER_WARPED, u, v, (c_tors,)))
                if r + 1 < rows:
                    u, v = at(r, cc), at(r + 1, cc)
                    joints.append((SIMPLE_SPRING_DAMPER, u, v, (0, 0, 0, 0, 0, 0, dist(u, v), k, c, eps)))
        if rows >= 2 and cols >= 2:
            u, v = at(0, 0), at(1, 1)
            joints.append((CRITICALLY_DAMPED_WARPED, u, v, (0, 0, 0, 0, 0, 0, dist(u, v), 0.5 * k, c, eps)))
            joints.append((PERPENDICULAR_DAMPER_WARPED, at(0, 1), at(1, 0), (0.5 * c, eps)))
        sizes.append(m)
        body += m
    for m in chains:
        p, links = chain_bodies(st, 1, m, radius)
        mass, tens, rot, vb = rigid_draws(st, m)
        positions.append(p); rots.append(rot); masses.append(mass); tensors.append(tens); vbs.append(vb)
        joints += chain_joints(body, links[0], k, c, c_tors, eps)
        sizes.append(m)
        body += m
    return _assemble(np.concatenate(positions), np.concatenate(rots), np.concatenate(masses),
                     np.concatenate(tensors), np.concatenate(vbs), np.array(sizes, dtype=np.int64), joints,
                     {"workload": "cloth_grids", "shapes": [list(s) for s in shapes], "chains": list(chains),
                      "seed": seed, "radius": radius})


def mixed_world(n_bodies: int, seed: int, radius: float | None = None, rigs: bool = False, center=(0.0, 0.0, 0.0),
                k: float = 1e3, c: float = 10.0, c_tors: float = 5.0, eps: float = 1e-6) -> Scene:
    """60% free gravitating bodies (singleton islands) and 40% jointed 8-body chains;
    with `rigs`, part of the chain budget goes to 4-body rigs (a ring of axial
    springs plus torsional dampers)."""
    st = Stream(seed, 3)
    n_chain_bodies = int(round(0.4 * n_bodies / 8.0)) * 8
    n_rig_bodies = 0
    if rigs:
        n_rig_bodies = (n_chain_bodies // 2) // 4 * 4
        n_chain_bodies -= n_rig_bodies
    n_free = n_bodies - n_chain_bodies - n_rig_bodies
    if radius is None:
        radius = 50.0 * (n_bodies / 4096.0) ** (1.0 / 3.0)
    positions, rots, masses, tensors, vbs, sizes, joints = [], [], [], [], [], [], []
    body = 0
    if n_free:
        p = st.in_ball(n_free, radius) + np.asarray(center)[None]
        mass, tens, rot, vb = rigid_draws(st, n_free)
        positions.append(p); rots.append(rot); masses.append(mass); tensors.append(tens); vbs.append(vb)
        sizes += [1] * n_free
        body += n_free
    nch = n_chain_bodies // 8
    if nch:
        p, links = chain_bodies(st, nch, 8, radius, center)
        mass, tens, rot, vb = rigid_draws(st, nch * 8)
        positions.append(p); rots.append(rot); masses.append(mass); tensors.append(tens); vbs.append(vb)
        for ch in range(nch):
            joints += chain_joints(body + ch * 8, links[ch], k, c, c_tors, eps)
            sizes.append(8)
        body += nch * 8
    nrig = n_rig_bodies // 4
    if nrig:
        # a rig: four bodies on a unit square, springs around the ring and across
        centers = st.in_ball(nrig, radius) + np.asarray(center)[None]
        corner = np.array([[0.5, 0.5, 0.0], [-0.5, 0.5, 0.0], [-0.5, -0.5, 0.0], [0.5, -0.5, 0.0]])
        R = rotation_matrix(st.rotations(nrig))
        p = (centers[:, None, :] + np.einsum("nij,kj->nki", R, corner)).reshape(-1, 3)
        mass, tens, rot, vb = rigid_draws(st, nrig * 4)
        positions.append(p); rots.append(rot); masses.append(mass); tensors.append(tens); vbs.append(vb)
        for r in range(nrig):
            b0 = body + 4 * r
            for (u, v, L) in ((0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0), (3, 0, 1.0), (0, 2, 2 ** 0.5)):
                joints.append((SIMPLE_SPRING_DAMPER, b0 + u, b0 + v, (0, 0, 0, 0, 0, 0, L, k, c, eps)))
                joints.append((TORSIONAL_DAMPER_WARPED, b0 + u, b0 + v, (c_tors,)))
            sizes.append(4)
        body += nrig * 4
    pos = np.concatenate(positions)
    rot = np.concatenate(rots)
    mass = np.concatenate(masses)
    tens = np.concatenate(tensors)
    vb = np.concatenate(vbs)
    # joints per island, declaration order — islands are laid out free, chains, rigs
    return _assemble(pos, rot, mass, tens, vb, np.array(sizes, dtype=np.int64), joints,
                     {"workload": "mixed_world", "n": n_bodies, "seed": seed, "rigs": rigs})


def batched_world_sizes(n_worlds: int = 512, seed: int = 0, lo: int = 64, hi: int = 8192, median: int = 512):
    """Body counts lognormal(ln median, ln(16)/3) clipped to [lo, hi], multiples of 8."""
    st = Stream(seed, 4)
    sizes = np.exp(np.log(median) + (np.log(16.0) / 3.0) * st.normal(n_worlds))
    sizes = np.clip(sizes, lo, hi)
    return (np.round(sizes / 8.0) * 8).astype(np.int64)
