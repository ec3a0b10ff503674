"""The reference's row-driven accelerator contract on the device: typed slot storage
(keys bit-exact with the oracle's RawMap) and one kernel per MessageKind, each
checked against the oracle's function on the same inputs (f64, ≤1e-12 relative)."""
import numpy as np
import pytest

import pyoracle as O
from exgame_amd import store as S

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(1.0, np.abs(b).max()))


def rmotor():
    return O.twist_exp(RNG.normal(size=6), 1.0)


def row(st, words):
    return st.insert(S.ROW, words)


def test_keys_are_bit_exact_with_rawmap():
    st = S.Store(capacity=4)
    ref = O.OracleRawMap(4)
    live = []
    for step in range(400):
        if live and RNG.uniform() < 0.45:
            k = live.pop(int(RNG.integers(len(live))))
            got = st.remove(S.SCALAR, S.Key(*k))
            want = ref.remove(*k)
            assert (got is None) == (want is None)
            assert got is None or got[0] == float(want)
            # a stale key is refused by both
            assert st.read(S.SCALAR, S.Key(*k)) is None and ref.get(*k) is None
        else:
            v = int(RNG.integers(1 << 40))
            k1 = st.insert(S.SCALAR, float(v))
            k2 = ref.insert(v)
            assert (k1.index, k1.generation) == k2
            live.append(k2)
    info, rinfo = st.info(S.SCALAR), ref.info()
    assert info["capacity"] == rinfo["capacity"] and info["generation"] == rinfo["generation"]
    assert info["len"] == rinfo["len"]
    for k in live:
        assert st.read(S.SCALAR, S.Key(*k))[0] == float(ref.get(*k))


def test_pre_and_joint_kernels():
    st = S.Store(capacity=64)
    acc = S.Accelerator(st)
    pose = rmotor()
    vel = RNG.normal(size=6)
    kv, kp = st.insert(S.TWIST, vel), st.insert(S.MOTOR, pose)
    kmid, ksv = st.insert(S.MOTOR, np.zeros(8)), st.insert(S.TWIST, np.zeros(6))
    kr = st.insert(S.SCALAR, 0.0123)
    acc.pre([row(st, [kv.index, kp.index, kmid.index, ksv.index, kr.index])])
    assert rel(st.read(S.MOTOR, kmid), O.pre(pose, vel, 0.0123)) < 1e-14
    assert np.array_equal(st.read(S.TWIST, ksv), vel)

    dt, warp = st.insert(S.SCALAR, 0.01), st.insert(S.SCALAR, 1.3)
    params = {0: [0.3, -0.2, 0.15, -0.1, 0.25, -0.05, 1.3, 7.0, 2.0, 1e-6], 1: [0.3, -0.2, 0.15, -0.1, 0.25, -0.05, 1.3, 7.0, 2.0, 1e-6],
              2: [0.7, 1e-3], 3: [5.0]}
    for jk, prm in params.items():
        pa, pb = rmotor(), rmotor()
        va, vb = RNG.normal(size=6), RNG.normal(size=6)
        keys = [st.insert(S.TWIST, va), st.insert(S.TWIST, vb), st.insert(S.MOTOR, pa), st.insert(S.MOTOR, pb)]
        pk = [st.insert(S.SCALAR, p).index for p in prm]
        conn = st.insert(S.WRENCH_PAIR, np.zeros(12))
        jac = st.insert(S.JACOBIAN, np.zeros(288))
        conn2 = st.insert(S.WRENCH_PAIR, np.zeros(12))
        base = [k.index for k in keys]
        acc.eval_joints(jk, False, [row(st, base + [conn.index, dt.index, warp.index] + pk)])
        acc.eval_joints(jk, True, [row(st, base + [jac.index, conn2.index, dt.index, warp.index] + pk)])
        want = O.joint_plain(jk, prm, np.concatenate([pa, pb]), np.concatenate([va, vb]), 0.01, 1.3)
        wv, wj = O.joint_jacobian(jk, prm, np.concatenate([pa, pb]), np.concatenate([va, vb]), 0.01, 1.3)
        assert rel(st.read(S.WRENCH_PAIR, conn).reshape(2, 6), want) < 1e-13
        assert rel(st.read(S.WRENCH_PAIR, conn2).reshape(2, 6), wv) < 1e-13
        assert rel(st.read(S.JACOBIAN, jac).reshape(2, 24, 6), wj) < 1e-12


def test_gather_gemm_matvec_reduce_copy_sum():
    st = S.Store(capacity=64)
    acc = S.Accelerator(st)
    ext = RNG.normal(size=6)
    pairs = [RNG.normal(size=12) for _ in range(3)]
    kext, kout = st.insert(S.WRENCH, ext), st.insert(S.WRENCH, np.zeros(6))
    kp = [st.insert(S.WRENCH_PAIR, p) for p in pairs]
    terms = [kp[0].index * 2 + 1, kp[1].index * 2 + 0, kp[2].index * 2 + 1]
    acc.gather([row(st, [kout.index, kext.index, 0, 3] + terms)])
    assert rel(st.read(S.WRENCH, kout), ext + pairs[0][6:] + pairs[1][:6] + pairs[2][6:]) < 1e-15
    # second round accumulates onto out
    acc.gather([row(st, [kout.index, kext.index, 1, 1, kp[0].index * 2])])
    assert rel(st.read(S.WRENCH, kout), ext + pairs[0][6:] + pairs[1][:6] + pairs[2][6:] + pairs[0][:6]) < 1e-15

    A = [RNG.normal(size=(6, 6)) for _ in range(2)]
    B = [RNG.normal(size=(6, 6)) for _ in range(2)]
    ka, kb = [st.insert(S.BLOCK, a) for a in A], [st.insert(S.BLOCK, b) for b in B]
    kc = st.insert(S.BLOCK, np.zeros(36))
    kd = st.insert(S.BLOCK, np.zeros(36))
    plus, minus = S.f32_bits(1.0), S.f32_bits(-1.0)
    acc.gemm([[row(st, [kc.index, 0, 2, plus, ka[0].index, kb[0].index * 2, ka[1].index, kb[1].index * 2])],
              [row(st, [kd.index, 0, 2, minus, ka[0].index, kb[0].index * 2 + 1, ka[1].index, kb[1].index * 2])]])
    assert rel(st.read(S.BLOCK, kc).reshape(6, 6), A[0] @ B[0] + A[1] @ B[1]) < 1e-13
    assert rel(st.read(S.BLOCK, kd).reshape(6, 6), A[0] @ (2 * np.eye(6) - B[0]) + A[1] @ (-B[1])) < 1e-13

    rhs = [RNG.normal(size=6) for _ in range(2)]
    kr = [st.insert(S.WRENCH, r) for r in rhs]
    kdv = st.insert(S.TWIST, np.zeros(6))
    acc.block_matvec([[row(st, [kdv.index, 0, 2, ka[0].index, kr[0].index, ka[1].index, kr[1].index])]])
    assert rel(st.read(S.TWIST, kdv), A[0] @ rhs[0] + A[1] @ rhs[1]) < 1e-13

    kpart = st.insert(S.SCALAR, 0.0)
    acc.block_reduce([row(st, [kpart.index, 2, ka[0].index * 2 + 1, ka[1].index * 2])])
    want = np.sum((np.eye(6) - A[0]) ** 2) + np.sum(A[1] ** 2)
    assert abs(st.read(S.SCALAR, kpart)[0] - want) < 1e-12 * want

    kcopy = st.insert(S.BLOCK, np.zeros(36))
    acc.block_copy([row(st, [kcopy.index, ka[1].index])])
    assert np.array_equal(st.read(S.BLOCK, kcopy).reshape(6, 6), A[1])

    x, y, z = st.insert(S.SCALAR, 1.25), st.insert(S.SCALAR, -3.5), st.insert(S.SCALAR, 0.0)
    acc.simple_sum([row(st, [x.index, y.index, z.index])])
    assert st.read(S.SCALAR, z)[0] == -2.25


@pytest.mark.parametrize("diagonal", [True, False])
def test_body_post_and_gathered(diagonal):
    st = S.Store(capacity=64)
    acc = S.Accelerator(st)
    half, floor2 = st.insert(S.SCALAR, 0.005), st.insert(S.SCALAR, 1e-18)
    rows = []
    want = []
    for b in range(3):
        m = rmotor()
        v = RNG.normal(size=6)
        mass = float(RNG.uniform(0.5, 2))
        if diagonal:
            d = RNG.uniform(0.5, 2, 3)
            t9 = np.diag(d).reshape(-1)
            kang = st.insert(S.DIAG_INERTIA, d)
        else:
            R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
            t9 = (R @ np.diag(RNG.uniform(0.5, 2, 3)) @ R.T).reshape(-1)
            kang = st.insert(S.FULL_INERTIA, t9)
        snap, tot = RNG.normal(size=6), RNG.normal(size=6)
        ks = [st.insert(S.MOTOR, m), st.insert(S.TWIST, v), st.insert(S.SCALAR, mass), kang, st.insert(S.WRENCH, snap),
              st.insert(S.WRENCH, tot), st.insert(S.BLOCK, np.zeros(36)), st.insert(S.WRENCH, np.zeros(6)),
              st.insert(S.WRENCH, np.zeros(6))]
        rows += [k.index for k in ks]
        want.append((ks, O.body_post(m, v, mass, t9, snap, tot, 0.005)))
    acc.body_post([row(st, [3, half.index, floor2.index] + rows)], diagonal)
    for ks, (wb, wr, wsc) in want:
        assert rel(st.read(S.BLOCK, ks[6]).reshape(6, 6), wb) < 1e-13
        assert rel(st.read(S.WRENCH, ks[7]), wr) < 1e-12
        assert rel(st.read(S.WRENCH, ks[8]), wsc) < 1e-13
    # gathered: total = external + incident pairs
    ks, _ = want[0]
    ext = RNG.normal(size=6)
    pr = RNG.normal(size=12)
    kext, kpr = st.insert(S.WRENCH, ext), st.insert(S.WRENCH_PAIR, pr)
    m, v = st.read(S.MOTOR, ks[0]), st.read(S.TWIST, ks[1])
    mass = st.read(S.SCALAR, ks[2])[0]
    t9 = np.diag(st.read(S.DIAG_INERTIA, ks[3])).reshape(-1) if diagonal else st.read(S.FULL_INERTIA, ks[3])
    snap = st.read(S.WRENCH, ks[4])
    outs = [st.insert(S.BLOCK, np.zeros(36)), st.insert(S.WRENCH, np.zeros(6)), st.insert(S.WRENCH, np.zeros(6))]
    acc.body_post_gathered([row(st, [half.index, floor2.index, ks[0].index, ks[1].index, ks[2].index, ks[3].index,
                                     ks[4].index, kext.index] + [o.index for o in outs] + [1, kpr.index * 2 + 1])],
                           diagonal)
    wb, wr, _ = O.body_post(m, v, mass, t9, snap, ext + pr[6:], 0.005)
    assert rel(st.read(S.BLOCK, outs[0]).reshape(6, 6), wb) < 1e-13
    assert rel(st.read(S.WRENCH, outs[1]), wr) < 1e-12


def test_assemble_block():
    st = S.Store(capacity=64)
    acc = S.Accelerator(st)
    mass = RNG.normal(size=(6, 6))
    J = RNG.normal(size=(2, 24, 6))
    km, kj = st.insert(S.BLOCK, mass), st.insert(S.JACOBIAN, J)
    kout = st.insert(S.BLOCK, np.zeros(36))
    pf, vf = st.insert(S.SCALAR, -(0.005 ** 2)), st.insert(S.SCALAR, -0.005)
    term = kj.index * 4 + 1 * 2 + 0   # row_end = b, col_end = a
    acc.assemble([[row(st, [kout.index, km.index, S.f32_bits(1.0), pf.index, vf.index, 0, 1, term])]])
    P = J[1, 0:6, :].T      # column col = jac[row_end][col_end·6 + col]
    V = J[1, 12:18, :].T
    want = mass + (-(0.005 ** 2)) * P + (-0.005) * V
    assert rel(st.read(S.BLOCK, kout).reshape(6, 6), want) < 1e-14


def test_dispatch_refuses_bad_rows_and_two_writers():
    st = S.Store(capacity=8)
    acc = S.Accelerator(st)
    a = st.insert(S.BLOCK, np.eye(6))
    out = st.insert(S.BLOCK, np.zeros(36))
    with pytest.raises(Exception, match="past its storage"):
        acc.block_copy([row(st, [out.index, 5000])])
    with pytest.raises(Exception, match="ledger"):
        acc.block_copy([row(st, [out.index, a.index]), row(st, [out.index, a.index])])
    with pytest.raises(Exception, match="no kernel"):
        st.dispatch(77, [row(st, [0])])
    # nothing was launched: out is untouched
    assert np.array_equal(st.read(S.BLOCK, out), np.zeros(36))
