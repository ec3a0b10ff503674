"""The reference's row-driven accelerator contract on the device: typed slot storage
(keys bit-exact with the oracle's RawMap) and one kernel per MessageKind, each
checked against the oracle's function on the same inputs.

Slots hold the reference's Pod layouts: the tests build every screw slot with the
oracle's own map from constructor coordinates to the grade-2 store (derived from its
sign law) and read results back through it. Storage scalars f64 (≤1e-12 relative)
and f32 (the reference's `usual::<f32>()`: the oracle is evaluated at the
f32-rounded inputs, outputs agree to f32 rounding, ≤1e-6 relative)."""
import numpy as np
import pytest

import pyoracle as O
from exgame_amd import store as S

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)
POD, COORD = O.screw_to_pod, O.screw_from_pod
F32_TOL = 1e-6                       # f32 outputs: one rounding to f32 (≈ 6e-8) with margin


def tol(scalar, f64_tol):
    return f64_tol if scalar == "f64" else F32_TOL


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(1.0, np.abs(b).max()))


def rnd(st, x):
    """x as the store holds it (f32 rounding for an f32 store)."""
    return np.asarray(np.asarray(x, dtype=st.dtype), dtype=np.float64)


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


def test_pod_screw_layout_is_the_oracles():
    # a Pre row copies its twist slot through the kernel's coordinate conversion:
    # what comes out must be the Pod bytes that went in, and the midpoint must be
    # the oracle's for the twist those bytes mean
    st = S.Store(capacity=16)
    acc = S.Accelerator(st)
    vel = RNG.normal(size=6)
    pose = rmotor()
    kv, kp = st.insert(S.TWIST, POD(vel)), st.insert(S.MOTOR, pose)
    kmid, ksv, kr = st.insert(S.MOTOR, np.zeros(8)), st.insert(S.TWIST, np.zeros(6)), st.insert(S.SCALAR, 0.7)
    acc.pre([row(st, [kv.index, kp.index, kmid.index, ksv.index, kr.index])])
    assert np.array_equal(st.read(S.TWIST, ksv), POD(vel))
    assert rel(st.read(S.MOTOR, kmid), O.pre(pose, vel, 0.7)) < 1e-14
    # the map is a signed permutation: (lin, ang) → (−a_z, a_y, −a_x, −l_x, −l_y, −l_z)
    np.testing.assert_array_equal(POD(np.arange(1.0, 7.0)), [-6, 5, -4, -1, -2, -3])


@pytest.mark.parametrize("scalar", ["f64", "f32"])
def test_pre_and_joint_kernels(scalar):
    st = S.Store(capacity=64, scalar=scalar)
    acc = S.Accelerator(st)
    pose = rnd(st, rmotor())
    vel = rnd(st, POD(RNG.normal(size=6)))
    kv, kp = st.insert(S.TWIST, vel), st.insert(S.MOTOR, pose)
    kmid, ksv = st.insert(S.MOTOR, np.zeros(8)), st.insert(S.TWIST, np.zeros(6))
    kr = st.insert(S.SCALAR, 0.0123)
    acc.pre([row(st, [kv.index, kp.index, kmid.index, ksv.index, kr.index])])
    assert rel(st.read(S.MOTOR, kmid), O.pre(pose, COORD(vel), float(rnd(st, 0.0123)))) < tol(scalar, 1e-14)
    assert np.array_equal(st.read(S.TWIST, ksv), vel)

    dt, warp = st.insert(S.SCALAR, 0.01), st.insert(S.SCALAR, 1.3)
    dtv, warpv = float(rnd(st, 0.01)), float(rnd(st, 1.3))
    params = {0: [0.3, -0.2, 0.15, -0.1, 0.25, -0.05, 1.3, 7.0, 2.0, 1e-6], 1: [0.3, -0.2, 0.15, -0.1, 0.25, -0.05, 1.3, 7.0, 2.0, 1e-6],
              2: [0.7, 1e-3], 3: [5.0]}
    for jk, prm in params.items():
        prm = list(rnd(st, prm))
        pa, pb = rnd(st, rmotor()), rnd(st, rmotor())
        va, vb = rnd(st, RNG.normal(size=6)), rnd(st, RNG.normal(size=6))   # coordinates, f32-exact
        keys = [st.insert(S.TWIST, POD(va)), st.insert(S.TWIST, POD(vb)), st.insert(S.MOTOR, pa),
                st.insert(S.MOTOR, pb)]
        pk = [st.insert(S.SCALAR, p).index for p in prm]
        conn = st.insert(S.WRENCH_PAIR, np.zeros(12))
        jac = st.insert(S.JACOBIAN, np.zeros(288))
        conn2 = st.insert(S.WRENCH_PAIR, np.zeros(12))
        base = [k.index for k in keys]
        acc.eval_joints(jk, False, [row(st, base + [conn.index, dt.index, warp.index] + pk)])
        acc.eval_joints(jk, True, [row(st, base + [jac.index, conn2.index, dt.index, warp.index] + pk)])
        want = O.joint_plain(jk, prm, np.concatenate([pa, pb]), np.concatenate([va, vb]), dtv, warpv)
        wv, wj = O.joint_jacobian(jk, prm, np.concatenate([pa, pb]), np.concatenate([va, vb]), dtv, warpv)
        assert rel(COORD(st.read(S.WRENCH_PAIR, conn).reshape(2, 6)), want) < tol(scalar, 1e-13)
        assert rel(COORD(st.read(S.WRENCH_PAIR, conn2).reshape(2, 6)), wv) < tol(scalar, 1e-13)
        assert rel(COORD(st.read(S.JACOBIAN, jac).reshape(2, 24, 6)), wj) < tol(scalar, 1e-12)


@pytest.mark.parametrize("scalar", ["f64", "f32"])
def test_gather_gemm_matvec_reduce_copy_sum(scalar):
    st = S.Store(capacity=64, scalar=scalar)
    acc = S.Accelerator(st)
    ext = rnd(st, RNG.normal(size=6))
    pairs = [rnd(st, RNG.normal(size=12)) for _ in range(3)]
    kext, kout = st.insert(S.WRENCH, POD(ext)), st.insert(S.WRENCH, np.zeros(6))
    kp = [st.insert(S.WRENCH_PAIR, POD(p.reshape(2, 6)).reshape(-1)) for p in pairs]
    terms = [kp[0].index * 2 + 1, kp[1].index * 2 + 0, kp[2].index * 2 + 1]
    acc.gather([row(st, [kout.index, kext.index, 0, 3] + terms)])
    sum1 = ext + pairs[0][6:] + pairs[1][:6] + pairs[2][6:]
    assert rel(COORD(st.read(S.WRENCH, kout)), sum1) < tol(scalar, 1e-15)
    # second round accumulates onto out
    acc.gather([row(st, [kout.index, kext.index, 1, 1, kp[0].index * 2])])
    assert rel(COORD(st.read(S.WRENCH, kout)), sum1 + pairs[0][:6]) < tol(scalar, 1e-15)

    A = [rnd(st, RNG.normal(size=(6, 6))) for _ in range(2)]
    B = [rnd(st, RNG.normal(size=(6, 6))) for _ in range(2)]
    ka, kb = [st.insert(S.BLOCK, a) for a in A], [st.insert(S.BLOCK, b) for b in B]
    kc = st.insert(S.BLOCK, np.zeros(36))
    kd = st.insert(S.BLOCK, np.zeros(36))
    plus, minus = S.f32_bits(1.0), S.f32_bits(-1.0)
    acc.gemm([[row(st, [kc.index, 0, 2, plus, ka[0].index, kb[0].index * 2, ka[1].index, kb[1].index * 2])],
              [row(st, [kd.index, 0, 2, minus, ka[0].index, kb[0].index * 2 + 1, ka[1].index, kb[1].index * 2])]])
    assert rel(st.read(S.BLOCK, kc).reshape(6, 6), A[0] @ B[0] + A[1] @ B[1]) < tol(scalar, 1e-13)
    assert rel(st.read(S.BLOCK, kd).reshape(6, 6), A[0] @ (2 * np.eye(6) - B[0]) + A[1] @ (-B[1])) < tol(scalar, 1e-13)

    rhs = [rnd(st, RNG.normal(size=6)) for _ in range(2)]
    kr = [st.insert(S.WRENCH, POD(r)) for r in rhs]
    kdv = st.insert(S.TWIST, np.zeros(6))
    acc.block_matvec([[row(st, [kdv.index, 0, 2, ka[0].index, kr[0].index, ka[1].index, kr[1].index])]])
    dv1 = A[0] @ rhs[0] + A[1] @ rhs[1]
    assert rel(COORD(st.read(S.TWIST, kdv)), dv1) < tol(scalar, 1e-13)
    acc.block_matvec([[row(st, [kdv.index, 1, 1, ka[1].index, kr[0].index])]])   # accumulate
    assert rel(COORD(st.read(S.TWIST, kdv)), dv1 + A[1] @ rhs[0]) < tol(scalar, 1e-13)

    kpart = st.insert(S.SCALAR, 0.0)
    acc.block_reduce([row(st, [kpart.index, 2, ka[0].index * 2 + 1, ka[1].index * 2])])
    want = np.sum((np.eye(6) - A[0]) ** 2) + np.sum(A[1] ** 2)
    assert abs(st.read(S.SCALAR, kpart)[0] - want) < tol(scalar, 1e-12) * want

    kcopy = st.insert(S.BLOCK, np.zeros(36))
    acc.block_copy([row(st, [kcopy.index, ka[1].index])])
    assert np.array_equal(st.read(S.BLOCK, kcopy).reshape(6, 6), A[1])

    x, y, z = st.insert(S.SCALAR, 1.25), st.insert(S.SCALAR, -3.5), st.insert(S.SCALAR, 0.0)
    acc.simple_sum([row(st, [x.index, y.index, z.index])])
    assert st.read(S.SCALAR, z)[0] == -2.25


@pytest.mark.parametrize("scalar", ["f64", "f32"])
@pytest.mark.parametrize("diagonal", [True, False])
def test_body_post_and_gathered(diagonal, scalar):
    st = S.Store(capacity=64, scalar=scalar)
    acc = S.Accelerator(st)
    half, floor2 = st.insert(S.SCALAR, 0.005), st.insert(S.SCALAR, 1e-18)
    hv, fv = float(rnd(st, 0.005)), float(rnd(st, 1e-18))
    rows = []
    want = []
    for b in range(3):
        m = rnd(st, rmotor())
        v = rnd(st, RNG.normal(size=6))
        mass = float(rnd(st, RNG.uniform(0.5, 2)))
        if diagonal:
            d = rnd(st, RNG.uniform(0.5, 2, 3))
            t9 = np.diag(d).reshape(-1)
            kang = st.insert(S.DIAG_INERTIA, d)
        else:
            R = np.linalg.qr(RNG.normal(size=(3, 3)))[0]
            t9 = rnd(st, (R @ np.diag(RNG.uniform(0.5, 2, 3)) @ R.T).reshape(-1))
            kang = st.insert(S.FULL_INERTIA, t9)
        snap, tot = rnd(st, RNG.normal(size=6)), rnd(st, RNG.normal(size=6))
        ks = [st.insert(S.MOTOR, m), st.insert(S.TWIST, POD(v)), st.insert(S.SCALAR, mass), kang,
              st.insert(S.WRENCH, POD(snap)), st.insert(S.WRENCH, POD(tot)), st.insert(S.BLOCK, np.zeros(36)),
              st.insert(S.WRENCH, np.zeros(6)), st.insert(S.WRENCH, np.zeros(6))]
        rows += [k.index for k in ks]
        want.append((ks, O.body_post(m, v, mass, t9, snap, tot, hv), (m, v, mass, t9, snap)))
    acc.body_post([row(st, [3, half.index, floor2.index] + rows)], diagonal)
    for ks, (wb, wr, wsc), _ in want:
        assert rel(st.read(S.BLOCK, ks[6]).reshape(6, 6), wb) < tol(scalar, 1e-13)
        assert rel(COORD(st.read(S.WRENCH, ks[7])), wr) < tol(scalar, 1e-12)
        assert rel(COORD(st.read(S.WRENCH, ks[8])), wsc) < tol(scalar, 1e-13)
    # gathered: total = external + incident pairs
    ks, _, (m, v, mass, t9, snap) = want[0]
    ext = rnd(st, RNG.normal(size=6))
    pr = rnd(st, RNG.normal(size=12))
    kext, kpr = st.insert(S.WRENCH, POD(ext)), st.insert(S.WRENCH_PAIR, POD(pr.reshape(2, 6)).reshape(-1))
    outs = [st.insert(S.BLOCK, np.zeros(36)), st.insert(S.WRENCH, np.zeros(6)), st.insert(S.WRENCH, np.zeros(6))]
    acc.body_post_gathered([row(st, [half.index, floor2.index, ks[0].index, ks[1].index, ks[2].index, ks[3].index,
                                     ks[4].index, kext.index] + [o.index for o in outs] + [1, kpr.index * 2 + 1])],
                           diagonal)
    wb, wr, _ = O.body_post(m, v, mass, t9, snap, ext + pr[6:], hv)
    assert rel(st.read(S.BLOCK, outs[0]).reshape(6, 6), wb) < tol(scalar, 1e-13)
    assert rel(COORD(st.read(S.WRENCH, outs[1])), wr) < tol(scalar, 1e-12)


@pytest.mark.parametrize("scalar", ["f64", "f32"])
def test_assemble_block(scalar):
    st = S.Store(capacity=64, scalar=scalar)
    acc = S.Accelerator(st)
    mass = rnd(st, RNG.normal(size=(6, 6)))
    J = rnd(st, RNG.normal(size=(2, 24, 6)))      # coordinates; stored as Pod wrenches
    km, kj = st.insert(S.BLOCK, mass), st.insert(S.JACOBIAN, POD(J).reshape(-1))
    kout = st.insert(S.BLOCK, np.zeros(36))
    pf, vf = st.insert(S.SCALAR, -(0.005 ** 2)), st.insert(S.SCALAR, -0.005)
    term = kj.index * 4 + 1 * 2 + 0   # row_end = b, col_end = a
    acc.assemble([[row(st, [kout.index, km.index, S.f32_bits(1.0), pf.index, vf.index, 0, 1, term])]])
    P = J[1, 0:6, :].T      # column col = jac[row_end][col_end·6 + col]
    V = J[1, 12:18, :].T
    want = mass + float(rnd(st, -(0.005 ** 2))) * P + float(rnd(st, -0.005)) * V
    assert rel(st.read(S.BLOCK, kout).reshape(6, 6), want) < tol(scalar, 1e-14)


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


def test_fatal_rows_are_marked_and_their_outputs_kept():
    # a connection row whose velocity slot holds a NaN yields non-finite wrenches: the
    # row is marked fatal (EXG_ERR_FATAL), every row's outputs are written
    st = S.Store(capacity=32)
    acc = S.Accelerator(st)
    dt, warp = st.insert(S.SCALAR, 0.01), st.insert(S.SCALAR, 1.0)
    prm = [0.3, 0, 0, -0.3, 0, 0, 1.0, 5.0, 1.0, 1e-6]
    rows, outs = [], []
    for vel in (RNG.normal(size=6), np.full(6, np.nan)):
        keys = [st.insert(S.TWIST, POD(vel)), st.insert(S.TWIST, np.zeros(6)), st.insert(S.MOTOR, rmotor()),
                st.insert(S.MOTOR, rmotor())]
        pk = [st.insert(S.SCALAR, p).index for p in prm]
        outs.append(st.insert(S.WRENCH_PAIR, np.zeros(12)))
        rows.append(row(st, [k.index for k in keys] + [outs[-1].index, dt.index, warp.index] + pk))
    with pytest.raises(Exception) as e:
        acc.eval_joints(0, False, rows)
    assert getattr(e.value, "code", None) == 3          # EXG_ERR_FATAL
    assert st.fatals(S.KIND["EvalSimpleSpringDamperPlain"]) == 1
    assert np.all(np.isfinite(st.read(S.WRENCH_PAIR, outs[0])))
    assert not np.all(np.isfinite(st.read(S.WRENCH_PAIR, outs[1])))
    acc.eval_joints(0, False, rows[:1])                 # a clean dispatch clears the count
    assert st.fatals(S.KIND["EvalSimpleSpringDamperPlain"]) == 0
