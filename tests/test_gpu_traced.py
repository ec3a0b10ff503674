"""The joint and Pre row kernels generated by tracing the scalar-generic laws
(exgame_amd/viete, csrc/viete_gen.h; the reference's build.rs + viete route,
crates/newton/build.rs:673-821) against the hand-written kernels and the oracle, on
the same rows: many random rows per law, both storage scalars.

Tolerances: traced vs hand-written ≤ 1e-13 relative to the largest magnitude (both
evaluate the same f64 operations; they differ only where the FMA pass and the
compiler fused different products), traced vs the oracle ≤ 1e-12 (as the
hand-written kernels are held)."""
import numpy as np
import pytest

import pyoracle as O
from exgame_amd import store as S

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2024)
POD, COORD = O.screw_to_pod, O.screw_from_pod


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(1.0, np.abs(b).max()))


def rmotor():
    return O.twist_exp(RNG.normal(size=6), 1.0)


def params_of(jk):
    if jk in (0, 1):
        return list(0.3 * RNG.normal(size=6)) + [RNG.uniform(0.5, 2.0), RNG.uniform(1e2, 1e4), RNG.uniform(1, 20), 1e-6]
    if jk == 2:
        return [RNG.uniform(0.5, 5.0), 1e-6]
    return [RNG.uniform(0.5, 5.0)]


def joint_rows(st, jk, n, jac):
    """n random rows of law jk; returns (row keys, output keys, oracle inputs)."""
    dt, warp = st.insert(S.SCALAR, 0.01), st.insert(S.SCALAR, 1.3)
    rows, outs, cases = [], [], []
    for _ in range(n):
        prm = params_of(jk)
        pa, pb = rmotor(), rmotor()
        va, vb = RNG.normal(size=6), RNG.normal(size=6)
        base = [st.insert(S.TWIST, POD(va)).index, st.insert(S.TWIST, POD(vb)).index, st.insert(S.MOTOR, pa).index,
                st.insert(S.MOTOR, pb).index]
        pk = [st.insert(S.SCALAR, p).index for p in prm]
        if jac:
            kj, kc = st.insert(S.JACOBIAN, np.zeros(288)), st.insert(S.WRENCH_PAIR, np.zeros(12))
            rows.append(st.insert(S.ROW, base + [kj.index, kc.index, dt.index, warp.index] + pk))
            outs.append((kj, kc))
        else:
            kc = st.insert(S.WRENCH_PAIR, np.zeros(12))
            rows.append(st.insert(S.ROW, base + [kc.index, dt.index, warp.index] + pk))
            outs.append((kc,))
        cases.append((prm, np.concatenate([pa, pb]), np.concatenate([va, vb])))
    return rows, outs, cases


@pytest.mark.parametrize("jk", [0, 1, 2, 3])
@pytest.mark.parametrize("jac", [False, True])
def test_traced_joint_rows_equal_handwritten_and_oracle(jk, jac):
    st = S.Store(capacity=8192)
    acc = S.Accelerator(st)
    print(f"law {jk} jacobian={jac}: 48 rows", flush=True)
    rows, outs, cases = joint_rows(st, jk, 48, jac)
    st.set_codegen("handwritten")
    acc.eval_joints(jk, jac, rows)
    hand = [[st.read(S.JACOBIAN if (jac and i == 0) else S.WRENCH_PAIR, k) for i, k in enumerate(o)] for o in outs]
    st.set_codegen("traced")
    for o in outs:   # clear the outputs so that the traced kernel must write them
        for i, k in enumerate(o):
            st.write(S.JACOBIAN if (jac and i == 0) else S.WRENCH_PAIR, k, np.zeros(288 if (jac and i == 0) else 12))
    acc.eval_joints(jk, jac, rows)
    for o, h, (prm, poses, vels) in zip(outs, hand, cases):
        got = [st.read(S.JACOBIAN if (jac and i == 0) else S.WRENCH_PAIR, k) for i, k in enumerate(o)]
        for g, w in zip(got, h):
            assert rel(g, w) < 1e-13
        if jac:
            wv, wj = O.joint_jacobian(jk, prm, poses, vels, 0.01, 1.3)
            assert rel(COORD(got[0].reshape(2, 24, 6)), wj) < 1e-12
            assert rel(COORD(got[1].reshape(2, 6)), wv) < 1e-12
        else:
            assert rel(COORD(got[0].reshape(2, 6)), O.joint_plain(jk, prm, poses, vels, 0.01, 1.3)) < 1e-12


@pytest.mark.parametrize("scalar", ["f64", "f32"])
def test_traced_pre_rows_across_the_seams(scalar):
    # rotation parts of zero, 1e-6, 1e-3 and 1: every branch of the decision trie
    st = S.Store(capacity=1024, scalar=scalar)
    acc = S.Accelerator(st)
    rows, outs, cases = [], [], []
    for s in (0.0, 1e-6, 1e-3, 1.0) * 8:
        v = np.concatenate([RNG.normal(size=3), s * RNG.normal(size=3)])
        pose = rmotor()
        r = float(RNG.uniform(0.001, 0.6))
        kv, kp = st.insert(S.TWIST, POD(v)), st.insert(S.MOTOR, pose)
        km, ks, kr = st.insert(S.MOTOR, np.zeros(8)), st.insert(S.TWIST, np.zeros(6)), st.insert(S.SCALAR, r)
        rows.append(st.insert(S.ROW, [kv.index, kp.index, km.index, ks.index, kr.index]))
        outs.append((km, ks))
        cases.append((pose, v, r))
    st.set_codegen("handwritten")
    acc.pre(rows)
    hand = [(st.read(S.MOTOR, km), st.read(S.TWIST, ks)) for km, ks in outs]
    for km, ks in outs:
        st.write(S.MOTOR, km, np.zeros(8))
    st.set_codegen("traced")
    acc.pre(rows)
    f32 = scalar == "f32"
    for (km, ks), (hm, hs), (pose, v, r) in zip(outs, hand, cases):
        gm = st.read(S.MOTOR, km)
        assert rel(gm, hm) < (1e-6 if f32 else 1e-14)
        assert np.array_equal(st.read(S.TWIST, ks), hs)
        if not f32:
            assert rel(gm, O.pre(pose, v, r)) < 1e-14


def test_codegen_mode_is_validated():
    st = S.Store(capacity=4)
    with pytest.raises(KeyError):
        st.set_codegen("glsl")
    from exgame_amd import _lib
    with pytest.raises(_lib.ExgError):
        _lib.check(st.L.exg_store_set_codegen(st.h, 7))
