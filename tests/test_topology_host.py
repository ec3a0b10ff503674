"""The per-island topology record the Newton kernels load every step
(exgame_amd/csrc/topology.cpp), built on the host and checked against what the joint
graph fixes: neighbour sets (the block sparsity of A), every (joint, end, end) term
and (joint, end) incidence exactly once, joints grouped by kind, slab masks equal to
the non-zero pattern of A, and MFMA tile lists that cover every output tile once."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libtopo.so")
SRC = os.path.join(HERE, "csrc", "topo_host.cpp")
DEPS = [SRC] + [os.path.join(HERE, "..", "exgame_amd", "csrc", f) for f in ("topology.cpp", "island.h")]

# record layout (island.h, struct IslandTopo)
MAXM, MAXJ, WAVES, PER_WAVE = 8, 16, 4, 4
FIELDS = [("kmask", "<u2", 4), ("tcount", "u1", 2 * WAVES), ("tlist", "u1", 2 * WAVES * PER_WAVE),
          ("nbr_start", "u1", MAXM + 1), ("nbr", "u1", MAXM * MAXM), ("term_start", "u1", MAXM * MAXM + 1),
          ("terms", "u1", 4 * MAXJ), ("inc_start", "u1", MAXM + 1), ("inc", "u1", 2 * MAXJ), ("jorder", "u1", MAXJ)]


@pytest.fixture(scope="module")
def tp():
    if not os.path.exists(SO) or any(os.path.getmtime(d) > os.path.getmtime(SO) for d in DEPS):
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        subprocess.run(["hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared", SRC,
                        "-o", SO], check=True)
    L = C.CDLL(SO)
    assert L.tp_size() == 320
    return L


def build(L, m, ja, jb, kind, waves, per_wave, md=None):
    J = len(ja)
    buf = np.zeros(320, dtype=np.uint8)
    arr = [np.ascontiguousarray(x, dtype=np.int32) for x in (ja, jb, kind)]
    ip = C.POINTER(C.c_int)
    rc = L.tp_build_md(m, m if md is None else md, J, *[x.ctypes.data_as(ip) for x in arr], waves, per_wave,
                       buf.ctypes.data_as(C.POINTER(C.c_ubyte)))
    out, off = {}, 0
    for name, dt, cnt in FIELDS:
        size = np.dtype(dt).itemsize * cnt
        out[name] = np.frombuffer(buf[off:off + size].tobytes(), dtype=dt).astype(np.int64)
        off += size
    out["tcount"] = out["tcount"].reshape(2, WAVES)
    out["tlist"] = out["tlist"].reshape(2, WAVES, PER_WAVE)
    return rc, out


def shape_of(m, J):
    return (2, 2) if (m <= 4 and J <= 8) else (4, 3)


def graphs():
    rng = np.random.default_rng(7)
    out = []
    for m in range(2, 9):   # chains
        ja = list(range(m - 1)); jb = list(range(1, m))
        out.append((m, ja, jb, [i % 4 for i in range(m - 1)]))
        # chain with a torsional damper per link (two joints per pair) and braces
        ja2 = ja + ja + [i for i in range(m - 2)]
        jb2 = jb + jb + [i + 2 for i in range(m - 2)]
        if len(ja2) <= MAXJ:
            out.append((m, ja2, jb2, list(rng.integers(0, 4, len(ja2)))))
    out.append((4, [0, 0, 0, 1, 2], [1, 2, 3, 2, 3], [0, 1, 2, 3, 0]))   # a 4-body rig
    for _ in range(40):   # random connected graphs
        m = int(rng.integers(2, 9))
        J = int(rng.integers(m - 1, MAXJ + 1))
        ja, jb = [], []
        for i in range(1, m):   # spanning tree first
            p = int(rng.integers(0, i))
            if rng.random() < 0.5: ja.append(p); jb.append(i)
            else: ja.append(i); jb.append(p)
        while len(ja) < J:
            x, y = rng.choice(m, 2, replace=False)
            ja.append(int(x)); jb.append(int(y))
        out.append((m, ja, jb, list(rng.integers(0, 4, J))))
    return out


@pytest.mark.parametrize("case", range(len(graphs())))
def test_topology_record(tp, case):
    m, ja, jb, kind = graphs()[case]
    J = len(ja)
    waves, per_wave = shape_of(m, J)
    rc, t = build(tp, m, ja, jb, kind, waves, per_wave)
    assert rc == 0
    # neighbour lists: the body itself first, then the other ends in joint order, once each
    for i in range(m):
        lst = list(t["nbr"][t["nbr_start"][i]:t["nbr_start"][i + 1]])
        want = [i]
        for a, b in zip(ja, jb):
            for me, other in ((a, b), (b, a)):
                if me == i and other not in want:
                    want.append(other)
        assert lst == want
    # every (joint, row end, column end) exactly once, filed under its block, in joint order
    seen = []
    for blk in range(m * m):
        bi, bj = divmod(blk, m)
        terms = list(t["terms"][t["term_start"][blk]:t["term_start"][blk + 1]])
        assert terms == sorted(terms)
        for e in terms:
            j, re, ce = e >> 2, (e >> 1) & 1, e & 1
            ends = (ja[j], jb[j])
            assert ends[re] == bi and ends[ce] == bj
            seen.append(e)
    assert sorted(seen) == [(j << 2) | k for j in range(J) for k in range(4)]
    # incidence: every (joint, end) once, under the body at that end
    seen = []
    for i in range(m):
        for e in t["inc"][t["inc_start"][i]:t["inc_start"][i + 1]]:
            j, end = e >> 1, e & 1
            assert (ja[j], jb[j])[end] == i
            seen.append(e)
    assert sorted(seen) == list(range(2 * J))
    # joints grouped by kind, stable within a kind
    order = list(t["jorder"][:J])
    assert sorted(order) == list(range(J))
    assert order == sorted(range(J), key=lambda j: (kind[j], j))
    # slab masks equal the non-zero pattern of A (6×6 blocks of coupled bodies)
    n = 6 * m
    T = (n + 15) // 16
    nz = np.zeros((16 * T, 16 * T), dtype=bool)
    for i in range(m):
        nz[6 * i:6 * i + 6, 6 * i:6 * i + 6] = True
    for a, b in zip(ja, jb):
        nz[6 * a:6 * a + 6, 6 * b:6 * b + 6] = True
        nz[6 * b:6 * b + 6, 6 * a:6 * a + 6] = True
    for ti in range(T):
        want = 0
        for ks in range(4 * T):
            if nz[16 * ti:16 * ti + 16, 4 * ks:4 * ks + 4].any():
                want |= 1 << ks
        assert t["kmask"][ti] == want, (ti, bin(t["kmask"][ti]), bin(want))
    # tile lists: each of the T² output tiles once per product, within the wave slots
    for k in range(2):
        tiles = []
        for w in range(WAVES):
            c = t["tcount"][k][w]
            assert c <= (per_wave if w < waves else 0)
            for u in range(c):
                e = t["tlist"][k][w][u]
                ti, tj = e & 15, e >> 4
                assert ti < T and tj < T
                tiles.append(ti * T + tj)
        assert sorted(tiles) == list(range(T * T))
    # the sparse product's lists: no wave carries more non-zero slabs than the
    # largest-first bound (mean load + the largest single tile)
    cost = [bin(int(t["kmask"][x // T])).count("1") for x in range(T * T)]
    loads = [sum(cost[(e & 15) * T + (e >> 4)] for e in t["tlist"][0][w][:t["tcount"][0][w]]) for w in range(waves)]
    assert max(loads) <= sum(cost) / waves + max(cost)


def test_topology_rejects_oversize(tp):
    rc, _ = build(tp, 9, [0], [1], [0], 4, 3)
    assert rc == -1
    rc, _ = build(tp, 2, [0] * 17, [1] * 17, [0] * 17, 4, 3)
    assert rc == -1
    rc, _ = build(tp, 3, [0], [3], [0], 4, 3)   # end outside the island
    assert rc == -1


# ───── the mechanism-sized record (island.h, struct LargeTopo) ─────
LMAXM, LMAXJ = 48, 128
LMAXNBR, LMAXT = LMAXM + 2 * LMAXJ, 6 * LMAXM // 16


def build_large(L, m, ja, jb, md=None):
    lay = np.zeros(9, dtype=np.int64)
    L.tp_large_layout(lay.ctypes.data_as(C.POINTER(C.c_longlong)))
    buf = np.zeros(int(lay[8]), dtype=np.uint8)
    arr = [np.ascontiguousarray(x, dtype=np.int32) for x in (ja, jb)]
    ip = C.POINTER(C.c_int)
    rc = L.tp_build_large_md(m, m if md is None else md, len(ja), *[x.ctypes.data_as(ip) for x in arr],
                             buf.ctypes.data_as(C.POINTER(C.c_ubyte)))
    fields = [("nbr_start", "<u2", LMAXM + 1), ("nbr", "u1", LMAXNBR), ("nbr_row", "u1", LMAXNBR),
              ("term_start", "<u2", LMAXNBR + 1), ("terms", "<u2", 4 * LMAXJ), ("inc_start", "<u2", LMAXM + 1),
              ("inc", "<u2", 2 * LMAXJ), ("kmask", "<u4", 3 * LMAXT)]
    out = {}
    for (name, dt, cnt), off in zip(fields, lay[:8]):
        size = np.dtype(dt).itemsize * cnt
        out[name] = np.frombuffer(buf[off:off + size].tobytes(), dtype=dt).astype(np.int64)
    out["kmask"] = out["kmask"].reshape(LMAXT, 3)
    return rc, out


def large_graphs():
    from exgame_amd import scenes
    rng = np.random.default_rng(11)
    out = []
    for shapes in (((4, 4),), ((5, 9),), ((6, 8),), ((3, 3),)):
        s = scenes.cloth_grids(shapes=shapes)
        out.append((s.n_bodies, list(s.joint_a), list(s.joint_b)))
    for m in (9, 12, 20, 48):   # chains with a torsional damper per link
        ja = list(range(m - 1)) * 2
        jb = list(range(1, m)) * 2
        out.append((m, ja[:LMAXJ], jb[:LMAXJ]))
    for _ in range(20):   # random connected graphs up to the record's limits
        m = int(rng.integers(9, LMAXM + 1))
        J = int(rng.integers(m - 1, LMAXJ + 1))
        ja, jb = [], []
        for i in range(1, m):
            p = int(rng.integers(0, i))
            ja.append(p); jb.append(i)
        while len(ja) < J:
            x, y = rng.choice(m, 2, replace=False)
            ja.append(int(x)); jb.append(int(y))
        out.append((m, ja, jb))
    return out


@pytest.mark.parametrize("case", range(len(large_graphs())))
def test_large_topology_record(tp, case):
    m, ja, jb = large_graphs()[case]
    J = len(ja)
    rc, t = build_large(tp, m, ja, jb)
    assert rc == 0
    nnz = t["nbr_start"][m]
    assert nnz <= LMAXNBR
    blocks = []
    for i in range(m):
        lst = list(t["nbr"][t["nbr_start"][i]:t["nbr_start"][i + 1]])
        want = [i]
        for a, b in zip(ja, jb):
            for me, other in ((a, b), (b, a)):
                if me == i and other not in want:
                    want.append(other)
        assert lst == want
        assert list(t["nbr_row"][t["nbr_start"][i]:t["nbr_start"][i + 1]]) == [i] * len(want)
        blocks += [(i, k) for k in want]
    # every (joint, row end, column end) exactly once, under its block, in joint order
    seen = []
    for q, (bi, bj) in enumerate(blocks):
        terms = list(t["terms"][t["term_start"][q]:t["term_start"][q + 1]])
        assert terms == sorted(terms)
        for e in terms:
            j, re, ce = e >> 2, (e >> 1) & 1, e & 1
            ends = (ja[j], jb[j])
            assert ends[re] == bi and ends[ce] == bj
            seen.append(e)
    assert sorted(seen) == [(j << 2) | k for j in range(J) for k in range(4)]
    seen = []
    for i in range(m):
        for e in t["inc"][t["inc_start"][i]:t["inc_start"][i + 1]]:
            j, end = e >> 1, e & 1
            assert (ja[j], jb[j])[end] == i
            seen.append(e)
    assert sorted(seen) == list(range(2 * J))
    # slab masks equal the non-zero pattern of A, tile by tile
    n = 6 * m
    T = (n + 15) // 16
    nz = np.zeros((16 * T, 16 * T), dtype=bool)
    for bi, bk in blocks:
        nz[6 * bi:6 * bi + 6, 6 * bk:6 * bk + 6] = True
    for ti in range(LMAXT):
        for w in range(3):
            want = 0
            for b in range(32):
                ks = 32 * w + b
                if ti < T and 4 * ks < 16 * T and nz[16 * ti:16 * ti + 16, 4 * ks:4 * ks + 4].any():
                    want |= 1 << b
            assert t["kmask"][ti][w] == want, (ti, w)
    n16 = 16 * T
    tp.tp_large_workspace.restype = C.c_longlong
    assert tp.tp_large_workspace(m, J) == 4 * n16 * n16 + 288 * J


def test_large_topology_rejects_oversize(tp):
    assert build_large(tp, 49, [0], [1])[0] == -1
    assert build_large(tp, 10, [0] * 129, [1] * 129)[0] == -1
    assert build_large(tp, 10, [0], [10])[0] == -1   # end outside the island
    assert build_large(tp, 10, [3], [3])[0] == -1    # a joint from a body to itself


@pytest.mark.parametrize("m,k", [(2, 1), (5, 1), (8, 1), (8, 2), (4, 4)])
def test_kinematic_bodies_have_no_rows_or_columns(tp, m, k):
    # the last k bodies are kinematic (references, not unknowns: newton.rs:806-830):
    # the system's tables are those of the dynamic bodies alone with the joints among
    # them; the incidence still lists every joint at both ends
    md = m - k
    rng = np.random.default_rng(m * 10 + k)
    ja = list(range(md - 1)) if md > 1 else []
    jb = list(range(1, md)) if md > 1 else []
    kind = [int(x) for x in rng.integers(0, 4, len(ja))]
    dyn = (list(ja), list(jb), list(kind))
    for q in range(k):   # joints to the kinematic bodies last
        ja.append(md + q); jb.append(int(rng.integers(0, md)) if md else (md + q + 1) % m); kind.append(q % 4)
    waves, per_wave = shape_of(m, len(ja))
    rc, t = build(tp, m, ja, jb, kind, waves, per_wave, md=md)
    assert rc == 0
    if md > 0:
        rc0, t0 = build(tp, md, *dyn, waves, per_wave)
        assert rc0 == 0
        for f in ("kmask", "tcount", "tlist"):
            np.testing.assert_array_equal(t[f], t0[f], err_msg=f)
        np.testing.assert_array_equal(t["nbr_start"][:md + 1], t0["nbr_start"][:md + 1])
        np.testing.assert_array_equal(t["nbr"][:t0["nbr_start"][md]], t0["nbr"][:t0["nbr_start"][md]])
    else:
        assert not t["kmask"].any() and not t["tcount"].any()
    # kinematic rows are empty lists; no term names a kinematic end
    assert (t["nbr_start"][md:m + 1] == t["nbr_start"][md]).all()
    for b_ in range(m * m):
        bi, bj = divmod(b_, m)
        n_terms = t["term_start"][b_ + 1] - t["term_start"][b_]
        if bi >= md or bj >= md:
            assert n_terms == 0
    for code in t["terms"][:t["term_start"][m * m]]:
        j, re, ce = code >> 2, (code >> 1) & 1, code & 1
        ends = (ja[j], jb[j])
        assert ends[re] < md and ends[ce] < md
    # incidence: every (joint, end) once, kinematic ends included
    inc = sorted(int(c) for c in t["inc"][:t["inc_start"][m]])
    assert inc == sorted([j << 1 for j in range(len(ja))] + [(j << 1) | 1 for j in range(len(ja))])


@pytest.mark.parametrize("m,k", [(12, 1), (20, 3), (48, 1)])
def test_large_record_kinematic_bodies_have_no_rows_or_columns(tp, m, k):
    # the mechanism-sized record with the last k bodies kinematic has the blocks of the
    # dynamic bodies alone (joints among them first, the joints to the kinematic bodies
    # last); a joint to a kinematic body adds one term, its dynamic end's own, to that
    # end's diagonal block; the incidence lists every joint at both ends
    md = m - k
    ja, jb = list(range(md - 1)), list(range(1, md))
    ja += ja[::2]; jb += jb[::2]   # a second joint on every other link
    dyn = (list(ja), list(jb))
    for q in range(k):
        ja.append(md + q); jb.append(q % md)
    rc, t = build_large(tp, m, ja, jb, md=md)
    rc0, t0 = build_large(tp, md, *dyn)
    assert rc == 0 and rc0 == 0
    nnz = int(t0["nbr_start"][md])
    assert (t["nbr_start"][md:m + 1] == nnz).all()
    np.testing.assert_array_equal(t["nbr_start"][:md + 1], t0["nbr_start"][:md + 1])
    for f in ("nbr", "nbr_row"):
        np.testing.assert_array_equal(t[f][:nnz], t0[f][:nnz])
    np.testing.assert_array_equal(t["kmask"], t0["kmask"])
    for q in range(nnz):
        got = [int(c) for c in t["terms"][t["term_start"][q]:t["term_start"][q + 1]]]
        want = [int(c) for c in t0["terms"][t0["term_start"][q]:t0["term_start"][q + 1]]]
        bi, bj = int(t["nbr_row"][q]), int(t["nbr"][q])
        for j in range(len(dyn[0]), len(ja)):   # joint j: end a kinematic, end b dynamic
            if bi == bj == jb[j]:
                want.append((j << 2) | (1 << 1) | 1)
        assert got == want, (q, got, want)
    inc = sorted(int(c) for c in t["inc"][:t["inc_start"][m]])
    assert inc == sorted([j << 1 for j in range(len(ja))] + [(j << 1) | 1 for j in range(len(ja))])
