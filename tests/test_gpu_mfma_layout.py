"""Lane layout of the FP64 MFMA builtin (16x16x4) on gfx950, read back in bounds.

One-hot probes recover, for every lane, which A entry (row i, depth k) and which B
entry (depth k, column j) it feeds and which C entries its four accumulators hold;
the Newton–Schulz tiles of the island kernel are written against the layout the
last test asserts."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libmfma_probe.so")


def run(av, bv):
    lib = C.CDLL(SO)
    lib.mfma_probe_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    av = np.ascontiguousarray(av, dtype=np.float64).reshape(-1, 64)
    bv = np.ascontiguousarray(bv, dtype=np.float64).reshape(-1, 64)
    out = np.zeros((av.shape[0], 64, 4))
    assert lib.mfma_probe_run(av.ctypes.data, bv.ctypes.data, out.ctypes.data, av.shape[0]) == 0
    return out


def discover():
    """Returns (a_of_lane, b_of_lane, c_of_slot) as index maps, up to a relabelling
    of rows, columns and depth that is consistent across A, B and C."""
    eye = np.eye(64)
    ones = np.ones((64, 64))
    ramp = np.tile(np.arange(1, 65, dtype=np.float64), (64, 1))
    # A one-hot at lane p, B = lane index + 1: C slots of A's row get B[k_p][j] = lane(k_p, j) + 1
    out_a = run(eye, ramp)
    # B one-hot at lane q, A = 1: C slots of B's column get 1 per k match (exactly one)
    out_b = run(ones, eye)
    rows, cols = {}, {}
    for p in range(64):
        rows[p] = frozenset(zip(*np.nonzero(out_a[p])))
    for q in range(64):
        cols[q] = frozenset(zip(*np.nonzero(out_b[q])))
    row_classes = sorted(set(rows.values()), key=lambda s: min(s))
    col_classes = sorted(set(cols.values()), key=lambda s: min(s))
    assert len(row_classes) == 16 and len(col_classes) == 16
    ri = {s: i for i, s in enumerate(row_classes)}
    cj = {s: j for j, s in enumerate(col_classes)}
    c_of_slot = {}
    for s in row_classes:
        for slot in s:
            c_of_slot[slot] = [ri[s], None]
    for s in col_classes:
        for slot in s:
            c_of_slot[slot][1] = cj[s]
    # depth: A lanes of one row pair with B lanes of one column through k
    a_of_lane, b_of_lane = {}, {}
    for q in range(64):
        b_of_lane[q] = [None, cj[cols[q]]]
    depth = {}
    for i, s in enumerate(row_classes):
        lanes = [p for p in range(64) if rows[p] == s]
        assert len(lanes) == 4
        for p in lanes:
            key = frozenset(int(round(out_a[p][t])) - 1 for t in s)
            depth.setdefault(key, len(depth))
            a_of_lane[p] = [i, depth[key]]
            for t in s:
                b_of_lane[int(round(out_a[p][t])) - 1][0] = depth[key]
    return a_of_lane, b_of_lane, c_of_slot


def test_layout_is_a_permutation_and_multiplies():
    a_of, b_of, c_of = discover()
    assert sorted(map(tuple, a_of.values())) == [(i, k) for i in range(16) for k in range(4)]
    assert sorted(map(tuple, b_of.values())) == [(k, j) for k in range(4) for j in range(16)]
    assert sorted(map(tuple, c_of.values())) == [(i, j) for i in range(16) for j in range(16)]
    rng = np.random.default_rng(1)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    av = np.array([A[tuple(a_of[l])] for l in range(64)])
    bv = np.array([B[tuple(b_of[l])] for l in range(64)])
    out = run(av, bv)[0]
    ref = A @ B
    for (l, j), (r, c) in c_of.items():
        assert abs(out[l, j] - ref[r, c]) <= 1e-14 * (1 + abs(ref[r, c]))


def test_documented_layout():
    """lane l feeds A[l%16][l/16] and B[l/16][l%16]; accumulator j of lane l holds
    C[l/16 + 4j][l%16] (measured on gfx950: the four accumulators of a lane are rows
    4 apart, not 4 consecutive rows)."""
    a_of, b_of, c_of = discover()
    # relabel so A's row of lane l is l%16 and its depth l/16
    row_map = {a_of[l][0]: l % 16 for l in range(64)}
    dep_map = {a_of[l][1]: l // 16 for l in range(64)}
    col_map = {b_of[l][1]: l % 16 for l in range(64)}
    for l in range(64):
        assert (row_map[a_of[l][0]], dep_map[a_of[l][1]]) == (l % 16, l // 16)
        assert (dep_map[b_of[l][0]], col_map[b_of[l][1]]) == (l // 16, l % 16), l
    table = {(l, j): (row_map[r], col_map[c]) for (l, j), (r, c) in c_of.items()}
    print("C layout:", " ".join(f"{l}.{j}={r},{c}" for (l, j), (r, c) in sorted(table.items())))
    for l in range(64):
        for j in range(4):
            assert table[(l, j)] == (l // 16 + 4 * j, l % 16), (l, j, table[(l, j)])
