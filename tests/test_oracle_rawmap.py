"""Pins the oracle's restatement of aristotle::RawMap to the reference's own world
tests (crates/aristotle/src/world.rs:694-909, crates/aristotle/tests/ar0101_world_growth.rs)."""
from pyoracle import OracleRawMap


def test_insert_get_remove():
    m = OracleRawMap(100)
    k1 = m.insert(42)
    k2 = m.insert(27)
    assert m.get(*k1) == 42 and m.get(*k2) == 27
    assert m.remove(*k1) == 42 and m.get(*k2) == 27 and m.remove(*k2) == 27


def test_stale_key_and_slot_reuse_with_new_generation():
    m = OracleRawMap(100)
    k = m.insert(100)
    assert m.remove(*k) == 100
    assert m.get(*k) is None
    k2 = m.insert(200)
    assert k2[0] == k[0] and k2[1] != k[1]
    assert m.get(*k) is None and m.get(*k2) == 200


def test_remove_twice_is_none():
    m = OracleRawMap(100)
    k = m.insert(5)
    assert m.remove(*k) == 5 and m.remove(*k) is None


def test_free_list_reuses_lowest_slot_first():
    m = OracleRawMap(100)
    k = [m.insert(v) for v in range(4)]
    m.remove(*k[1])
    m.remove(*k[3])
    m.remove(*k[0])
    assert m.insert(10)[0] == k[0][0]
    assert m.insert(11)[0] == k[1][0]
    assert m.insert(13)[0] == k[3][0]
    assert m.get(*k[2]) == 2
    assert m.insert(14)[0] == 4


def test_high_water_mark_and_growth():
    m = OracleRawMap(100)
    k0 = m.insert(11)
    m.insert(22)
    assert m.info()["len"] == 2
    m.remove(*k0)
    assert m.info()["len"] == 2
    m.insert(33)
    assert m.info()["len"] == 2
    g = OracleRawMap(2)
    keys = [g.insert(i * 10) for i in range(9)]
    info = g.info()
    assert info["capacity"] >= 9 and info["generation"] == 3      # 2 → 4 → 8 → 16
    for i, k in enumerate(keys):
        assert k == (i, 0) and g.get(*k) == i * 10


def test_reuse_does_not_grow_and_generations_bump_per_remove():
    m = OracleRawMap(2)
    k0 = m.insert(1)
    m.insert(2)
    m.remove(*k0)
    k = m.insert(3)
    assert m.info()["capacity"] == 2 and m.info()["generation"] == 0
    assert k == (0, 1)
    m.remove(*k)
    assert m.insert(4) == (0, 2)


def test_growth_keeps_indices_past_initial_capacity():
    m = OracleRawMap(8)
    keys = [m.insert(i * 3 + 1) for i in range(8 * 5 + 3)]
    for i, k in enumerate(keys):
        assert k[0] == i and m.get(*k) == i * 3 + 1
    assert m.info()["generation"] > 0
