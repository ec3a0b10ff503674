// ORACLE — test infrastructure only. Restatement of the reference's generational
// slot map `aristotle::RawMap` (crates/aristotle/src/world.rs:17-66, 300-385):
//   insert  pop `first_free` (or append, growing capacity ×2 with a generation bump
//           of the storage) and return {index, slot generation}
//   remove  check the key's generation, splice the slot into the ADDRESS-ORDERED
//           free list, bump the slot generation (wrapping)
// Only the key/index arithmetic is restated; payloads are u64 here.
#pragma once
#include <cstdint>
#include <optional>
#include <vector>

namespace oracle {

struct RawMapKey { uint32_t is_value; uint32_t next; uint32_t generation; };
struct MapKey { uint32_t value; uint32_t generation; };

struct RawMap {
    std::vector<RawMapKey> map;
    std::vector<uint64_t> values;
    uint64_t capacity;
    uint64_t storage_generation = 0;
    uint32_t first_free = 0;

    explicit RawMap(uint64_t cap) : capacity(cap < 1 ? 1 : cap) { values.resize(capacity); }

    void grow(uint64_t min) {  // world.rs:176-200
        uint64_t c = capacity * 2;
        capacity = min > c ? min : c;
        values.resize(capacity);
        storage_generation++;
    }

    MapKey insert(uint64_t value) {  // world.rs:300-331
        uint32_t position = first_free;
        if (position >= map.size()) {
            if (position >= capacity) grow((uint64_t)position + 1);
            first_free = position + 1;
            map.push_back({1, 0, 0});
            values[position] = value;
            return {position, 0};
        }
        RawMapKey k = map[position];
        // the free list only ever points at free entries
        uint32_t next = k.next, gen = k.generation;
        first_free = next;
        map[position] = {1, 0, gen};
        values[position] = value;
        return {position, gen};
    }

    std::optional<uint64_t> get(MapKey key) const {  // world.rs:333-342
        if (key.value >= map.size()) return std::nullopt;
        const RawMapKey& k = map[key.value];
        if (k.is_value != 1 || k.generation != key.generation) return std::nullopt;
        return values[key.value];
    }

    bool set(MapKey key, uint64_t value) {
        if (!get(key)) return false;
        values[key.value] = value;
        return true;
    }

    std::optional<uint64_t> remove(MapKey key) {  // world.rs:353-385
        auto out = get(key);
        if (!out) return std::nullopt;
        uint32_t position = key.value;
        uint32_t it = first_free;
        bool has_prev = false;
        uint32_t prev = 0;
        while (true) {
            if (position < it) {
                uint32_t new_gen = key.generation + 1u;  // wrapping_add
                map[position] = {0, it, new_gen};
                if (!has_prev) first_free = position;
                else {
                    uint32_t g = map[prev].generation;
                    map[prev] = {0, position, g};
                }
                break;
            }
            prev = it;
            has_prev = true;
            it = map[it].next;
        }
        return out;
    }
};

}  // namespace oracle
