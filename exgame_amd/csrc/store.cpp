// Typed slot storage with generational keys in HIP device buffers, and the
// row-driven dispatch over it — the device half of aristotle::World
// (crates/aristotle/src/world.rs:17-66, 300-385) plus Accelerator's per-kind
// dispatch and one-writer Ledger (crates/newton/src/accelerator/shaders.rs:287-431).
//
// Keys are bit-exact with RawMap: insert pops `first_free` or appends (growing
// the buffer ×2, copying every slot to the same index and bumping the storage
// generation); remove checks the generation, splices the slot into the
// address-ordered free list and bumps the slot generation (wrapping).
//
// Before a dispatch every slot each row names is checked against its storage's
// high-water mark on the host (a bad row is refused, never launched), and the
// output slots go through a ledger: two writers of one slot refuse the whole
// dispatch, as the reference refuses a flush.
#include <hip/hip_runtime.h>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>
#include "../../include/exgame.h"
#include "island.h"
#include "rows.h"

namespace {

int sfail(int code, const std::string& msg) { return exg::set_error(code, msg.c_str()); }

// scalars per element of each storage (the ROW storage holds 128 u32 words)
constexpr size_t ELEM_SCALARS[exg::N_STORE_TYPES] = {1, 8, 6, 6, 12, 288, 36, 3, 9, 0};

size_t elem_bytes(int type, int scalar_bytes) {
    return type == exg::T_ROW ? exg::ROW_WORDS * 4 : ELEM_SCALARS[type] * (size_t)scalar_bytes;
}

struct SlotKey {
    uint32_t is_value, next, generation;
};

struct TypedMap {
    std::vector<SlotKey> keys;
    uint32_t first_free = 0;
    uint64_t capacity = 1;
    uint64_t generation = 0;
    void* dev = nullptr;
    size_t elem = 8;
    std::vector<uint32_t> host_rows;  // the ROW storage keeps a host mirror (rows are baked on the host)
};

}  // namespace

struct exg_store {
    TypedMap maps[exg::N_STORE_TYPES];
    int scalar_bytes = 8;                 // 4: f32 storages, 8: f64
    int codegen = EXG_CODEGEN_HANDWRITTEN;  // joint and Pre rows: hand-written or traced kernels
    hipStream_t stream = nullptr;
    uint32_t* dev_slots = nullptr;
    size_t dev_slots_cap = 0;
    unsigned* dev_fatal = nullptr;        // rows of the running dispatch with non-finite outputs
    uint32_t fatals[exg::N_KINDS] = {};   // that count for the last dispatch of each kind
};

namespace {

hipError_t grow(exg_store* s, int type, uint64_t min) {
    TypedMap& m = s->maps[type];
    uint64_t cap = m.capacity * 2;
    if (min > cap) cap = min;
    void* fresh = nullptr;
    hipError_t e = hipMalloc(&fresh, cap * m.elem);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(fresh, 0, cap * m.elem, s->stream);
    if (e == hipSuccess && !m.keys.empty())
        e = hipMemcpyAsync(fresh, m.dev, m.keys.size() * m.elem, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        (void)hipFree(fresh);
        return e;
    }
    (void)hipFree(m.dev);
    m.dev = fresh;
    m.capacity = cap;
    m.generation += 1;
    if (type == exg::T_ROW) m.host_rows.resize(cap * exg::ROW_WORDS, 0u);
    return hipSuccess;
}

bool live(const TypedMap& m, uint32_t index, uint32_t generation) {
    return index < m.keys.size() && m.keys[index].is_value == 1 && m.keys[index].generation == generation;
}

// (storage type, slot, is output) for every slot a row of `kind` names.
struct SlotRef {
    int type;
    uint32_t slot;
    bool out;
};

int joint_params_of(int kind) {
    switch (kind) {
        case exg::K_CDW_PLAIN: case exg::K_CDW_JAC: case exg::K_SSD_PLAIN: case exg::K_SSD_JAC: return 10;
        case exg::K_PERP_PLAIN: case exg::K_PERP_JAC: return 2;
        default: return 1;
    }
}

// Decode a row; false when its own counts are out of range.
bool row_refs(int kind, const uint32_t* w, std::vector<SlotRef>& out) {
    using namespace exg;
    out.clear();
    auto add = [&](int t, uint32_t s, bool o) { out.push_back({t, s, o}); };
    switch (kind) {
        case K_SIMPLE_SUM: add(T_SCALAR, w[0], false); add(T_SCALAR, w[1], false); add(T_SCALAR, w[2], true); return true;
        case K_PRE:
            add(T_TWIST, w[0], false); add(T_MOTOR, w[1], false); add(T_MOTOR, w[2], true); add(T_TWIST, w[3], true);
            add(T_SCALAR, w[4], false);
            return true;
        case K_CDW_PLAIN: case K_PERP_PLAIN: case K_SSD_PLAIN: case K_TORS_PLAIN: {
            add(T_TWIST, w[0], false); add(T_TWIST, w[1], false); add(T_MOTOR, w[2], false); add(T_MOTOR, w[3], false);
            add(T_WRENCH_PAIR, w[4], true); add(T_SCALAR, w[5], false); add(T_SCALAR, w[6], false);
            for (int p = 0; p < joint_params_of(kind); ++p) add(T_SCALAR, w[7 + p], false);
            return true;
        }
        case K_CDW_JAC: case K_PERP_JAC: case K_SSD_JAC: case K_TORS_JAC: {
            add(T_TWIST, w[0], false); add(T_TWIST, w[1], false); add(T_MOTOR, w[2], false); add(T_MOTOR, w[3], false);
            add(T_JACOBIAN, w[4], true); add(T_WRENCH_PAIR, w[5], true); add(T_SCALAR, w[6], false);
            add(T_SCALAR, w[7], false);
            for (int p = 0; p < joint_params_of(kind); ++p) add(T_SCALAR, w[8 + p], false);
            return true;
        }
        case K_GATHER: {
            if (w[2] > 1u || w[3] > (uint32_t)(ROW_WORDS - 4)) return false;
            add(T_WRENCH, w[0], true); add(T_WRENCH, w[1], false);
            for (uint32_t q = 0; q < w[3]; ++q) add(T_WRENCH_PAIR, w[4 + q] >> 1, false);
            return true;
        }
        case K_GEMM: {
            if (w[1] > 1u || w[2] > (uint32_t)((ROW_WORDS - 4) / 2)) return false;
            add(T_BLOCK, w[0], true);
            for (uint32_t q = 0; q < w[2]; ++q) { add(T_BLOCK, w[4 + 2 * q], false); add(T_BLOCK, w[5 + 2 * q] >> 1, false); }
            return true;
        }
        case K_BODY_POST_DIAG: case K_BODY_POST_FULL: {
            if (w[0] > 13u) return false;
            add(T_SCALAR, w[1], false); add(T_SCALAR, w[2], false);
            int ang = kind == K_BODY_POST_DIAG ? T_DIAG_INERTIA : T_FULL_INERTIA;
            for (uint32_t b = 0; b < w[0]; ++b) {
                const uint32_t* o = w + 3 + 9 * b;
                add(T_MOTOR, o[0], false); add(T_TWIST, o[1], false); add(T_SCALAR, o[2], false); add(ang, o[3], false);
                add(T_WRENCH, o[4], false); add(T_WRENCH, o[5], false); add(T_BLOCK, o[6], true);
                add(T_WRENCH, o[7], true); add(T_WRENCH, o[8], true);
            }
            return true;
        }
        case K_BODY_POST_GATHERED_DIAG: case K_BODY_POST_GATHERED_FULL: {
            if (w[11] > (uint32_t)(ROW_WORDS - 12)) return false;
            int ang = kind == K_BODY_POST_GATHERED_DIAG ? T_DIAG_INERTIA : T_FULL_INERTIA;
            add(T_SCALAR, w[0], false); add(T_SCALAR, w[1], false); add(T_MOTOR, w[2], false); add(T_TWIST, w[3], false);
            add(T_SCALAR, w[4], false); add(ang, w[5], false); add(T_WRENCH, w[6], false); add(T_WRENCH, w[7], false);
            add(T_BLOCK, w[8], true); add(T_WRENCH, w[9], true); add(T_WRENCH, w[10], true);
            for (uint32_t q = 0; q < w[11]; ++q) add(T_WRENCH_PAIR, w[12 + q] >> 1, false);
            return true;
        }
        case K_ASSEMBLE: {
            if (w[5] > 1u || w[6] > (uint32_t)(ROW_WORDS - 7)) return false;
            add(T_BLOCK, w[0], true); add(T_BLOCK, w[1], false); add(T_SCALAR, w[3], false); add(T_SCALAR, w[4], false);
            for (uint32_t q = 0; q < w[6]; ++q) add(T_JACOBIAN, w[7 + q] >> 2, false);
            return true;
        }
        case K_MATVEC: {
            if (w[1] > 1u || w[2] > (uint32_t)((ROW_WORDS - 3) / 2)) return false;
            add(T_TWIST, w[0], true);
            for (uint32_t q = 0; q < w[2]; ++q) { add(T_BLOCK, w[3 + 2 * q], false); add(T_WRENCH, w[4 + 2 * q], false); }
            return true;
        }
        case K_REDUCE: {
            if (w[1] > (uint32_t)(ROW_WORDS - 2)) return false;
            add(T_SCALAR, w[0], true);
            for (uint32_t q = 0; q < w[1]; ++q) add(T_BLOCK, w[2 + q] >> 1, false);
            return true;
        }
        case K_COPY: add(T_BLOCK, w[0], true); add(T_BLOCK, w[1], false); return true;
        default: return false;
    }
}

}  // namespace

extern "C" {

int exg_store_create_scalar(uint64_t capacity, int scalar, exg_store** out) {
    if (!out) return sfail(EXG_ERR_INVALID, "exg_store_create: null output");
    *out = nullptr;
    if (capacity < 1) return sfail(EXG_ERR_INVALID, "exg_store_create: world storage capacity must be at least 1");
    if (scalar != EXG_SCALAR_F32 && scalar != EXG_SCALAR_F64)
        return sfail(EXG_ERR_INVALID, "exg_store_create: storage scalar must be EXG_SCALAR_F32 or EXG_SCALAR_F64");
    auto* s = new exg_store();
    s->scalar_bytes = scalar == EXG_SCALAR_F32 ? 4 : 8;
    hipError_t e = hipStreamCreate(&s->stream);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&s->dev_fatal), sizeof(unsigned));
    for (int t = 0; t < exg::N_STORE_TYPES && e == hipSuccess; ++t) {
        TypedMap& m = s->maps[t];
        m.elem = elem_bytes(t, s->scalar_bytes);
        m.capacity = capacity;
        e = hipMalloc(&m.dev, capacity * m.elem);
        if (e == hipSuccess) e = hipMemset(m.dev, 0, capacity * m.elem);
        if (t == exg::T_ROW) m.host_rows.assign(capacity * exg::ROW_WORDS, 0u);
    }
    if (e != hipSuccess) {
        for (auto& m : s->maps) if (m.dev) (void)hipFree(m.dev);
        if (s->dev_fatal) (void)hipFree(s->dev_fatal);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return sfail(EXG_ERR_BACKEND, std::string("exg_store_create: ") + hipGetErrorString(e));
    }
    *out = s;
    return EXG_OK;
}

int exg_store_create(uint64_t capacity, exg_store** out) { return exg_store_create_scalar(capacity, EXG_SCALAR_F64, out); }

int exg_store_set_codegen(exg_store* s, int codegen) {
    if (!s) return sfail(EXG_ERR_INVALID, "exg_store_set_codegen: null store");
    if (codegen != EXG_CODEGEN_HANDWRITTEN && codegen != EXG_CODEGEN_TRACED)
        return sfail(EXG_ERR_INVALID, "exg_store_set_codegen: unknown mode " + std::to_string(codegen));
    s->codegen = codegen;
    return EXG_OK;
}

int exg_store_scalar(exg_store* s, int* scalar) {
    if (!s || !scalar) return sfail(EXG_ERR_INVALID, "exg_store_scalar: null argument");
    *scalar = s->scalar_bytes == 4 ? EXG_SCALAR_F32 : EXG_SCALAR_F64;
    return EXG_OK;
}

int exg_store_destroy(exg_store* s) {
    if (!s) return EXG_OK;
    (void)hipStreamSynchronize(s->stream);
    for (auto& m : s->maps) if (m.dev) (void)hipFree(m.dev);
    if (s->dev_slots) (void)hipFree(s->dev_slots);
    if (s->dev_fatal) (void)hipFree(s->dev_fatal);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return EXG_OK;
}

int exg_slot_insert(exg_store* s, int type, const void* value, uint32_t* index, uint32_t* generation) {
    if (!s || type < 0 || type >= exg::N_STORE_TYPES || !value || !index || !generation)
        return sfail(EXG_ERR_INVALID, "exg_slot_insert: bad arguments");
    TypedMap& m = s->maps[type];
    uint32_t position = m.first_free;
    uint32_t gen = 0;
    if (position >= m.keys.size()) {
        if (position >= m.capacity) {
            hipError_t e = grow(s, type, (uint64_t)position + 1);
            if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_slot_insert: ") + hipGetErrorString(e));
        }
        m.first_free = position + 1;
        m.keys.push_back({1, 0, 0});
    } else {
        SlotKey k = m.keys[position];
        if (k.is_value) return sfail(EXG_ERR_BACKEND, "exg_slot_insert: free list points to a non-free element");
        m.first_free = k.next;
        gen = k.generation;
        m.keys[position] = {1, 0, gen};
    }
    hipError_t e = hipMemcpy(static_cast<char*>(m.dev) + (size_t)position * m.elem, value, m.elem, hipMemcpyHostToDevice);
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_slot_insert: ") + hipGetErrorString(e));
    if (type == exg::T_ROW) std::memcpy(&m.host_rows[(size_t)position * exg::ROW_WORDS], value, m.elem);
    *index = position;
    *generation = gen;
    return EXG_OK;
}

int exg_slot_read(exg_store* s, int type, uint32_t index, uint32_t generation, void* out, int* found) {
    if (!s || type < 0 || type >= exg::N_STORE_TYPES || !out || !found)
        return sfail(EXG_ERR_INVALID, "exg_slot_read: bad arguments");
    TypedMap& m = s->maps[type];
    *found = live(m, index, generation) ? 1 : 0;
    if (!*found) return EXG_OK;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess)
        e = hipMemcpy(out, static_cast<char*>(m.dev) + (size_t)index * m.elem, m.elem, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_slot_read: ") + hipGetErrorString(e));
    return EXG_OK;
}

int exg_slot_write(exg_store* s, int type, uint32_t index, uint32_t generation, const void* value, int* found) {
    if (!s || type < 0 || type >= exg::N_STORE_TYPES || !value || !found)
        return sfail(EXG_ERR_INVALID, "exg_slot_write: bad arguments");
    TypedMap& m = s->maps[type];
    *found = live(m, index, generation) ? 1 : 0;
    if (!*found) return EXG_OK;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess)
        e = hipMemcpy(static_cast<char*>(m.dev) + (size_t)index * m.elem, value, m.elem, hipMemcpyHostToDevice);
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_slot_write: ") + hipGetErrorString(e));
    if (type == exg::T_ROW) std::memcpy(&m.host_rows[(size_t)index * exg::ROW_WORDS], value, m.elem);
    return EXG_OK;
}

int exg_slot_remove(exg_store* s, int type, uint32_t index, uint32_t generation, void* value_out, int* found) {
    if (!s || type < 0 || type >= exg::N_STORE_TYPES || !found)
        return sfail(EXG_ERR_INVALID, "exg_slot_remove: bad arguments");
    TypedMap& m = s->maps[type];
    *found = live(m, index, generation) ? 1 : 0;
    if (!*found) return EXG_OK;
    if (value_out) {
        int f2 = 0;
        int rc = exg_slot_read(s, type, index, generation, value_out, &f2);
        if (rc != EXG_OK) return rc;
    }
    uint32_t it = m.first_free;
    bool has_prev = false;
    uint32_t prev = 0;
    while (true) {
        if (index < it) {
            m.keys[index] = {0, it, generation + 1u};
            if (!has_prev) m.first_free = index;
            else m.keys[prev] = {0, index, m.keys[prev].generation};
            break;
        }
        if (it >= m.keys.size() || m.keys[it].is_value)
            return sfail(EXG_ERR_BACKEND, "exg_slot_remove: non-free element in the free list");
        prev = it;
        has_prev = true;
        it = m.keys[it].next;
    }
    return EXG_OK;
}

int exg_storage_info(exg_store* s, int type, uint64_t* capacity, uint64_t* len, uint64_t* generation) {
    if (!s || type < 0 || type >= exg::N_STORE_TYPES) return sfail(EXG_ERR_INVALID, "exg_storage_info: bad arguments");
    const TypedMap& m = s->maps[type];
    if (capacity) *capacity = m.capacity;
    if (len) *len = m.keys.size();
    if (generation) *generation = m.generation;
    return EXG_OK;
}

int exg_storage_device_ptr(exg_store* s, int type, void** ptr) {
    if (!s || type < 0 || type >= exg::N_STORE_TYPES || !ptr) return sfail(EXG_ERR_INVALID, "exg_storage_device_ptr: bad arguments");
    *ptr = s->maps[type].dev;
    return EXG_OK;
}

int exg_dispatch(exg_store* s, int kind, const uint32_t* row_slots, uint32_t count) {
    if (!s || (!row_slots && count)) return sfail(EXG_ERR_INVALID, "exg_dispatch: bad arguments");
    if (kind < 0 || kind >= exg::N_KINDS) return sfail(EXG_ERR_NO_SHADER, "exg_dispatch: no kernel for kind " + std::to_string(kind));
    if (count == 0) return EXG_OK;
    const TypedMap& rows = s->maps[exg::T_ROW];
    std::vector<SlotRef> refs;
    std::unordered_set<uint64_t> ledger;
    for (uint32_t r = 0; r < count; ++r) {
        uint32_t rs = row_slots[r];
        if (rs >= rows.keys.size() || !rows.keys[rs].is_value)
            return sfail(EXG_ERR_INVALID, "exg_dispatch: row " + std::to_string(r) + " names no live row slot");
        const uint32_t* w = &rows.host_rows[(size_t)rs * exg::ROW_WORDS];
        if (!row_refs(kind, w, refs))
            return sfail(EXG_ERR_INVALID, "exg_dispatch: row " + std::to_string(r) + " has counts out of range");
        for (const SlotRef& ref : refs) {
            if (ref.slot >= s->maps[ref.type].keys.size())
                return sfail(EXG_ERR_INVALID, "exg_dispatch: row " + std::to_string(r) + " names slot " +
                                                  std::to_string(ref.slot) + " past its storage");
            if (ref.out && !ledger.insert(((uint64_t)ref.type << 32) | ref.slot).second)
                return sfail(EXG_ERR_BACKEND, "exg_dispatch: ledger refuses two writers for slot " +
                                                  std::to_string(ref.slot) + " of storage " + std::to_string(ref.type));
        }
    }
    if (s->dev_slots_cap < count) {
        if (s->dev_slots) (void)hipFree(s->dev_slots);
        s->dev_slots = nullptr;
        size_t cap = count < 1024 ? 1024 : count;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&s->dev_slots), cap * sizeof(uint32_t));
        if (e != hipSuccess) {
            s->dev_slots_cap = 0;
            return sfail(EXG_ERR_BACKEND, std::string("exg_dispatch: ") + hipGetErrorString(e));
        }
        s->dev_slots_cap = cap;
    }
    hipError_t e = hipMemcpyAsync(s->dev_slots, row_slots, count * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->dev_fatal, 0, sizeof(unsigned), s->stream);
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_dispatch: ") + hipGetErrorString(e));
    exg::StorePtrs p;
    p.scalar = s->maps[exg::T_SCALAR].dev;
    p.motor = s->maps[exg::T_MOTOR].dev;
    p.twist = s->maps[exg::T_TWIST].dev;
    p.wrench = s->maps[exg::T_WRENCH].dev;
    p.pair = s->maps[exg::T_WRENCH_PAIR].dev;
    p.jac = s->maps[exg::T_JACOBIAN].dev;
    p.block = s->maps[exg::T_BLOCK].dev;
    p.diag = s->maps[exg::T_DIAG_INERTIA].dev;
    p.full = s->maps[exg::T_FULL_INERTIA].dev;
    p.row = static_cast<const uint32_t*>(s->maps[exg::T_ROW].dev);
    p.fatal = s->dev_fatal;
    p.scalar_bytes = s->scalar_bytes;
    if (!(s->codegen == EXG_CODEGEN_TRACED && exg::launch_rows_traced(kind, p, s->dev_slots, count, s->stream)))
        exg::launch_rows(kind, p, s->dev_slots, count, s->stream);
    e = hipGetLastError();
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_dispatch: ") + hipGetErrorString(e));
    // the slot list is staged in one buffer: wait so the next dispatch can reuse it
    e = hipStreamSynchronize(s->stream);
    unsigned fatal = 0;
    if (e == hipSuccess) e = hipMemcpy(&fatal, s->dev_fatal, sizeof(unsigned), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_dispatch: ") + hipGetErrorString(e));
    s->fatals[kind] = fatal;
    if (fatal)   // mark-and-continue: the outputs are written; the message fails
        return sfail(EXG_ERR_FATAL, "exg_dispatch: " + std::to_string(fatal) + " row(s) of kind " +
                                        std::to_string(kind) + " produced non-finite outputs");
    return EXG_OK;
}

int exg_flush(exg_store* s) {
    if (!s) return sfail(EXG_ERR_INVALID, "exg_flush: null store");
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return sfail(EXG_ERR_BACKEND, std::string("exg_flush: ") + hipGetErrorString(e));
    return EXG_OK;
}

int exg_fatals(exg_store* s, int kind, uint32_t* count) {
    if (!s || !count || kind < 0 || kind >= exg::N_KINDS) return sfail(EXG_ERR_INVALID, "exg_fatals: bad arguments");
    *count = s->fatals[kind];
    return EXG_OK;
}

}  // extern "C"
