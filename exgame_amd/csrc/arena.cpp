// Many independent worlds stepped together on one device: the batched and served
// forms of the per-step hot path.
//
// The reference steps several mechanisms of one Driver epoch against one
// accelerator, flushing their rows together (crates/newton/src/accelerator/mod.rs:161-277,
// 613-630; crates/newton/src/driver.rs). Here every admitted world lives in one
// device pool — bodies, islands, joints and accepted inverses are ranges of shared
// HBM arrays handed out first-fit — and one step is one gravity pass over a work
// table of all worlds (each world's sources stay its own) plus one launch per
// island class over the island ids of all worlds. Nothing is exchanged between
// worlds, and every per-world quantity (island kernel, solver budgets, and the
// gravity chunking, fixed when the world is admitted) depends only on that world and
// the arena's settings at its admission, so a world's trajectory is bit-identical
// alone or in any mix.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>
#include "../../include/exgame.h"
#include "devbuf.h"
#include "island.h"

using exg::DevBuf;

namespace {

int afail(int code, const std::string& msg) { return exg::set_error(code, msg.c_str()); }

#define A_TRY(expr)                                                                                      \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return afail(EXG_ERR_BACKEND, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// First-fit allocator of contiguous index ranges [0, capacity), coalescing on free.
struct RangeAlloc {
    std::vector<std::pair<long long, long long>> free_;  // (start, length), sorted by start
    long long capacity = 0, used = 0;
    void init(long long cap) {
        capacity = cap;
        used = 0;
        free_.clear();
        if (cap > 0) free_.push_back({0, cap});
    }
    long long take(long long n) {
        if (n == 0) return 0;
        for (size_t k = 0; k < free_.size(); ++k) {
            if (free_[k].second >= n) {
                long long s = free_[k].first;
                free_[k].first += n;
                free_[k].second -= n;
                if (free_[k].second == 0) free_.erase(free_.begin() + (long)k);
                used += n;
                return s;
            }
        }
        return -1;
    }
    bool fits(long long n) const {
        if (n == 0) return true;
        for (auto& f : free_)
            if (f.second >= n) return true;
        return false;
    }
    void give(long long s, long long n) {
        if (n == 0) return;
        used -= n;
        auto it = std::lower_bound(free_.begin(), free_.end(), std::make_pair(s, 0LL));
        it = free_.insert(it, {s, n});
        size_t k = (size_t)(it - free_.begin());
        if (k + 1 < free_.size() && free_[k].first + free_[k].second == free_[k + 1].first) {
            free_[k].second += free_[k + 1].second;
            free_.erase(free_.begin() + (long)k + 1);
        }
        if (k > 0 && free_[k - 1].first + free_[k - 1].second == free_[k].first) {
            free_[k - 1].second += free_[k].second;
            free_.erase(free_.begin() + (long)k);
        }
    }
};

struct ArenaWorld {
    bool live = false;
    long long b0 = 0, nb = 0, i0 = 0, ni = 0, j0 = 0, nj = 0, x0 = 0, nx = 0;
    double G = 1.0;
    int gravity_on = 1;
    int chunk = 0;   // sources per gravity chunk, fixed at admission (it sets the order of the force sums)
    int mapping = EXG_JOINT_MAPPING_SPLIT;   // joint mapping of its 8-body islands, fixed at admission
};

}  // namespace

struct exg_arena {
    int Bc = 0, Ic = 0, Jc = 0;
    long long Xc = 0;
    int cus = 256;
    int grav_chunk = 0;   // chunk length of worlds admitted from now on; 0: the single-world rule
    double tol2 = 1e-8;
    int ns_warm = 2, ns_seed = 24, ns_engine = EXG_NS_MFMA, grav_prec = EXG_PREC_F64, jac_prec = EXG_PREC_F64;
    int joint_mapping = EXG_JOINT_MAPPING_SPLIT;
    hipStream_t stream = nullptr;
    exg::ClassStreams classes;     // the island classes run side by side
    DevBuf<double> pose, mom, mass, tens, size, ext, origin, jparams, xacc, charges[2], partial, island_G, grav_stage;
    DevBuf<int> island_start, island_end, island_jstart, island_jend, jkind, ja, jb, seeded, stats;
    DevBuf<int> fatal;                // island-steps that ended with a non-finite state
    DevBuf<int> requeue;              // single-body islands whose step subdivides (capacity: every island)
    DevBuf<unsigned> requeue_ctl;     // the requeue list's length and workgroup counter (0 between launches)
    long long fatal_reported = 0;
    DevBuf<long long> xacc_off;
    DevBuf<exg::IslandTopo> topo;
    DevBuf<exg::GravEntry> entries;
    DevBuf<exg::GravTile> tiles;
    size_t n_entries = 0, n_tiles = 0;
    int stage_n = 0;   // bodies the mixed-precision source stage covers: up to the last live world's end
    DevBuf<int> class_ids[exg::ISLAND_CLASSES];
    int class_n[exg::ISLAND_CLASSES] = {};
    std::vector<int> island_class_of;   // per island slot
    RangeAlloc bodies, islands, joints, xaccs;
    std::vector<ArenaWorld> worlds;
    bool dirty = true;
    int front = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    static constexpr int N_MARKS = 8;
    hipEvent_t marks[N_MARKS] = {};   // completion marks of exg_arena_mark, a ring
    uint64_t marks_issued = 0;
    bool use_graph = false;
    int settings_epoch = 0;   // bumped by anything baked into captured kernel arguments
    struct Graph {
        hipGraphExec_t exec = nullptr;
        double dt = 0.0, warp = 0.0;
        int settings = -1;
    } graphs[2];              // by the charge buffer the two-step graph starts from

    void release() {
        for (auto* b : {&pose, &mom, &mass, &tens, &size, &ext, &origin, &jparams, &xacc, &charges[0], &charges[1],
                        &partial, &island_G, &grav_stage})
            b->release();
        for (auto* b : {&island_start, &island_end, &island_jstart, &island_jend, &jkind, &ja, &jb, &seeded, &stats,
                        &fatal, &requeue})
            b->release();
        requeue_ctl.release();
        xacc_off.release();
        topo.release();
        entries.release();
        tiles.release();
        for (auto& c : class_ids) c.release();
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : marks)
            if (e) (void)hipEventDestroy(e);
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        classes.destroy();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // Sources per chunk for a world of n bodies under chunk setting `setting`: a fixed
    // length, or (0) what a stand-alone world of n bodies uses (exg_world_create).
    int chunk_for(long long n, int setting) const {
        if (setting > 0) return setting;
        std::vector<exg::GravEntry> en;
        std::vector<exg::GravTile> tl;
        exg::gravity_add_world(en, tl, 0, (int)n, 1.0, cus);
        return en.empty() ? (int)n : en[0].src1 - en[0].src0;
    }
    // Gravity work table and island-class id lists of the live worlds; with
    // `override_chunk` ≥ 0 every world is chunked by that setting instead of its own
    // (timing only: exg_arena_autotune_gravity).
    // Work tables and class lists of the live worlds; override_chunk / override_mapping
    // (≥ 0) replace every world's own setting (the autotuners' trials).
    int rebuild(int override_chunk = -1, int override_mapping = -1) {
        std::vector<exg::GravEntry> en;
        std::vector<exg::GravTile> tl;
        std::vector<int> ids[exg::ISLAND_CLASSES];
        long long hw = 0;
        for (auto& w : worlds) {
            if (!w.live) continue;
            hw = std::max(hw, w.b0 + w.nb);
            if (w.gravity_on) {
                const int ch = override_chunk >= 0 ? chunk_for(w.nb, override_chunk) : w.chunk;
                exg::gravity_add_world_chunked(en, tl, (int)w.b0, (int)w.nb, w.G, ch);
            } else {
                // tiles without entries: the reduce writes a zero wrench
                for (long long t = 0; t < w.nb; t += exg::GRAV_TILE) {
                    exg::GravTile z{};
                    z.tgt0 = (int)(w.b0 + t);
                    z.ntgt = (int)std::min<long long>(exg::GRAV_TILE, w.nb - t);
                    z.e0 = 0;
                    z.ne = 0;
                    z.g_inv = 1.0;
                    tl.push_back(z);
                }
            }
            const int mapping = override_mapping >= 0 ? override_mapping : w.mapping;
            for (long long i = w.i0; i < w.i0 + w.ni; ++i) {
                int cls = island_class_of[(size_t)i];
                if (cls == exg::ISLAND_CLASS_M8J14 && mapping == EXG_JOINT_MAPPING_FLAT) cls = exg::ISLAND_CLASS_M8J14_FLAT;
                ids[cls].push_back((int)i);
            }
        }
        if (en.size() > entries.n) {
            entries.release();
            partial.release();
            A_TRY(entries.alloc(en.size()));
            A_TRY(partial.alloc(en.size() * exg::GRAV_PARTIAL_PER_ENTRY));
        }
        if (tl.size() > tiles.n) {
            tiles.release();
            A_TRY(tiles.alloc(tl.size()));
        }
        A_TRY(entries.upload(en.data(), en.size()));
        A_TRY(tiles.upload(tl.data(), tl.size()));
        n_entries = en.size();
        stage_n = (int)hw;
        n_tiles = tl.size();
        for (int c = 0; c < exg::ISLAND_CLASSES; ++c) {
            class_n[c] = (int)ids[c].size();
            if (ids[c].empty()) continue;
            if (ids[c].size() > class_ids[c].n) {
                class_ids[c].release();
                A_TRY(class_ids[c].alloc(std::max<size_t>(ids[c].size(), 64)));
            }
            A_TRY(class_ids[c].upload(ids[c].data(), ids[c].size()));
        }
        dirty = false;
        settings_epoch++;
        return EXG_OK;
    }
};

extern "C" {

int exg_arena_create(int body_capacity, int island_capacity, int joint_capacity, long long inverse_capacity,
                     exg_arena** out) {
    if (!out) return afail(EXG_ERR_INVALID, "exg_arena_create: null output");
    *out = nullptr;
    if (body_capacity <= 0 || island_capacity <= 0 || joint_capacity < 0 || inverse_capacity <= 0)
        return afail(EXG_ERR_INVALID, "exg_arena_create: capacities must be positive");
    auto* r = new exg_arena();
    auto bail = [&](int code) {
        r->release();
        delete r;
        return code;
    };
#define R_TRY(expr)                                                                                    \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            afail(EXG_ERR_BACKEND, std::string(#expr) + ": " + hipGetErrorString(e_));                 \
            return bail(EXG_ERR_BACKEND);                                                              \
        }                                                                                              \
    } while (0)
    r->Bc = body_capacity;
    r->Ic = island_capacity;
    r->Jc = joint_capacity;
    r->Xc = inverse_capacity;
    R_TRY(hipDeviceGetAttribute(&r->cus, hipDeviceAttributeMultiprocessorCount, 0));
    R_TRY(hipStreamCreate(&r->stream));
    R_TRY(r->classes.create());
    for (auto& e : r->ev) R_TRY(hipEventCreate(&e));
    for (auto& e : r->marks) R_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const size_t B = (size_t)body_capacity, I = (size_t)island_capacity, J = (size_t)std::max(joint_capacity, 1);
    R_TRY(r->pose.alloc(B * 8));
    R_TRY(r->mom.alloc(B * 6));
    R_TRY(r->mass.alloc(B));
    R_TRY(r->tens.alloc(B * 9));
    R_TRY(r->size.alloc(B));
    R_TRY(r->ext.alloc(B * 6));
    R_TRY(r->charges[0].alloc(B * 8));
    R_TRY(r->charges[1].alloc(B * 8));
    R_TRY(r->grav_stage.alloc(B * exg::GRAV_STAGE_BYTES / sizeof(double)));
    R_TRY(r->origin.alloc(I * 3));
    R_TRY(r->island_G.alloc(I));
    R_TRY(r->island_start.alloc(I));
    R_TRY(r->island_end.alloc(I));
    R_TRY(r->island_jstart.alloc(I));
    R_TRY(r->island_jend.alloc(I));
    R_TRY(r->seeded.alloc(I));
    R_TRY(r->stats.alloc(I * exg::ISLAND_STATS));
    R_TRY(r->fatal.alloc(1));
    R_TRY(hipMemset(r->fatal.p, 0, sizeof(int)));
    R_TRY(r->requeue.alloc(I));
    R_TRY(r->requeue_ctl.alloc(2));
    R_TRY(hipMemset(r->requeue_ctl.p, 0, 2 * sizeof(unsigned)));
    R_TRY(r->xacc_off.alloc(I));
    R_TRY(r->topo.alloc(I));
    R_TRY(r->xacc.alloc((size_t)inverse_capacity));
    R_TRY(r->jkind.alloc(J));
    R_TRY(r->ja.alloc(J));
    R_TRY(r->jb.alloc(J));
    R_TRY(r->jparams.alloc(J * 10));
#undef R_TRY
    r->island_class_of.assign(I, 0);
    r->bodies.init(body_capacity);
    r->islands.init(island_capacity);
    r->joints.init(joint_capacity);
    r->xaccs.init(inverse_capacity);
    *out = r;
    return EXG_OK;
}

int exg_arena_destroy(exg_arena* r) {
    if (!r) return EXG_OK;
    r->release();
    delete r;
    return EXG_OK;
}

int exg_arena_set_solver(exg_arena* r, double tol2, int ns_warm, int ns_seed, int ns_engine, int gravity_precision) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_set_solver: null arena");
    if (ns_warm < 0 || ns_seed < 1) return afail(EXG_ERR_INVALID, "exg_arena_set_solver: bad Newton–Schulz budgets");
    if (ns_engine != EXG_NS_MFMA && ns_engine != EXG_NS_VALU_ORDERED)
        return afail(EXG_ERR_INVALID, "exg_arena_set_solver: unknown Newton–Schulz engine");
    if (gravity_precision != EXG_PREC_F64 && gravity_precision != EXG_PREC_F32_COMPENSATED)
        return afail(EXG_ERR_INVALID, "exg_arena_set_solver: unknown gravity precision");
    r->tol2 = tol2;
    r->ns_warm = ns_warm;
    r->ns_seed = ns_seed;
    r->ns_engine = ns_engine;
    r->grav_prec = gravity_precision;
    r->settings_epoch++;
    return EXG_OK;
}

int exg_arena_set_precision(exg_arena* r, int gravity, int jacobian) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_set_precision: null arena");
    auto known = [](int p) { return p == EXG_PREC_F64 || p == EXG_PREC_F32_COMPENSATED; };
    if (!known(gravity) || !known(jacobian)) return afail(EXG_ERR_INVALID, "exg_arena_set_precision: unknown precision");
    r->grav_prec = gravity;
    r->jac_prec = jacobian;
    r->settings_epoch++;
    return EXG_OK;
}

int exg_arena_set_gravity_chunk(exg_arena* r, int chunk_len) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_set_gravity_chunk: null arena");
    if (chunk_len != 0 && chunk_len < 64)
        return afail(EXG_ERR_INVALID, "exg_arena_set_gravity_chunk: chunk below 64 sources");
    r->grav_chunk = chunk_len;   // worlds already admitted keep their chunking
    return EXG_OK;
}

int exg_arena_admit(exg_arena* r, int n_bodies, const double* pose, const double* momentum, const double* mass,
                    const double* inertia, const double* size, int n_islands, const int* island_start,
                    const double* origin, int n_joints, const int* kind, const int* a, const int* b,
                    const double* params, double G, int gravity_on, int* world_id) {
    if (!r || !world_id) return afail(EXG_ERR_INVALID, "exg_arena_admit: null argument");
    *world_id = -1;
    if (n_bodies <= 0 || n_islands <= 0 || n_joints < 0)
        return afail(EXG_ERR_INVALID, "exg_arena_admit: need at least one body and one island");
    if (!pose || !momentum || !mass || !inertia || !size || !island_start || !origin ||
        (n_joints > 0 && (!kind || !a || !b || !params)))
        return afail(EXG_ERR_INVALID, "exg_arena_admit: null array");
    std::vector<int> jstart;
    int max_m = 0, max_j = 0;
    if (int rc = exg::check_world("exg_arena_admit", n_bodies, n_islands, island_start, n_joints, kind, a, b, jstart,
                                  max_m, max_j))
        return rc;
    if (int rc = exg::check_bodies("exg_arena_admit", n_bodies, pose, momentum, mass, inertia, size, n_islands, origin,
                                   n_joints, kind, params, G))
        return rc;
    std::vector<int> md;
    if (int rc = exg::dynamic_counts("exg_arena_admit", n_islands, island_start, mass, md)) return rc;
    std::vector<exg::IslandTopo> topo(n_islands);
    for (int i = 0; i < n_islands; ++i) {
        const int j0 = jstart[i], ib0 = island_start[i];
        if (exg::island_class(island_start[i + 1] - ib0, jstart[i + 1] - j0) == exg::ISLAND_CLASS_LARGE)
            return afail(EXG_ERR_UNSUPPORTED, "exg_arena_admit: arenas take islands of up to 8 bodies and 16 joints "
                                              "(larger ones run in a world of their own)");
        if (exg::island_topology_of(island_start[i + 1] - ib0, md[i], jstart[i + 1] - j0, kind + j0, a + j0, b + j0,
                                    ib0, topo[i]) != 0)
            return afail(EXG_ERR_UNSUPPORTED, "exg_arena_admit: an island exceeds the kernels' topology record");
    }
    long long nx = 0;
    std::vector<long long> xo(n_islands);
    for (int i = 0; i < n_islands; ++i) {
        long long mm = island_start[i + 1] - island_start[i];
        xo[i] = nx;
        nx += mm * mm * 36;
    }
    if (!r->bodies.fits(n_bodies) || !r->islands.fits(n_islands) || !r->joints.fits(n_joints) || !r->xaccs.fits(nx))
        return afail(EXG_ERR_FULL, "exg_arena_admit: no contiguous room for the world (retire one and retry)");
    ArenaWorld w;
    w.live = true;
    w.nb = n_bodies;
    w.ni = n_islands;
    w.nj = n_joints;
    w.nx = nx;
    w.b0 = r->bodies.take(n_bodies);
    w.i0 = r->islands.take(n_islands);
    w.j0 = r->joints.take(n_joints);
    w.x0 = r->xaccs.take(nx);
    // every failure from here on gives the four ranges back
    struct Release {
        exg_arena* r;
        const ArenaWorld& w;
        bool armed = true;
        ~Release() {
            if (!armed) return;
            r->bodies.give(w.b0, w.nb);
            r->islands.give(w.i0, w.ni);
            r->joints.give(w.j0, w.nj);
            r->xaccs.give(w.x0, w.nx);
        }
    } release_on_failure{r, w};
    w.G = G;
    w.gravity_on = gravity_on != 0;
    w.chunk = r->chunk_for(n_bodies, r->grav_chunk);
    w.mapping = r->joint_mapping;
    const size_t B = (size_t)n_bodies, I = (size_t)n_islands, J = (size_t)n_joints;
    auto up = [&](void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
    A_TRY(up(r->pose.p + w.b0 * 8, pose, B * 8 * sizeof(double)));
    A_TRY(up(r->mom.p + w.b0 * 6, momentum, B * 6 * sizeof(double)));
    A_TRY(up(r->mass.p + w.b0, mass, B * sizeof(double)));
    A_TRY(up(r->tens.p + w.b0 * 9, inertia, B * 9 * sizeof(double)));
    A_TRY(up(r->size.p + w.b0, size, B * sizeof(double)));
    A_TRY(up(r->origin.p + w.i0 * 3, origin, I * 3 * sizeof(double)));
    std::vector<int> is(I), ie(I), js(I), je(I), zero_i(I * exg::ISLAND_STATS, 0);
    std::vector<double> gi(I, G);
    std::vector<long long> xoff(I);
    for (size_t i = 0; i < I; ++i) {
        is[i] = (int)w.b0 + island_start[i];
        ie[i] = (int)w.b0 + island_start[i + 1];
        js[i] = (int)w.j0 + jstart[i];
        je[i] = (int)w.j0 + jstart[i + 1];
        xoff[i] = w.x0 + xo[i];
        r->island_class_of[(size_t)w.i0 + i] =
            exg::island_class(island_start[i + 1] - island_start[i], jstart[i + 1] - jstart[i]);
    }
    A_TRY(up(r->island_start.p + w.i0, is.data(), I * sizeof(int)));
    A_TRY(up(r->island_end.p + w.i0, ie.data(), I * sizeof(int)));
    A_TRY(up(r->island_jstart.p + w.i0, js.data(), I * sizeof(int)));
    A_TRY(up(r->island_jend.p + w.i0, je.data(), I * sizeof(int)));
    A_TRY(up(r->xacc_off.p + w.i0, xoff.data(), I * sizeof(long long)));
    A_TRY(up(r->topo.p + w.i0, topo.data(), I * sizeof(exg::IslandTopo)));
    A_TRY(up(r->island_G.p + w.i0, gi.data(), I * sizeof(double)));
    A_TRY(up(r->seeded.p + w.i0, zero_i.data(), I * sizeof(int)));
    A_TRY(up(r->stats.p + w.i0 * exg::ISLAND_STATS, zero_i.data(), I * exg::ISLAND_STATS * sizeof(int)));
    if (J) {
        std::vector<int> ga(J), gb(J);
        for (size_t q = 0; q < J; ++q) {
            ga[q] = (int)w.b0 + a[q];
            gb[q] = (int)w.b0 + b[q];
        }
        A_TRY(up(r->jkind.p + w.j0, kind, J * sizeof(int)));
        A_TRY(up(r->ja.p + w.j0, ga.data(), J * sizeof(int)));
        A_TRY(up(r->jb.p + w.j0, gb.data(), J * sizeof(int)));
        A_TRY(up(r->jparams.p + w.j0 * 10, params, J * 10 * sizeof(double)));
    }
    // zero charges (com 0, G·m 0, size 1) in both buffers: nothing gravitates on the
    // world's first step, as on the reference's first frame (gravity.rs:620-626)
    std::vector<double> zc(B * 8, 0.0);
    for (size_t i = 0; i < B; ++i) zc[i * 8 + 7] = 1.0;
    A_TRY(up(r->charges[0].p + w.b0 * 8, zc.data(), B * 8 * sizeof(double)));
    A_TRY(up(r->charges[1].p + w.b0 * 8, zc.data(), B * 8 * sizeof(double)));
    A_TRY(hipMemset(r->xacc.p + w.x0, 0, (size_t)nx * sizeof(double)));
    // ids are never reused, so a retired world's id stays invalid
    const int id = (int)r->worlds.size();
    r->worlds.push_back(w);
    release_on_failure.armed = false;
    r->dirty = true;
    *world_id = id;
    return EXG_OK;
}

static ArenaWorld* live_world(exg_arena* r, int id) {
    if (!r || id < 0 || (size_t)id >= r->worlds.size() || !r->worlds[(size_t)id].live) return nullptr;
    return &r->worlds[(size_t)id];
}

int exg_arena_retire(exg_arena* r, int world_id) {
    ArenaWorld* w = live_world(r, world_id);
    if (!w) return afail(EXG_ERR_INVALID, "exg_arena_retire: no such world");
    A_TRY(hipStreamSynchronize(r->stream));
    r->bodies.give(w->b0, w->nb);
    r->islands.give(w->i0, w->ni);
    r->joints.give(w->j0, w->nj);
    r->xaccs.give(w->x0, w->nx);
    w->live = false;
    r->dirty = true;
    return EXG_OK;
}

// One step of every live world, enqueued on the arena's stream.
static int enqueue_arena_step(exg_arena* r, double dt, double warp) {
    double* front = r->charges[r->front].p;
    double* back = r->charges[1 - r->front].p;
    exg::launch_gravity(front, r->entries.p, (int)r->n_entries, r->tiles.p, (int)r->n_tiles, r->grav_prec,
                        r->partial.p, r->ext.p, r->grav_stage.p, r->stage_n, r->stream);
    A_TRY(hipGetLastError());
    exg::IslandArgs a{};
    a.pose = r->pose.p;
    a.mom = r->mom.p;
    a.mass = r->mass.p;
    a.tens = r->tens.p;
    a.size = r->size.p;
    a.ext = r->ext.p;
    a.island_start = r->island_start.p;
    a.island_end = r->island_end.p;
    a.island_jstart = r->island_jstart.p;
    a.island_jend = r->island_jend.p;
    a.origin = r->origin.p;
    a.jkind = r->jkind.p;
    a.ja = r->ja.p;
    a.jb = r->jb.p;
    a.jparams = r->jparams.p;
    a.xacc = r->xacc.p;
    a.xacc_off = r->xacc_off.p;
    a.topo = r->topo.p;
    a.seeded = r->seeded.p;
    a.charges_out = back;
    a.stats = r->stats.p;
    a.fatal = r->fatal.p;
    a.requeue = r->requeue.p;
    a.requeue_ctl = r->requeue_ctl.p;
    a.dt = dt;
    a.warp = warp;
    a.tol2 = r->tol2;
    a.G = 1.0;
    a.G_of_island = r->island_G.p;  // each world publishes G·m with its own G
    a.ns_warm = r->ns_warm;
    a.ns_seed = r->ns_seed;
    a.ns_engine = r->ns_engine;
    a.jac_f32 = r->jac_prec == EXG_PREC_F32_COMPENSATED;
    int* ids[exg::ISLAND_CLASSES];
    for (int c = 0; c < exg::ISLAND_CLASSES; ++c) ids[c] = r->class_ids[c].p;
    const exg::ClassLaunch cl{r->class_n, ids, r->stream, r->classes.side, r->classes.fork, r->classes.join};
    A_TRY(exg::launch_classes(a, cl));
    r->front = 1 - r->front;
    return EXG_OK;
}

// Graph of two consecutive steps from charge buffer `front` (one round trip of the
// double buffer), re-captured when dt, warp, a setting or the set of worlds changed.
static int arena_graph(exg_arena* r, double dt, double warp, hipGraphExec_t* out) {
    auto& g = r->graphs[r->front];
    if (g.exec && g.dt == dt && g.warp == warp && g.settings == r->settings_epoch) {
        *out = g.exec;
        return EXG_OK;
    }
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    const int front0 = r->front;
    A_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_arena_step(r, dt, warp);
    if (rc == EXG_OK) rc = enqueue_arena_step(r, dt, warp);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(r->stream, &graph);
    r->front = front0;
    if (rc != EXG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return afail(EXG_ERR_BACKEND, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return afail(EXG_ERR_BACKEND, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    g.dt = dt;
    g.warp = warp;
    g.settings = r->settings_epoch;
    *out = g.exec;
    return EXG_OK;
}

int exg_arena_step(exg_arena* r, int n_steps, double dt, double warp) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_step: null arena");
    if (n_steps < 0) return afail(EXG_ERR_INVALID, "exg_arena_step: negative step count");
    if (!(dt > 0.0) || !std::isfinite(dt) || !(warp >= 1.0) || !std::isfinite(warp))
        return afail(EXG_ERR_INVALID, "exg_arena_step: dt must be positive and warp at least 1 (both finite)");
    if (r->dirty)
        if (int rc = r->rebuild()) return rc;
    bool any = false;
    for (auto& w : r->worlds) any |= w.live;
    if (!any || n_steps == 0) return EXG_OK;
    A_TRY(hipEventRecord(r->ev[0], r->stream));
    int s = 0;
    if (r->use_graph && n_steps >= 2) {
        hipGraphExec_t exec = nullptr;
        if (int rc = arena_graph(r, dt, warp, &exec)) return rc;
        for (; s + 1 < n_steps; s += 2) A_TRY(hipGraphLaunch(exec, r->stream));
    }
    for (; s < n_steps; ++s)
        if (int rc = enqueue_arena_step(r, dt, warp)) return rc;
    A_TRY(hipEventRecord(r->ev[1], r->stream));
    return EXG_OK;
}

int exg_arena_autotune_gravity(exg_arena* r, int* chosen_chunk) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_autotune_gravity: null arena");
    bool any = false;
    for (auto& w : r->worlds) any |= w.live && w.gravity_on;
    if (!any) {
        if (chosen_chunk) *chosen_chunk = r->grav_chunk;
        return EXG_OK;
    }
    // each candidate setting: the live worlds' work table chunked by it, one warm pass,
    // then the mean of five timed passes on the current charges (ext is recomputed by
    // the next step); the worlds' own chunking is restored afterwards
    static const int candidates[] = {0, 128, 256, 512, 1024, 2048, 4096, 8192};
    int best = r->grav_chunk;
    float best_ms = -1.0f;
    for (int c : candidates) {
        if (int rc = r->rebuild(c)) return rc;
        double* front = r->charges[r->front].p;
        exg::launch_gravity(front, r->entries.p, (int)r->n_entries, r->tiles.p, (int)r->n_tiles, r->grav_prec,
                            r->partial.p, r->ext.p, r->grav_stage.p, r->stage_n, r->stream);
        A_TRY(hipEventRecord(r->ev[0], r->stream));
        for (int k = 0; k < 5; ++k)
            exg::launch_gravity(front, r->entries.p, (int)r->n_entries, r->tiles.p, (int)r->n_tiles, r->grav_prec,
                                r->partial.p, r->ext.p, r->grav_stage.p, r->stage_n, r->stream);
        A_TRY(hipEventRecord(r->ev[1], r->stream));
        A_TRY(hipGetLastError());
        A_TRY(hipEventSynchronize(r->ev[1]));
        float ms = 0.0f;
        A_TRY(hipEventElapsedTime(&ms, r->ev[0], r->ev[1]));
        if (best_ms < 0.0f || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    r->grav_chunk = best;   // for worlds admitted from now on
    if (int rc = r->rebuild()) return rc;
    if (chosen_chunk) *chosen_chunk = best;
    return EXG_OK;
}

int exg_arena_set_joint_mapping(exg_arena* r, int mapping) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_set_joint_mapping: null arena");
    if (mapping != EXG_JOINT_MAPPING_SPLIT && mapping != EXG_JOINT_MAPPING_FLAT)
        return afail(EXG_ERR_INVALID, "exg_arena_set_joint_mapping: unknown mapping " + std::to_string(mapping));
    r->joint_mapping = mapping;   // for worlds admitted from now on
    return EXG_OK;
}

int exg_arena_autotune_joints(exg_arena* r, int* chosen_mapping) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_autotune_joints: null arena");
    if (r->dirty)
        if (int rc = r->rebuild()) return rc;
    bool any = false;
    for (auto& w : r->worlds) any |= w.live;
    if (!any) {
        if (chosen_mapping) *chosen_mapping = r->joint_mapping;
        return EXG_OK;
    }
    A_TRY(hipStreamSynchronize(r->stream));
    // what a step writes, saved and put back after every trial step (the mappings
    // give the same results, so the arena's worlds end where they started)
    struct Saved {
        void* dev;
        size_t bytes;
        void* copy;
    };
    std::vector<Saved> saved;
    auto keep = [&](void* dev, size_t bytes) { saved.push_back({dev, bytes, nullptr}); };
    keep(r->pose.p, r->pose.n * sizeof(double));
    keep(r->mom.p, r->mom.n * sizeof(double));
    keep(r->origin.p, r->origin.n * sizeof(double));
    keep(r->xacc.p, r->xacc.n * sizeof(double));
    keep(r->charges[0].p, r->charges[0].n * sizeof(double));
    keep(r->charges[1].p, r->charges[1].n * sizeof(double));
    keep(r->seeded.p, r->seeded.n * sizeof(int));
    keep(r->stats.p, r->stats.n * sizeof(int));
    keep(r->fatal.p, r->fatal.n * sizeof(int));
    auto release = [&]() {
        for (auto& x : saved)
            if (x.copy) (void)hipFree(x.copy);
    };
    for (auto& x : saved) {
        if (!x.bytes) continue;
        hipError_t e = hipMalloc(&x.copy, x.bytes);
        if (e == hipSuccess) e = hipMemcpyAsync(x.copy, x.dev, x.bytes, hipMemcpyDeviceToDevice, r->stream);
        if (e != hipSuccess) {
            release();
            return afail(EXG_ERR_BACKEND, std::string("exg_arena_autotune_joints: ") + hipGetErrorString(e));
        }
    }
    const int front0 = r->front, mapping0 = r->joint_mapping;
    auto restore = [&]() -> hipError_t {
        for (auto& x : saved)
            if (x.bytes) {
                const hipError_t e = hipMemcpyAsync(x.dev, x.copy, x.bytes, hipMemcpyDeviceToDevice, r->stream);
                if (e != hipSuccess) return e;
            }
        r->front = front0;
        return hipSuccess;
    };
    // each mapping: one warm step, then the best of three timed steps, each from the
    // saved state, at the step of the mixed-workload benchmarks (h = 1e-2)
    static const int candidates[] = {EXG_JOINT_MAPPING_SPLIT, EXG_JOINT_MAPPING_FLAT};
    int best = mapping0;
    float best_ms = -1.0f;
    int rc = EXG_OK;
    for (int c : candidates) {
        rc = r->rebuild(-1, c);   // every live world's 8-body islands on mapping c
        for (int rep = 0; rep < 4 && rc == EXG_OK; ++rep) {
            hipError_t e = restore();
            if (e == hipSuccess) e = hipEventRecord(r->ev[0], r->stream);
            if (e == hipSuccess) rc = enqueue_arena_step(r, 1e-2, 1.0);
            if (rc == EXG_OK && e == hipSuccess) e = hipEventRecord(r->ev[1], r->stream);
            if (rc == EXG_OK && e == hipSuccess) e = hipEventSynchronize(r->ev[1]);
            float ms = 0.0f;
            if (rc == EXG_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, r->ev[0], r->ev[1]);
            if (e != hipSuccess) rc = afail(EXG_ERR_BACKEND, std::string("exg_arena_autotune_joints: ") + hipGetErrorString(e));
            if (rc == EXG_OK && rep > 0 && (best_ms < 0.0f || ms < best_ms)) {
                best_ms = ms;
                best = c;
            }
        }
        if (rc != EXG_OK) break;
    }
    const hipError_t e = restore();
    const hipError_t es = hipStreamSynchronize(r->stream);
    release();
    if (int rc2 = r->rebuild()) return rc2;   // the worlds' own mappings again
    if (rc != EXG_OK) return rc;
    if (e != hipSuccess || es != hipSuccess)
        return afail(EXG_ERR_BACKEND, std::string("exg_arena_autotune_joints: ") +
                                          hipGetErrorString(e != hipSuccess ? e : es));
    if (int rc2 = exg_arena_set_joint_mapping(r, best)) return rc2;
    if (chosen_mapping) *chosen_mapping = best;
    return EXG_OK;
}

int exg_arena_set_graph(exg_arena* r, int enable) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_set_graph: null arena");
    r->use_graph = enable != 0;
    return EXG_OK;
}

int exg_arena_mark(exg_arena* r, uint64_t* ticket) {
    if (!r || !ticket) return afail(EXG_ERR_INVALID, "exg_arena_mark: null argument");
    const uint64_t t = r->marks_issued;
    hipEvent_t e = r->marks[t % exg_arena::N_MARKS];
    if (t >= (uint64_t)exg_arena::N_MARKS) A_TRY(hipEventSynchronize(e));   // the ticket it carried has passed
    A_TRY(hipEventRecord(e, r->stream));
    r->marks_issued = t + 1;
    *ticket = t;
    return EXG_OK;
}

int exg_arena_mark_done(exg_arena* r, uint64_t ticket, int* done) {
    if (!r || !done) return afail(EXG_ERR_INVALID, "exg_arena_mark_done: null argument");
    if (ticket >= r->marks_issued) return afail(EXG_ERR_INVALID, "exg_arena_mark_done: ticket not issued");
    if (ticket + exg_arena::N_MARKS <= r->marks_issued) {
        *done = 1;   // its ring slot was reused, which waited for it
        return EXG_OK;
    }
    const hipError_t e = hipEventQuery(r->marks[ticket % exg_arena::N_MARKS]);
    if (e == hipSuccess) *done = 1;
    else if (e == hipErrorNotReady) *done = 0;
    else return afail(EXG_ERR_BACKEND, std::string("exg_arena_mark_done: ") + hipGetErrorString(e));
    return EXG_OK;
}

int exg_arena_mark_wait(exg_arena* r, uint64_t ticket) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_mark_wait: null arena");
    if (ticket >= r->marks_issued) return afail(EXG_ERR_INVALID, "exg_arena_mark_wait: ticket not issued");
    if (ticket + exg_arena::N_MARKS <= r->marks_issued) return EXG_OK;
    A_TRY(hipEventSynchronize(r->marks[ticket % exg_arena::N_MARKS]));
    return EXG_OK;
}

int exg_arena_sync(exg_arena* r) {
    if (!r) return afail(EXG_ERR_INVALID, "exg_arena_sync: null arena");
    A_TRY(hipStreamSynchronize(r->stream));
    int c = 0;
    A_TRY(hipMemcpy(&c, r->fatal.p, sizeof(int), hipMemcpyDeviceToHost));
    if (c > r->fatal_reported) {
        const long long fresh = c - r->fatal_reported;
        r->fatal_reported = c;
        return afail(EXG_ERR_FATAL, "exg_arena_sync: " + std::to_string(fresh) +
                                        " island-step(s) ended with a non-finite state (per-island counts in "
                                        "exg_arena_world_stats slot 6)");
    }
    return EXG_OK;
}

int exg_arena_download(exg_arena* r, int world_id, double* pose, double* momentum, double* origin) {
    ArenaWorld* w = live_world(r, world_id);
    if (!w) return afail(EXG_ERR_INVALID, "exg_arena_download: no such world");
    A_TRY(hipStreamSynchronize(r->stream));
    if (pose) A_TRY(hipMemcpy(pose, r->pose.p + w->b0 * 8, (size_t)w->nb * 8 * sizeof(double), hipMemcpyDeviceToHost));
    if (momentum)
        A_TRY(hipMemcpy(momentum, r->mom.p + w->b0 * 6, (size_t)w->nb * 6 * sizeof(double), hipMemcpyDeviceToHost));
    if (origin)
        A_TRY(hipMemcpy(origin, r->origin.p + w->i0 * 3, (size_t)w->ni * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return EXG_OK;
}

int exg_arena_world_stats(exg_arena* r, int world_id, int32_t* stats) {
    ArenaWorld* w = live_world(r, world_id);
    if (!w || !stats) return afail(EXG_ERR_INVALID, "exg_arena_world_stats: no such world");
    A_TRY(hipStreamSynchronize(r->stream));
    A_TRY(hipMemcpy(stats, r->stats.p + w->i0 * exg::ISLAND_STATS, (size_t)w->ni * exg::ISLAND_STATS * sizeof(int),
                    hipMemcpyDeviceToHost));
    return EXG_OK;
}

int exg_arena_last_ms(exg_arena* r, double* ms) {
    if (!r || !ms) return afail(EXG_ERR_INVALID, "exg_arena_last_ms: null argument");
    A_TRY(hipEventSynchronize(r->ev[1]));
    float t = 0.f;
    A_TRY(hipEventElapsedTime(&t, r->ev[0], r->ev[1]));
    *ms = t;
    return EXG_OK;
}

int exg_arena_usage(exg_arena* r, int64_t* used4, int64_t* capacity4) {
    if (!r || !used4 || !capacity4) return afail(EXG_ERR_INVALID, "exg_arena_usage: null argument");
    const RangeAlloc* al[4] = {&r->bodies, &r->islands, &r->joints, &r->xaccs};
    for (int k = 0; k < 4; ++k) {
        used4[k] = al[k]->used;
        capacity4[k] = al[k]->capacity;
    }
    return EXG_OK;
}

int exg_arena_footprint(int n_bodies, int n_islands, int n_joints, long long inverse_doubles, int gravity_chunk,
                        int64_t* bytes) {
    if (!bytes || n_bodies < 0 || n_islands < 0 || n_joints < 0 || inverse_doubles < 0 ||
        (gravity_chunk != 0 && gravity_chunk < 64))
        return afail(EXG_ERR_INVALID, "exg_arena_footprint: bad arguments");
    if (gravity_chunk == 0) {   // the single-world rule (at most 64 workgroups per CU of 256)
        std::vector<exg::GravEntry> en;
        std::vector<exg::GravTile> tl;
        int cus = 256;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) cus = 256;
        exg::gravity_add_world(en, tl, 0, std::max(n_bodies, 1), 1.0, cus);
        gravity_chunk = en.empty() ? std::max(n_bodies, 1) : std::max(1, en[0].src1 - en[0].src0);
    }
    const long long per_body = (8 + 6 + 1 + 9 + 1 + 6 + 16) * 8;                 // state, ext, two charges
    const long long per_island = 3 * 8 + 5 * 4 + exg::ISLAND_STATS * 4 + 8 +      // origin, ranges, stats, offset,
                                 (long long)sizeof(exg::IslandTopo) + 4;         // topology record, requeue slot
    const long long per_joint = 10 * 8 + 3 * 4;                                   // params, kind, ends
    const long long tiles = (n_bodies + 63) / 64;
    const long long chunks = (n_bodies + gravity_chunk - 1) / gravity_chunk;
    const long long grav = tiles * chunks * (exg::GRAV_PARTIAL_PER_ENTRY * 8 + (long long)sizeof(exg::GravEntry)) +
                           tiles * (long long)sizeof(exg::GravTile);
    *bytes = per_body * n_bodies + per_island * n_islands + per_joint * n_joints + inverse_doubles * 8 + grav;
    return EXG_OK;
}

}  // extern "C"
