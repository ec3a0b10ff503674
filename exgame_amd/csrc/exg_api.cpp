// Host side of the C ABI (include/exgame.h): device buffers of a world in HBM, the
// per-step launch sequence, error mapping. Compiled by hipcc into libexgame_hip.so.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/exgame.h"
#include "island.h"
#include "devbuf.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(EXG_ERR_BACKEND, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

using exg::DevBuf;

constexpr int MAX_ISLAND_BODIES = exg::LARGE_MAXM;
constexpr int MAX_ISLAND_JOINTS = exg::LARGE_MAXJ;
constexpr int EXPLICIT_MAX_BODIES = 8;    // the explicit kernel keeps an island in registers
constexpr int EXPLICIT_MAX_JOINTS = 16;

}  // namespace

int exg::set_error(int code, const char* msg) { return fail(code, msg); }

int exg::check_world(const char* who, int n_bodies, int n_islands, const int* island_start, int n_joints,
                     const int* kind, const int* a, const int* b, std::vector<int>& jstart, int& max_m, int& max_j) {
    if (island_start[0] != 0 || island_start[n_islands] != n_bodies)
        return fail(EXG_ERR_INVALID, std::string(who) + ": island_start must run from 0 to n_bodies");
    // joints grouped by island, in declaration order, both ends in one island
    std::vector<int> island_of(n_bodies);
    max_m = 0;
    for (int i = 0; i < n_islands; ++i) {
        int s = island_start[i], e = island_start[i + 1];
        if (e <= s) return fail(EXG_ERR_INVALID, std::string(who) + ": empty island " + std::to_string(i));
        max_m = std::max(max_m, e - s);
        for (int k = s; k < e; ++k) island_of[k] = i;
    }
    jstart.assign(n_islands + 1, 0);
    int prev = 0;
    for (int q = 0; q < n_joints; ++q) {
        if (a[q] < 0 || a[q] >= n_bodies || b[q] < 0 || b[q] >= n_bodies || a[q] == b[q])
            return fail(EXG_ERR_INVALID, std::string(who) + ": joint " + std::to_string(q) + " has a bad body index");
        int ia = island_of[a[q]];
        if (island_of[b[q]] != ia)
            return fail(EXG_ERR_INVALID, std::string(who) + ": joint " + std::to_string(q) + " spans two islands");
        if (ia < prev) return fail(EXG_ERR_INVALID, std::string(who) + ": joints must be grouped island by island");
        if (kind[q] < 0 || kind[q] > 3) return fail(EXG_ERR_NO_SHADER, std::string(who) + ": unknown joint kind");
        prev = ia;
        jstart[ia + 1]++;
    }
    max_j = 0;
    for (int i = 0; i < n_islands; ++i) {
        max_j = std::max(max_j, jstart[i + 1]);
        jstart[i + 1] += jstart[i];
    }
    if (max_m > MAX_ISLAND_BODIES || max_j > MAX_ISLAND_JOINTS)
        return fail(EXG_ERR_UNSUPPORTED, std::string(who) + ": islands up to " + std::to_string(MAX_ISLAND_BODIES) +
                                             " bodies and " + std::to_string(MAX_ISLAND_JOINTS) +
                                             " joints are supported (got " + std::to_string(max_m) + " bodies, " +
                                             std::to_string(max_j) + " joints)");

    return EXG_OK;
}

int exg::check_bodies(const char* who, int n_bodies, const double* pose, const double* momentum, const double* mass,
                      const double* inertia, const double* size, int n_islands, const double* origin, int n_joints,
                      const int* kind, const double* params, double G) {
    auto bad = [&](const std::string& what, long long i) {
        return fail(EXG_ERR_INVALID, std::string(who) + ": " + what + " (index " + std::to_string(i) + ")");
    };
    if (!std::isfinite(G)) return fail(EXG_ERR_INVALID, std::string(who) + ": G is not finite");
    for (long long i = 0; i < n_bodies; ++i) {
        // zero mass is Inertia::Kinematic (crates/newton/src/inertia.rs:103-151): a
        // prescribed reference whose tensor is never read
        const bool kinematic = mass[i] == 0.0;
        if (!kinematic && (!(mass[i] > 0.0) || !std::isfinite(mass[i])))
            return bad("mass must be positive (or zero: kinematic) and finite", i);
        const double* t = inertia + 9 * i;
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(t[k])) return bad("inertia tensor is not finite", i);
        // positive definite (Sylvester: leading minors > 0): the tensor is inverted per step
        const double m1 = t[0], m2 = t[0] * t[4] - t[1] * t[3];
        const double m3 = t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
                          t[2] * (t[3] * t[7] - t[4] * t[6]);
        if (!kinematic && !(m1 > 0.0 && m2 > 0.0 && m3 > 0.0)) return bad("inertia tensor is not positive definite", i);
        const double* p = pose + 8 * i;
        double r2 = 0.0;
        for (int k = 0; k < 8; ++k)
            if (!std::isfinite(p[k])) return bad("pose is not finite", i);
        for (int k = 0; k < 4; ++k) r2 += p[k] * p[k];   // 1, e12, e13, e23: the rotor part
        if (!(r2 > 0.0)) return bad("pose has no rotor part", i);
        for (int k = 0; k < 6; ++k)
            if (!std::isfinite(momentum[6 * i + k])) return bad("momentum is not finite", i);
        if (!std::isfinite(size[i]) || size[i] < 0.0) return bad("size must be finite and non-negative", i);
    }
    for (long long i = 0; i < 3LL * n_islands; ++i)
        if (!std::isfinite(origin[i])) return bad("island origin is not finite", i / 3);
    for (long long q = 0; q < n_joints; ++q) {
        const int np = (kind[q] == 0 || kind[q] == 1) ? 10 : (kind[q] == 2 ? 2 : 1);
        for (int k = 0; k < np; ++k)
            if (!std::isfinite(params[10 * q + k])) return bad("joint parameter is not finite", q);
    }
    return EXG_OK;
}

int exg::dynamic_counts(const char* who, int n_islands, const int* island_start, const double* mass,
                        std::vector<int>& md) {
    md.assign(n_islands, 0);
    for (int i = 0; i < n_islands; ++i) {
        int d = 0;
        for (int k = island_start[i]; k < island_start[i + 1]; ++k) {
            if (mass[k] == 0.0) continue;
            if (k - island_start[i] != d)
                return fail(EXG_ERR_INVALID, std::string(who) + ": island " + std::to_string(i) + " has dynamic body " +
                                                 std::to_string(k) + " after a kinematic one (kinematic bodies go last)");
            ++d;
        }
        md[i] = d;
    }
    return EXG_OK;
}

struct exg_world {
    int B = 0, I = 0, J = 0;
    double G = 1.0;
    bool gravity = true;
    int integrator = EXG_NEWTON;
    double tol2 = 1e-8;
    int ns_warm = 2, ns_seed = 24;
    int ns_engine = EXG_NS_MFMA;
    int joint_mapping = EXG_JOINT_MAPPING_SPLIT;
    int max_m = 0, max_j = 0;
    DevBuf<int> class_ids[exg::ISLAND_CLASSES];   // island ids routed to each kernel variant
    int class_n[exg::ISLAND_CLASSES] = {};
    hipStream_t stream = nullptr;
    exg::ClassStreams classes;       // the island classes run side by side
    DevBuf<double> pose, mom, mass, tens, size, ext, origin, jparams, xacc, charges[2], grav_partial, grav_stage;
    DevBuf<exg::GravEntry> grav_entries;
    DevBuf<exg::GravTile> grav_tiles;
    int grav_prec = EXG_PREC_F64, jac_prec = EXG_PREC_F64;
    bool grav_island_reduce = true;   // the islands add the gravity partial sums (where eligible)
    double field[3] = {0.0, 0.0, 0.0}; // UniformField acceleration
    bool field_on = false;
    bool island_reduce_ok = false;    // every island on-chip and every tile within the islands' chunk cap
    bool explicit_reduce_ok = false;  // every tile within the chunk cap (the explicit kernel)
    DevBuf<int> island_start, island_jstart, jkind, ja, jb, seeded, stats;
    DevBuf<int> fatal;              // island-steps that ended with a non-finite state
    DevBuf<int> requeue;            // single-body islands whose step subdivides (one per such island)
    DevBuf<unsigned> requeue_ctl;   // the requeue list's length and workgroup counter (0 between launches)
    long long fatal_reported = 0;   // the count the last exg_world_sync reported
    DevBuf<long long> xacc_off;
    DevBuf<exg::IslandTopo> topo;   // per island, Newton integrator only
    DevBuf<exg::LargeTopo> ltopo;   // per island when the world has large islands (others: unused)
    DevBuf<double> lws;             // the large islands' workspace
    DevBuf<long long> lws_off;      // per island
    int front = 0;
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    hipEvent_t* step_ev = nullptr;   // this step's four timing events (gravity start/stop, island start/stop)
    bool use_graph = false;
    bool graph_ms_valid = false;
    int settings_epoch = 0;      // bumped when a setting baked into kernel arguments changes
    struct Graph {
        hipGraphExec_t exec = nullptr;
        double dt = 0.0, warp = 0.0;
        int settings = -1;
    } graphs[2];                 // by the charge buffer the two-step graph starts from

    void release() {
        for (auto* b : {&pose, &mom, &mass, &tens, &size, &ext, &origin, &jparams, &xacc, &charges[0], &charges[1],
                        &grav_partial, &grav_stage})
            b->release();
        for (auto* b : {&island_start, &island_jstart, &jkind, &ja, &jb, &seeded, &stats, &fatal, &requeue}) b->release();
        requeue_ctl.release();
        grav_entries.release();
        grav_tiles.release();
        xacc_off.release();
        topo.release();
        ltopo.release();
        lws.release();
        lws_off.release();
        for (auto& c : class_ids) c.release();
        for (auto e : ev) (void)hipEventDestroy(e);
        ev.clear();
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        classes.destroy();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

extern "C" {

const char* exg_last_error(void) { return g_last_error.c_str(); }

int exg_device_info_get(exg_device_info* out) {
    if (!out) return fail(EXG_ERR_INVALID, "exg_device_info_get: null output");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, 0));
    out->compute_units = p.multiProcessorCount;
    out->wave_size = p.warpSize;
    out->lds_bytes_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    std::snprintf(out->name, sizeof(out->name), "%s", p.gcnArchName);
    return EXG_OK;
}

int exg_world_create(int n_bodies, const double* pose, const double* momentum, const double* mass,
                     const double* inertia, const double* size, int n_islands, const int* island_start,
                     const double* origin, int n_joints, const int* kind, const int* a, const int* b,
                     const double* params, double G, int gravity_on, exg_world** out) {
    if (!out) return fail(EXG_ERR_INVALID, "exg_world_create: null output");
    *out = nullptr;
    if (n_bodies <= 0 || n_islands <= 0 || n_joints < 0)
        return fail(EXG_ERR_INVALID, "exg_world_create: need at least one body and one island");
    if (!pose || !momentum || !mass || !inertia || !size || !island_start || !origin)
        return fail(EXG_ERR_INVALID, "exg_world_create: null state array");
    if (n_joints > 0 && (!kind || !a || !b || !params))
        return fail(EXG_ERR_INVALID, "exg_world_create: null joint array");
    std::vector<int> jstart;
    int max_m = 0, max_j = 0;
    if (int rc = exg::check_world("exg_world_create", n_bodies, n_islands, island_start, n_joints, kind, a, b, jstart,
                                  max_m, max_j))
        return rc;
    if (int rc = exg::check_bodies("exg_world_create", n_bodies, pose, momentum, mass, inertia, size, n_islands, origin,
                                   n_joints, kind, params, G))
        return rc;
    std::vector<int> md;
    if (int rc = exg::dynamic_counts("exg_world_create", n_islands, island_start, mass, md)) return rc;
    std::vector<exg::IslandTopo> topo(n_islands);
    for (int i = 0; i < n_islands; ++i) {
        const int j0 = jstart[i], b0 = island_start[i], mi = island_start[i + 1] - b0;
        if (exg::island_topology_of(mi, md[i], jstart[i + 1] - j0, kind + j0, a + j0, b + j0, b0, topo[i]) != 0)
            return fail(EXG_ERR_UNSUPPORTED, "exg_world_create: an island exceeds the kernels' topology record");
    }

    auto* w = new exg_world();
    w->B = n_bodies;
    w->I = n_islands;
    w->J = n_joints;
    w->G = G;
    w->gravity = gravity_on != 0;
    w->max_m = max_m;
    w->max_j = max_j;
    auto bail = [&](int code) { w->release(); delete w; return code; };
#define W_TRY(expr)                                                                                    \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            fail(EXG_ERR_BACKEND, std::string(#expr) + ": " + hipGetErrorString(e_));                  \
            return bail(EXG_ERR_BACKEND);                                                              \
        }                                                                                              \
    } while (0)
    W_TRY(hipStreamCreate(&w->stream));
    W_TRY(w->classes.create());
    const size_t B = n_bodies, I = n_islands, Jn = n_joints;
    W_TRY(w->pose.alloc(B * 8));
    W_TRY(w->mom.alloc(B * 6));
    W_TRY(w->mass.alloc(B));
    W_TRY(w->tens.alloc(B * 9));
    W_TRY(w->size.alloc(B));
    W_TRY(w->ext.alloc(B * 6));
    W_TRY(w->origin.alloc(I * 3));
    W_TRY(w->charges[0].alloc(B * 8));
    W_TRY(w->charges[1].alloc(B * 8));
    {
        int cus = 0;
        W_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
        std::vector<exg::GravEntry> entries;
        std::vector<exg::GravTile> tiles;
        exg::gravity_add_world(entries, tiles, 0, n_bodies, G, cus);
        W_TRY(w->grav_entries.alloc(entries.size()));
        W_TRY(w->grav_entries.upload(entries.data(), entries.size()));
        W_TRY(w->grav_tiles.alloc(tiles.size()));
        W_TRY(w->grav_tiles.upload(tiles.data(), tiles.size()));
        W_TRY(w->grav_partial.alloc(entries.size() * exg::GRAV_PARTIAL_PER_ENTRY));
        bool chunks_ok = true;
        for (const auto& tl : tiles) chunks_ok &= tl.ne <= exg::GRAV_ISLAND_REDUCE_MAX_CHUNKS;
        w->island_reduce_ok = chunks_ok;   // and every island on-chip (checked with the classes below)
        w->explicit_reduce_ok = chunks_ok;
        W_TRY(w->grav_stage.alloc(B * exg::GRAV_STAGE_BYTES / sizeof(double)));
    }
    W_TRY(w->island_start.alloc(I + 1));
    W_TRY(w->island_jstart.alloc(I + 1));
    W_TRY(w->seeded.alloc(I));
    W_TRY(w->stats.alloc(I * exg::ISLAND_STATS));
    W_TRY(w->fatal.alloc(1));
    W_TRY(w->xacc_off.alloc(I));
    {
        std::vector<int> ids[exg::ISLAND_CLASSES];
        for (size_t i = 0; i < I; ++i) {
            int c = exg::island_class(island_start[i + 1] - island_start[i], jstart[i + 1] - jstart[i]);
            ids[c].push_back((int)i);
        }
        for (int c = 0; c < exg::ISLAND_CLASSES; ++c) {
            w->class_n[c] = (int)ids[c].size();
            if (ids[c].empty()) continue;
            W_TRY(w->class_ids[c].alloc(ids[c].size()));
            W_TRY(hipMemcpy(w->class_ids[c].p, ids[c].data(), ids[c].size() * sizeof(int), hipMemcpyHostToDevice));
        }
        if (w->class_n[exg::ISLAND_CLASS_M1] > 0 || w->class_n[exg::ISLAND_CLASS_LARGE] > 0)
            w->island_reduce_ok = false;   // free bodies and mechanism-sized islands read ext
        if (w->class_n[exg::ISLAND_CLASS_M1] > 0) {
            W_TRY(w->requeue.alloc(w->class_n[exg::ISLAND_CLASS_M1]));
            W_TRY(w->requeue_ctl.alloc(2));
            W_TRY(hipMemset(w->requeue_ctl.p, 0, 2 * sizeof(unsigned)));
        }
    }
    std::vector<long long> xoff(I);
    long long tot = 0;
    for (size_t i = 0; i < I; ++i) {
        long long mm = island_start[i + 1] - island_start[i];
        xoff[i] = tot;
        tot += mm * mm * 36;
    }
    W_TRY(w->xacc.alloc((size_t)tot));
    if (Jn) {
        W_TRY(w->jkind.alloc(Jn));
        W_TRY(w->ja.alloc(Jn));
        W_TRY(w->jb.alloc(Jn));
        W_TRY(w->jparams.alloc(Jn * 10));
    }
    auto up = [&](void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
    W_TRY(up(w->pose.p, pose, B * 8 * sizeof(double)));
    W_TRY(up(w->mom.p, momentum, B * 6 * sizeof(double)));
    W_TRY(up(w->mass.p, mass, B * sizeof(double)));
    W_TRY(up(w->tens.p, inertia, B * 9 * sizeof(double)));
    W_TRY(up(w->size.p, size, B * sizeof(double)));
    W_TRY(up(w->origin.p, origin, I * 3 * sizeof(double)));
    W_TRY(up(w->island_start.p, island_start, (I + 1) * sizeof(int)));
    W_TRY(up(w->island_jstart.p, jstart.data(), (I + 1) * sizeof(int)));
    W_TRY(up(w->xacc_off.p, xoff.data(), I * sizeof(long long)));
    W_TRY(w->topo.alloc(I));
    W_TRY(w->topo.upload(topo.data(), I));
    if (w->class_n[exg::ISLAND_CLASS_LARGE]) {
        // large islands: a topology record and a workspace each (records of the other
        // islands stay zero and are never read)
        std::vector<exg::LargeTopo> lt(I);
        std::vector<long long> loff(I, 0);
        long long ltot = 0;
        for (size_t i = 0; i < I; ++i) {
            const int mm = island_start[i + 1] - island_start[i], jj = jstart[i + 1] - jstart[i];
            if (exg::island_class(mm, jj) != exg::ISLAND_CLASS_LARGE) {
                std::memset(&lt[i], 0, sizeof(exg::LargeTopo));
                continue;
            }
            std::vector<int> la(jj), lb(jj);
            for (int q = 0; q < jj; ++q) {
                la[q] = a[jstart[i] + q] - island_start[i];
                lb[q] = b[jstart[i] + q] - island_start[i];
            }
            if (exg::build_large_topology(mm, md[i], jj, la.data(), lb.data(), lt[i]) != 0) {
                fail(EXG_ERR_UNSUPPORTED, "exg_world_create: island " + std::to_string(i) + " exceeds the large-island record");
                return bail(EXG_ERR_UNSUPPORTED);
            }
            loff[i] = ltot;
            ltot += exg::large_workspace_doubles(mm, jj);
        }
        W_TRY(w->ltopo.alloc(I));
        W_TRY(w->ltopo.upload(lt.data(), I));
        W_TRY(w->lws.alloc((size_t)ltot));
        W_TRY(w->lws_off.alloc(I));
        W_TRY(w->lws_off.upload(loff.data(), I));
    }
    if (Jn) {
        W_TRY(up(w->jkind.p, kind, Jn * sizeof(int)));
        W_TRY(up(w->ja.p, a, Jn * sizeof(int)));
        W_TRY(up(w->jb.p, b, Jn * sizeof(int)));
        W_TRY(up(w->jparams.p, params, Jn * 10 * sizeof(double)));
    }
    // zero charges (com 0, G·m 0, size 1): nothing gravitates on the first step
    std::vector<double> zc(B * 8, 0.0);
    for (size_t i = 0; i < B; ++i) zc[i * 8 + 7] = 1.0;
    W_TRY(up(w->charges[0].p, zc.data(), B * 8 * sizeof(double)));
    W_TRY(up(w->charges[1].p, zc.data(), B * 8 * sizeof(double)));
    W_TRY(hipMemset(w->xacc.p, 0, (size_t)tot * sizeof(double)));
    W_TRY(hipMemset(w->seeded.p, 0, I * sizeof(int)));
    W_TRY(hipMemset(w->stats.p, 0, I * exg::ISLAND_STATS * sizeof(int)));
    W_TRY(hipMemset(w->fatal.p, 0, sizeof(int)));
    W_TRY(hipMemset(w->ext.p, 0, B * 6 * sizeof(double)));
#undef W_TRY
    *out = w;
    return EXG_OK;
}

int exg_world_destroy(exg_world* w) {
    if (!w) return EXG_OK;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    w->release();
    delete w;
    return EXG_OK;
}

int exg_world_set_solver(exg_world* w, int integrator, double tol2, int ns_warm, int ns_seed) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_solver: null world");
    if (integrator < EXG_EXPLICIT_EULER || integrator > EXG_LIE_EULER)
        return fail(EXG_ERR_INVALID, "exg_world_set_solver: unknown integrator");
    if (ns_warm < 0 || ns_seed < 1) return fail(EXG_ERR_INVALID, "exg_world_set_solver: bad Newton–Schulz budgets");
    w->integrator = integrator;
    w->tol2 = tol2;
    w->settings_epoch++;
    w->ns_warm = ns_warm;
    w->ns_seed = ns_seed;
    return EXG_OK;
}

int exg_world_set_precision(exg_world* w, int gravity, int jacobian) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_precision: null world");
    auto known = [](int p) { return p == EXG_PREC_F64 || p == EXG_PREC_F32_COMPENSATED; };
    if (!known(gravity) || !known(jacobian)) return fail(EXG_ERR_INVALID, "exg_world_set_precision: unknown precision");
    w->grav_prec = gravity;
    w->settings_epoch++;
    w->jac_prec = jacobian;
    return EXG_OK;
}

int exg_world_set_uniform_field(exg_world* w, double ax, double ay, double az) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_uniform_field: null world");
    if (!std::isfinite(ax) || !std::isfinite(ay) || !std::isfinite(az))
        return fail(EXG_ERR_INVALID, "exg_world_set_uniform_field: acceleration is not finite");
    w->field[0] = ax;
    w->field[1] = ay;
    w->field[2] = az;
    w->field_on = ax != 0.0 || ay != 0.0 || az != 0.0;
    w->settings_epoch++;
    return EXG_OK;
}

int exg_world_set_gravity_island_reduce(exg_world* w, int on) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_gravity_island_reduce: null world");
    w->grav_island_reduce = on != 0;
    w->settings_epoch++;
    return EXG_OK;
}

int exg_world_set_ns_engine(exg_world* w, int engine) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_ns_engine: null world");
    if (engine != EXG_NS_MFMA && engine != EXG_NS_VALU_ORDERED)
        return fail(EXG_ERR_INVALID, "exg_world_set_ns_engine: unknown engine " + std::to_string(engine));
    w->ns_engine = engine;
    w->settings_epoch++;
    return EXG_OK;
}

int exg_world_set_joint_mapping(exg_world* w, int mapping) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_joint_mapping: null world");
    if (mapping != EXG_JOINT_MAPPING_SPLIT && mapping != EXG_JOINT_MAPPING_FLAT)
        return fail(EXG_ERR_INVALID, "exg_world_set_joint_mapping: unknown mapping " + std::to_string(mapping));
    w->joint_mapping = mapping;
    w->settings_epoch++;
    return EXG_OK;
}

// Enqueue one world step on the world's stream (no event records: also used inside
// graph capture).
static int enqueue_world_step(exg_world* w, double dt, double warp) {
    double* front = w->charges[w->front].p;
    double* back = w->charges[1 - w->front].p;
    // timing events are recorded by the dispatches themselves (no marker packets)
    hipEvent_t* ev = w->step_ev;
    hipEvent_t g0 = ev ? ev[0] : nullptr, g1 = ev ? ev[1] : nullptr;
    // the island kernels add the f64 partial sums themselves (the reduce kernel is
    // skipped): under the Newton integrator every island in a 4- or 8-body class; the
    // explicit kernel takes every island it takes
    const bool island_reduce = w->gravity && w->grav_island_reduce && w->grav_prec == EXG_PREC_F64 &&
                               w->grav_entries.n > 0 &&
                               (w->integrator == EXG_NEWTON ? w->island_reduce_ok : w->explicit_reduce_ok);
    if (w->gravity)
        exg::launch_gravity(front, w->grav_entries.p, (int)w->grav_entries.n, w->grav_tiles.p, (int)w->grav_tiles.n,
                            w->grav_prec, w->grav_partial.p, w->ext.p, w->grav_stage.p, w->B, w->stream, g0, g1,
                            !island_reduce);
    else
        exg::launch_zero_wrench(w->ext.p, 0, w->B, w->stream, g0, g1);
    HIP_TRY(hipGetLastError());
    if (ev && !(w->gravity ? w->grav_tiles.n > 0 : w->B > 0)) {  // nothing launched: mark the (empty) phase
        HIP_TRY(hipEventRecord(g0, w->stream));
        HIP_TRY(hipEventRecord(g1, w->stream));
    }
    exg::IslandArgs a{};
    a.pose = w->pose.p;
    a.mom = w->mom.p;
    a.mass = w->mass.p;
    a.tens = w->tens.p;
    a.size = w->size.p;
    a.ext = w->ext.p;
    a.island_start = w->island_start.p;
    a.island_end = w->island_start.p + 1;
    a.island_jstart = w->island_jstart.p;
    a.island_jend = w->island_jstart.p + 1;
    a.origin = w->origin.p;
    a.jkind = w->jkind.p;
    a.ja = w->ja.p;
    a.jb = w->jb.p;
    a.jparams = w->jparams.p;
    a.topo = w->topo.p;
    a.ltopo = w->ltopo.p;
    a.lws = w->lws.p;
    a.lws_off = w->lws_off.p;
    a.xacc = w->xacc.p;
    a.xacc_off = w->xacc_off.p;
    a.seeded = w->seeded.p;
    a.charges_out = back;
    a.stats = w->stats.p;
    a.fatal = w->fatal.p;
    a.requeue = w->requeue.p;
    a.requeue_ctl = w->requeue_ctl.p;
    a.dt = dt;
    a.warp = warp;
    a.tol2 = w->tol2;
    a.G = w->G;
    a.ns_warm = w->ns_warm;
    a.ns_seed = w->ns_seed;
    a.ns_engine = w->ns_engine;
    a.jac_f32 = w->jac_prec == EXG_PREC_F32_COMPENSATED;
    a.n_islands = w->I;
    for (int k = 0; k < 3; ++k) a.field[k] = w->field[k];
    a.field_on = w->field_on ? 1 : 0;
    a.grav_tiles = island_reduce ? w->grav_tiles.p : nullptr;
    a.grav_partial = w->grav_partial.p;
    a.grav_charges = front;
    hipEvent_t i0 = ev ? ev[2] : nullptr, i1 = ev ? ev[3] : nullptr;
    if (w->integrator == EXG_NEWTON) {
        int* ids[exg::ISLAND_CLASSES];
        int n[exg::ISLAND_CLASSES];
        bool any = false;
        for (int c = 0; c < exg::ISLAND_CLASSES; ++c) {
            ids[c] = w->class_ids[c].p;
            n[c] = w->class_n[c];
            any |= n[c] > 0;
        }
        if (w->joint_mapping == EXG_JOINT_MAPPING_FLAT) {   // the m8j14 islands on the other mapping
            ids[exg::ISLAND_CLASS_M8J14_FLAT] = ids[exg::ISLAND_CLASS_M8J14];
            n[exg::ISLAND_CLASS_M8J14_FLAT] = n[exg::ISLAND_CLASS_M8J14];
            n[exg::ISLAND_CLASS_M8J14] = 0;
        }
        const exg::ClassLaunch cl{n, ids, w->stream, w->classes.side, w->classes.fork, w->classes.join};
        HIP_TRY(exg::launch_classes(a, cl, i0, i1));
        if (ev && !any) {
            HIP_TRY(hipEventRecord(i0, w->stream));
            HIP_TRY(hipEventRecord(i1, w->stream));
        }
    } else {
        a.island_ids = nullptr;
        exg::launch_explicit_step(a, w->integrator, w->stream, i0, i1);
    }
    HIP_TRY(hipGetLastError());
    w->front = 1 - w->front;
    return EXG_OK;
}

// Graph of two consecutive steps starting from charge buffer `front` (one round trip
// of the double buffer, so the graph replays without re-binding pointers).
static int world_graph(exg_world* w, double dt, double warp, hipGraphExec_t* out) {
    auto& g = w->graphs[w->front];
    if (g.exec && g.dt == dt && g.warp == warp && g.settings == w->settings_epoch) {
        *out = g.exec;
        return EXG_OK;
    }
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    const int front0 = w->front;
    w->step_ev = nullptr;
    HIP_TRY(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_world_step(w, dt, warp);
    if (rc == EXG_OK) rc = enqueue_world_step(w, dt, warp);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(w->stream, &graph);
    w->front = front0;
    if (rc != EXG_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(EXG_ERR_BACKEND, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(EXG_ERR_BACKEND, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    g.dt = dt;
    g.warp = warp;
    g.settings = w->settings_epoch;
    *out = g.exec;
    return EXG_OK;
}

int exg_world_step(exg_world* w, int n_steps, double dt, double warp) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_step: null world");
    if (n_steps < 0) return fail(EXG_ERR_INVALID, "exg_world_step: negative step count");
    // the engine's input facts (crates/newton/build.rs input_facts): dt > 0, warp ≥ 1
    if (!(dt > 0.0) || !std::isfinite(dt) || !(warp >= 1.0) || !std::isfinite(warp))
        return fail(EXG_ERR_INVALID, "exg_world_step: dt must be positive and warp at least 1 (both finite)");
    if (w->integrator != EXG_NEWTON && (w->max_m > EXPLICIT_MAX_BODIES || w->max_j > EXPLICIT_MAX_JOINTS))
        return fail(EXG_ERR_UNSUPPORTED, "exg_world_step: the explicit integrators take islands of up to 8 bodies "
                                         "and 16 joints");
    w->ev_used = 0;
    w->graph_ms_valid = false;
    if (w->use_graph && n_steps >= 2) {
        // whole-call timing only: the graph replays two steps per launch
        while (w->ev.size() < 2) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            w->ev.push_back(e);
        }
        hipGraphExec_t exec = nullptr;
        if (int rc = world_graph(w, dt, warp, &exec)) return rc;
        HIP_TRY(hipEventRecord(w->ev[0], w->stream));
        for (int k = 0; k + 1 < n_steps; k += 2) HIP_TRY(hipGraphLaunch(exec, w->stream));
        if (n_steps % 2) {
            w->step_ev = nullptr;
            if (int rc = enqueue_world_step(w, dt, warp)) return rc;
        }
        HIP_TRY(hipEventRecord(w->ev[1], w->stream));
        w->graph_ms_valid = true;
        return EXG_OK;
    }
    size_t need = (size_t)n_steps * 4;
    while (w->ev.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        w->ev.push_back(e);
    }
    for (int s = 0; s < n_steps; ++s) {
        w->step_ev = &w->ev[4 * (size_t)s];
        const int rc = enqueue_world_step(w, dt, warp);
        w->step_ev = nullptr;
        if (rc) return rc;
        w->ev_used = s + 1;
    }
    return EXG_OK;
}

int exg_world_set_graph(exg_world* w, int enable) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_set_graph: null world");
    w->use_graph = enable != 0;
    return EXG_OK;
}

// Island-steps since creation whose state ended non-finite (after a synchronization).
static int world_fatal_count(exg_world* w, long long& count) {
    int c = 0;
    HIP_TRY(hipStreamSynchronize(w->stream));
    HIP_TRY(hipMemcpy(&c, w->fatal.p, sizeof(int), hipMemcpyDeviceToHost));
    count = c;
    return EXG_OK;
}

int exg_world_sync(exg_world* w) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_sync: null world");
    long long c = 0;
    if (int rc = world_fatal_count(w, c)) return rc;
    if (c > w->fatal_reported) {
        const long long fresh = c - w->fatal_reported;
        w->fatal_reported = c;
        return fail(EXG_ERR_FATAL, "exg_world_sync: " + std::to_string(fresh) +
                                       " island-step(s) ended with a non-finite state (per-island counts in "
                                       "exg_world_stats slot 6)");
    }
    return EXG_OK;
}

int exg_world_fatals(exg_world* w, int64_t* count) {
    if (!w || !count) return fail(EXG_ERR_INVALID, "exg_world_fatals: null argument");
    long long c = 0;
    if (int rc = world_fatal_count(w, c)) return rc;
    *count = c;
    return EXG_OK;
}

int exg_world_last_timing(exg_world* w, double* out2) {
    if (!w || !out2) return fail(EXG_ERR_INVALID, "exg_world_last_timing: null argument");
    HIP_TRY(hipStreamSynchronize(w->stream));
    if (w->graph_ms_valid) {  // graph replay: one total, not split by kernel
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, w->ev[0], w->ev[1]));
        out2[0] = -1.0;
        out2[1] = t;
        return EXG_OK;
    }
    double g = 0.0, s = 0.0;
    for (int k = 0; k < w->ev_used; ++k) {
        float a = 0.f, b = 0.f;
        HIP_TRY(hipEventElapsedTime(&a, w->ev[4 * k], w->ev[4 * k + 1]));
        HIP_TRY(hipEventElapsedTime(&b, w->ev[4 * k + 2], w->ev[4 * k + 3]));
        g += a;
        s += b;
    }
    out2[0] = g;
    out2[1] = s;
    return EXG_OK;
}

int exg_world_download(exg_world* w, double* pose, double* momentum, double* origin) {
    if (!w) return fail(EXG_ERR_INVALID, "exg_world_download: null world");
    HIP_TRY(hipStreamSynchronize(w->stream));
    if (pose) HIP_TRY(hipMemcpy(pose, w->pose.p, (size_t)w->B * 8 * sizeof(double), hipMemcpyDeviceToHost));
    if (momentum) HIP_TRY(hipMemcpy(momentum, w->mom.p, (size_t)w->B * 6 * sizeof(double), hipMemcpyDeviceToHost));
    if (origin) HIP_TRY(hipMemcpy(origin, w->origin.p, (size_t)w->I * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return EXG_OK;
}

int exg_world_charges(exg_world* w, double* charges) {
    if (!w || !charges) return fail(EXG_ERR_INVALID, "exg_world_charges: null argument");
    HIP_TRY(hipStreamSynchronize(w->stream));
    HIP_TRY(hipMemcpy(charges, w->charges[w->front].p, (size_t)w->B * 8 * sizeof(double), hipMemcpyDeviceToHost));
    return EXG_OK;
}

int exg_world_external(exg_world* w, double* ext) {
    if (!w || !ext) return fail(EXG_ERR_INVALID, "exg_world_external: null argument");
    HIP_TRY(hipStreamSynchronize(w->stream));
    HIP_TRY(hipMemcpy(ext, w->ext.p, (size_t)w->B * 6 * sizeof(double), hipMemcpyDeviceToHost));
    return EXG_OK;
}

int exg_world_stats(exg_world* w, int32_t* stats) {
    if (!w || !stats) return fail(EXG_ERR_INVALID, "exg_world_stats: null argument");
    HIP_TRY(hipStreamSynchronize(w->stream));
    HIP_TRY(hipMemcpy(stats, w->stats.p, (size_t)w->I * exg::ISLAND_STATS * sizeof(int), hipMemcpyDeviceToHost));
    return EXG_OK;
}

int exg_gravity_accumulate(int n, const double* charges, double G, double* wrench) {
    return exg_gravity_accumulate_precision(n, charges, G, EXG_PREC_F64, wrench);
}

int exg_gravity_accumulate_precision(int n, const double* charges, double G, int precision, double* wrench) {
    if (n <= 0 || !charges || !wrench) return fail(EXG_ERR_INVALID, "exg_gravity_accumulate: bad arguments");
    if (precision != EXG_PREC_F64 && precision != EXG_PREC_F32_COMPENSATED)
        return fail(EXG_ERR_INVALID, "exg_gravity_accumulate: unknown precision");
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    std::vector<exg::GravEntry> entries;
    std::vector<exg::GravTile> tiles;
    exg::gravity_add_world(entries, tiles, 0, n, G, cus);
    DevBuf<double> dc, dw, dp, ds;
    DevBuf<exg::GravEntry> de;
    DevBuf<exg::GravTile> dt;
    hipError_t e = dc.alloc((size_t)n * 8);
    if (e == hipSuccess) e = dw.alloc((size_t)n * 6);
    if (e == hipSuccess) e = ds.alloc((size_t)n * exg::GRAV_STAGE_BYTES / sizeof(double));
    if (e == hipSuccess) e = dp.alloc(entries.size() * exg::GRAV_PARTIAL_PER_ENTRY);
    if (e == hipSuccess) e = de.alloc(entries.size());
    if (e == hipSuccess) e = dt.alloc(tiles.size());
    if (e == hipSuccess) e = dc.upload(charges, (size_t)n * 8);
    if (e == hipSuccess) e = de.upload(entries.data(), entries.size());
    if (e == hipSuccess) e = dt.upload(tiles.data(), tiles.size());
    if (e == hipSuccess) {
        exg::launch_gravity(dc.p, de.p, (int)entries.size(), dt.p, (int)tiles.size(), precision, dp.p, dw.p, ds.p, n,
                            nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(wrench, dw.p, (size_t)n * 6 * sizeof(double), hipMemcpyDeviceToHost);
    for (auto* b : {&dc, &dw, &dp, &ds}) b->release();
    de.release();
    dt.release();
    if (e != hipSuccess) return fail(EXG_ERR_BACKEND, std::string("exg_gravity_accumulate: ") + hipGetErrorString(e));
    return EXG_OK;
}

}  // extern "C"
