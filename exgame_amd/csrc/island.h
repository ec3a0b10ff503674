// Launch interface of the device kernels of the dynamics hot path.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <vector>

namespace exg {

// Record the calling thread's last error (read back by exg_last_error) and return `code`.
int set_error(int code, const char* msg);

// Validates a world's topology (bodies per island, joints grouped by island with
// both ends inside it, known kinds, on-chip size limits); fills the joint prefix
// jstart[I+1] and the largest island. Returns EXG_OK or the error (message recorded).
int check_world(const char* who, int n_bodies, int n_islands, const int* island_start, int n_joints,
                const int* kind, const int* a, const int* b, std::vector<int>& jstart, int& max_m, int& max_j);

// Validates the state of a world's bodies, islands and joints: finite masses, positive
// or zero (zero: a kinematic body, Inertia::Kinematic, whose momentum slot holds its
// body-frame velocity and whose inertia tensor is not read), positive definite finite
// inertia tensors of the others, finite poses with a rotor part, finite momenta, sizes,
// origins, joint parameters and G. Returns EXG_OK or the error (message recorded).
int check_bodies(const char* who, int n_bodies, const double* pose, const double* momentum, const double* mass,
                 const double* inertia, const double* size, int n_islands, const double* origin, int n_joints,
                 const int* kind, const double* params, double G);

// Dynamic bodies of each island (md[i]): the kernels take an island's kinematic bodies
// after its dynamic ones, so that the unknowns of its system are its first md bodies
// (the reference keeps the kinematic ones out of the system through an order map,
// crates/newton/src/integrator/implicit/newton.rs:806-830). An island with a dynamic
// body after a kinematic one gives EXG_ERR_INVALID. Returns EXG_OK or the error.
int dynamic_counts(const char* who, int n_islands, const int* island_start, const double* mass, std::vector<int>& md);

// Tables of one island that depend only on its joint graph (neighbour lists, assembly
// terms, incidence, joints by kind, A's non-zero slab masks and the MFMA tile lists
// of each wave), built on the host by build_island_topology (topology.cpp) when a
// world is set up, for islands of up to TOPO_MAXM bodies and TOPO_MAXJ joints.
constexpr int TOPO_MAXM = 8, TOPO_MAXJ = 16, TOPO_WAVES = 4, TOPO_PER_WAVE = 4, TOPO_MAXT = 3;
struct IslandTopo {
    uint16_t kmask[4];                              // per 16-row tile: 4-column slabs that can be non-zero
    uint8_t tcount[2][TOPO_WAVES];                  // [0] sparse product A·X, [1] dense product X·T
    uint8_t tlist[2][TOPO_WAVES][TOPO_PER_WAVE];    // row tile | column tile << 4
    uint8_t nbr_start[TOPO_MAXM + 1], nbr[TOPO_MAXM * TOPO_MAXM];
    uint8_t term_start[TOPO_MAXM * TOPO_MAXM + 1], terms[4 * TOPO_MAXJ];   // joint << 2 | row end << 1 | column end
    uint8_t inc_start[TOPO_MAXM + 1], inc[2 * TOPO_MAXJ];                 // joint << 1 | end
    uint8_t jorder[TOPO_MAXJ];
    uint8_t pad[13];
};
static_assert(sizeof(IslandTopo) == 320, "IslandTopo is a 320-byte record");
// m bodies of which the first md are dynamic, J joints with island-local ends ja/jb and
// kinds; `waves` and `per_wave` from island_class_shape. The system's tables (neighbour
// lists, assembly terms, slab masks, tile lists) cover the dynamic bodies only; the
// incidence lists cover every body. Returns 0, or -1 when the island exceeds the record.
int build_island_topology(int m, int md, int J, const int* ja, const int* jb, const int* kind, int waves,
                          int per_wave, IslandTopo& t);

// The same tables for mechanism-sized islands (more than 8 bodies or 16 joints, up
// to LARGE_MAXM bodies and LARGE_MAXJ joints), one record per such island. Neighbour
// entries are indexed in block-row order (the non-zero blocks of A); assembly terms
// are listed per neighbour entry.
constexpr int LARGE_MAXM = 48, LARGE_MAXJ = 128, LARGE_MAXN = 6 * LARGE_MAXM, LARGE_MAXT = LARGE_MAXN / 16;
constexpr int LARGE_MAXNBR = LARGE_MAXM + 2 * LARGE_MAXJ;
struct LargeTopo {
    uint16_t nbr_start[LARGE_MAXM + 1];
    uint8_t nbr[LARGE_MAXNBR], nbr_row[LARGE_MAXNBR];     // column and row block of each non-zero block
    uint16_t term_start[LARGE_MAXNBR + 1], terms[4 * LARGE_MAXJ];   // joint << 2 | row end << 1 | column end
    uint16_t inc_start[LARGE_MAXM + 1], inc[2 * LARGE_MAXJ];       // joint << 1 | end
    uint32_t kmask[LARGE_MAXT][3];                        // per 16-row tile: its non-zero 4-column slabs
};
// m bodies of which the first md are dynamic (A's blocks, and so the neighbour lists,
// assembly terms and slab masks, cover those; the incidence covers every body), J
// joints with island-local ends; returns 0, or -1 beyond the record.
int build_large_topology(int m, int md, int J, const int* ja, const int* jb, LargeTopo& t);
// Doubles of device workspace one large island of m bodies and J joints uses: the
// system matrices A, X, X', T (n16 × n16 each, n16 = 6m rounded up to 16) and the
// joint Jacobians (J × 288).
long long large_workspace_doubles(int m, int J);

constexpr int ISLAND_THREADS = 256;
// Per-island counters: spans, Newton iterations, NS iterations, reseeds, line-search
// rejects, unconverged commits, steps that ended with a non-finite state (the fatal
// mark), spans found non-finite at the depth limit and not committed.
constexpr int ISLAND_STATS = 8;
constexpr int STAT_FATAL = 6, STAT_NONFINITE_SPANS = 7;

struct GravTile;

// One implicit step of every island (one workgroup per island).
struct IslandArgs {
    double* pose;             // B×8 motors, island frame
    double* mom;              // B×6 body-frame momenta
    const double* mass;       // B
    const double* tens;       // B×9 body-frame inertia tensor
    const double* size;       // B effective size
    const double* ext;        // B×6 external wrench, island frame
    const int* island_start;  // first body of each island
    const int* island_end;    // one past its last body (a world passes island_start + 1)
    const int* island_jstart; // first joint of each island
    const int* island_jend;   // one past its last joint
    double* origin;           // I×3
    const int* jkind;         // J
    const int* ja;            // J, global body index
    const int* jb;            // J
    const double* jparams;    // J×10
    double* xacc;             // accepted approximate inverses, (6m)² per island, row-major
    const long long* xacc_off;// I
    int* seeded;              // I
    double* charges_out;      // B×8: com xyz, local xyz, G·m, size
    int* stats;               // I×ISLAND_STATS
    double dt, warp, tol2, G;
    int ns_warm, ns_seed, n_islands;
    int ns_engine;            // 0: Newton–Schulz products on MFMA; 1: VALU in the oracle's summation order
    const int* island_ids;    // islands this launch covers (n_islands of them); null: 0 .. n_islands-1
    const double* G_of_island;// per-island gravitational constant for the published charges; null: G
    int jac_f32;              // 1: FP32 Jacobian lanes and compensated FP32 assembly (mixed precision)
    const IslandTopo* topo;   // I records (islands of the single-body class need none)
    int* fatal;               // island-steps that ended with a non-finite state (one counter); may be null
    const LargeTopo* ltopo;   // per island, used by the large class only (others: unused entries)
    double* lws;              // workspace of the large islands
    const long long* lws_off; // per island: its offset into lws (large class only)
    int* requeue;             // single-body islands whose step subdivides, for the requeue kernel (capacity:
                              // the single-body islands of the launch); null: the one-lane kernel solves all
    unsigned* requeue_ctl;    // [0] the list's length, [1] requeue workgroups done; both 0 between launches
    // gravity reduced by the islands (4- and 8-body classes, f64 pairs): with grav_tiles
    // set, each island adds its bodies' partial sums (grav_partial, from the gravity
    // pass's partial kernel) and forms their wrenches from the published charges
    // (grav_charges) itself, and ext is not read; null: ext holds the wrenches
    const GravTile* grav_tiles;
    const double* grav_partial;
    const double* grav_charges;
    double field[3];          // UniformField acceleration (world frame), applied when field_on
    int field_on;
};

// UniformField (crates/newton/src/field.rs:16-44): a constant world acceleration adds
// F = m·a and a zero torque to a body's wrench after gravity (a kinematic body, mass 0,
// gets nothing). Wrench component k of a body of mass m; no add at all when the field
// is off, so that the wrenches keep their bits.
__device__ inline double field_add(const IslandArgs& a, double m, int k, double w) {
    return a.field_on ? w + (k < 3 ? a.field[k] * m : 0.0) : w;
}

// The f64 pair law's 1/sqrt: the hardware estimate (≈ 2^-24 relative) refined by one
// third-order step, y = y0 + y0·e·(1/2 + 3/8·e) with e = 1 − x·y0²: within 2 ulp of
// 1/sqrt(x) over 1e-40 … 2e30 (the bound of two Newton steps, in five operations
// instead of eight). Used by the gravity pass and by the islands' fused gravity.
__device__ inline double rsqrt_refined(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y0), y0, 1.0);
    const double c = fma(0.375, e, 0.5);
    return fma(y0 * e, c, y0);
}

// Mark-and-continue fatal report of one island-step (the state is written as it is):
// the island's STAT_FATAL counter and the launch-wide counter.
__device__ inline void mark_fatal(const IslandArgs& a, int isl) {
    a.stats[(size_t)isl * ISLAND_STATS + STAT_FATAL] += 1;
    if (a.fatal) atomicAdd(a.fatal, 1);
}

// Island classes of the fused Newton kernel and the launch of one class over the
// island ids in a.island_ids[0 .. a.n_islands).
// ISLAND_CLASS_M8J14_FLAT runs the islands of the m8j14 class with the other joint
// mapping (EXG_JOINT_MAPPING_FLAT): island_class never returns it; a world or an
// arena world that asks for that mapping routes its m8j14 islands there.
enum {
    ISLAND_CLASS_M1 = 0, ISLAND_CLASS_M4J10 = 1, ISLAND_CLASS_M8J14 = 2, ISLAND_CLASS_M8J16 = 3, ISLAND_CLASS_LARGE = 4,
    ISLAND_CLASS_M8J14_FLAT = 5, ISLAND_CLASSES = 6
};
int island_class(int m, int joints);   // -1 when no variant holds the island
// Waves per workgroup and MFMA tiles per wave of a class's kernel (for its tile lists).
void island_class_shape(int cls, int& waves, int& per_wave);
// build_island_topology for island i of a world: bodies [b0, b0 + m), the first md of
// them dynamic, joints with global body indices a/b; routes by island_class.
int island_topology_of(int m, int md, int J, const int* kind, const int* a, const int* b, int b0, IslandTopo& t);
// One kernel launch; with a start or stop event the dispatch itself records them
// (hipExtLaunchKernelGGL), otherwise a plain launch (which is also what stream
// capture records).
template <typename F, typename... A>
inline void launch_k(F kernel, dim3 grid, dim3 block, hipStream_t stream, hipEvent_t start, hipEvent_t stop, A... args) {
    if (start || stop) hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, start, stop, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
}

// Launchers take optional start/stop events that the dispatch itself records
// (hipExtLaunchKernelGGL: no marker packets between kernels); null records nothing.
void launch_island_class(const IslandArgs& a, int cls, hipStream_t stream, hipEvent_t start = nullptr,
                         hipEvent_t stop = nullptr);
// Every island class of one step, run side by side: the classes with islands are
// launched heaviest first (large, m8j14, m8j14 flat, m8j16, m4j10, singles), the first on `main`
// and each further one on its own side stream after `fork` (recorded on main), and
// `main` then waits for `join[c]` of each. Classes hold disjoint islands, so their
// order does not change any result. `class_ids[c]` lists class c's islands. With a
// single class, `start`/`stop` are recorded by its dispatch; with several, `start` by
// the first dispatch and `stop` on main after the joins. Also valid under stream
// capture (the fork and joins become graph edges).
struct ClassLaunch {
    const int* n;               // islands per class [ISLAND_CLASSES]
    int* const* ids;            // island ids per class
    hipStream_t main;
    const hipStream_t* side;    // one per class
    hipEvent_t fork;
    const hipEvent_t* join;     // one per class
};
hipError_t launch_classes(IslandArgs a, const ClassLaunch& c, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// The side streams (default priority) and fork/join events (no timing) of launch_classes.
struct ClassStreams {
    hipStream_t side[ISLAND_CLASSES] = {};
    hipEvent_t fork = nullptr, join[ISLAND_CLASSES] = {};
    hipError_t create();
    void destroy();
};

// Mechanism-sized islands, one workgroup of LARGE_THREADS per island (island_large.hip).
constexpr int LARGE_THREADS = 512;
void launch_island_large(const IslandArgs& a, hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// Single-body islands without joints (singles.hip): one lane per island, and, with a
// requeue list, eight lanes per island for those whose step subdivides (a second launch
// on the same stream).
void launch_singles(const IslandArgs& a, hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);

// Explicit integrators (1 = SymplecticEuler, 0 = ExplicitEuler), one thread per island.
void launch_explicit_step(const IslandArgs& a, int integrator, hipStream_t stream, hipEvent_t start = nullptr,
                          hipEvent_t stop = nullptr);

// Pairwise gravity over a work table (gravity.hip): entries pair a tile of ≤ 64
// targets with a chunk of sources of the same world; tiles list their entries in
// ascending chunk order. gravity_add_world appends one world's bodies [b0, b0+n)
// with a chunking that depends only on n (and the CU count).
struct GravEntry {
    int tgt0, ntgt, src0, src1;
};
struct GravTile {
    int tgt0, ntgt, e0, ne;
    double g_inv;
};
constexpr int GRAV_TILE = 64;                     // targets per tile
constexpr int GRAV_PARTIAL_PER_ENTRY = 3 * GRAV_TILE;   // doubles of partial sums per entry
constexpr int GRAV_ISLAND_REDUCE_MAX_CHUNKS = 64;        // chunks per tile the islands' reduction holds
// A target's wrench from its summed pair force f (Σ_j G·m_j·r⁻³·(c_j − c_i)): scaled by
// its charge G·m over G, with the moment about its island origin from its lever arm
// (charge record [x y z | lx ly lz | G·m | size]). Shared by the reduce kernel and the
// islands' reduction so that both give the same bits.
__device__ inline void gravity_wrench(const double* own, double g_inv, double f[3], double* out) {
    const double a = own[6] * g_inv;
    for (int k = 0; k < 3; ++k) f[k] *= a;
    const double lx = own[3], ly = own[4], lz = own[5];
    out[0] = f[0];
    out[1] = f[1];
    out[2] = f[2];
    out[3] = ly * f[2] - lz * f[1];
    out[4] = lz * f[0] - lx * f[2];
    out[5] = lx * f[1] - ly * f[0];
}
void gravity_add_world(std::vector<GravEntry>& entries, std::vector<GravTile>& tiles, int b0, int n, double G,
                       int compute_units);
// The chunk length gravity_add_world uses for a world of n bodies (a pure function).
int gravity_chunk_len(int n, int compute_units);
// The same with a fixed chunk length (≥ 64 sources).
void gravity_add_world_chunked(std::vector<GravEntry>& entries, std::vector<GravTile>& tiles, int b0, int n, double G,
                               int chunk_len);
// precision 0: f64 pairs; 1: f32 pairs on double-single positions, f64 accumulation.
// `stage` holds GRAV_STAGE_BYTES per body for bodies [0, n_bodies) (every body an
// entry names): the per-source records the f32 kernel reads.
constexpr int GRAV_STAGE_BYTES = 64;
// With reduce = false (f64 pairs, entries present) only the partial kernel runs: the
// islands add the partial sums themselves (IslandArgs::grav_tiles).
void launch_gravity(const double* charges, const GravEntry* entries, int n_entries, const GravTile* tiles,
                    int n_tiles, int precision, double* partial, double* ext, void* stage, int n_bodies,
                    hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr, bool reduce = true);
void launch_zero_wrench(double* ext, int b0, int n, hipStream_t stream, hipEvent_t start = nullptr,
                        hipEvent_t stop = nullptr);

}  // namespace exg
