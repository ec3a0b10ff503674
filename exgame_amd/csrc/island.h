// Launch interface of the device kernels of the dynamics hot path.
#pragma once
#include <hip/hip_runtime.h>

namespace exg {

// Record the calling thread's last error (read back by exg_last_error) and return `code`.
int set_error(int code, const char* msg);

constexpr int ISLAND_THREADS = 256;
constexpr int ISLAND_STATS = 8;  // spans, newton iterations, NS iterations, reseeds, line-search rejects, unconverged commits

// One implicit step of every island (one workgroup per island).
struct IslandArgs {
    double* pose;             // B×8 motors, island frame
    double* mom;              // B×6 body-frame momenta
    const double* mass;       // B
    const double* tens;       // B×9 body-frame inertia tensor
    const double* size;       // B effective size
    const double* ext;        // B×6 external wrench, island frame
    const int* island_start;  // I+1
    const int* island_jstart; // I+1
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
};

// Island classes of the fused Newton kernel and the launch of one class over the
// island ids in a.island_ids[0 .. a.n_islands).
enum { ISLAND_CLASS_M1 = 0, ISLAND_CLASS_M4J8 = 1, ISLAND_CLASS_M8J14 = 2, ISLAND_CLASS_M8J16 = 3, ISLAND_CLASSES = 4 };
int island_class(int m, int joints);   // -1 when no variant holds the island
void launch_island_class(const IslandArgs& a, int cls, hipStream_t stream);

// Explicit integrators (1 = SymplecticEuler, 0 = ExplicitEuler), one thread per island.
void launch_explicit_step(const IslandArgs& a, int integrator, hipStream_t stream);

// Pairwise gravity: ext[i] = (Σ_j F_ij, local_i × Σ_j F_ij), charges B×8. The
// sources are split into `chunks` chunks of `chunk_len`; partial holds
// partial_doubles doubles of per-chunk sums.
struct GravityPlan {
    int tiles, chunks, chunk_len;
    size_t partial_doubles;
};
GravityPlan gravity_plan(int n, int compute_units);
void launch_gravity_f64(const double* charges, int n, double g_inv, const GravityPlan& plan, double* partial,
                        double* ext, hipStream_t stream);
// Mixed precision: FP32 pairs with compensated accumulation; `narrow` holds 6n floats
// (x, y, z, G·m | s², 1/s per body, 16-byte aligned).
void launch_gravity_f32(const double* charges, int n, double g_inv, const GravityPlan& plan, double* partial,
                        float* narrow, double* ext, hipStream_t stream);

}  // namespace exg
