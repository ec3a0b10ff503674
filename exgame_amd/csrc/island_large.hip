// The implicit-midpoint Newton step of one mechanism-sized island per workgroup:
// islands of more than 8 bodies or 16 joints, up to 48 bodies (a 288 × 288 system)
// and 128 joints — cloth grids, long chains, rigs of rigs.
//
// The control flow is the reference's host loop, as in island.hip (newton.rs:209-753,
// cache.rs:1115-1136): Newton iterations with an 8-lane backtracking line search,
// Newton–Schulz on the block system with a transpose seed, the ×4 divergence guard
// and one reseed, dt subdivision by the stiffness signal, the midpoint commit and the
// accepted-inverse checkpoint. What differs is where the data lives and how the
// work is spread over the workgroup's 512 threads:
//
//   * the system matrices A, X, X' and T = 2I − A·X (n16 × n16 doubles each, 663 KB
//     at 48 bodies) and the joint Jacobians live in a per-island device workspace
//     (L2-resident during the step); the per-body state, the current iterate, one
//     line-search candidate, the connection wrenches and the topology tables live in
//     LDS (≈ 100 KB);
//   * the Newton–Schulz products run on FP64 MFMA 16×16×4 tiles, each wave taking
//     every eighth output tile; R = A·X skips the 4-column slabs of A that the
//     island's block sparsity makes zero; X·T goes to X' and the two swap (no
//     in-place product);
//   * line-search candidates are evaluated one at a time, in lane order, and the
//     first that lowers the residual is taken — the same choice as evaluating all
//     eight together.
// Every control decision is taken by all threads from values in LDS, so branches
// are uniform across the workgroup.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "island.h"
#include "physics_dev.h"
#include "body_rows.h"

namespace exg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NT = LARGE_THREADS;
constexpr int WAVES = NT / 64;
constexpr int MAXM = LARGE_MAXM, MAXJ = LARGE_MAXJ;
constexpr int LANES = 8;
constexpr int MAX_ITERATIONS = 32;
constexpr int MAX_DEPTH = 8;
constexpr int THRESHOLD_LOG2 = 8;
constexpr int PER_HALVING_LOG2 = 4;
static_assert(8 * MAXM <= NT, "eight threads per body in one pass");
static_assert(LARGE_MAXN <= NT, "one thread per row of the system in one pass");

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v) {
    long long b = __double_as_longlong(v);
    unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Per-body arrays of one evaluation point: slot 0 is the Newton iterate, slot 1 the
// line-search candidate under test.
struct Slot {
    double vmid[MAXM][6], mid[MAXM][8], total[MAXM][6], rhs[MAXM][6], scale[MAXM][2];
    double conn[MAXJ][2][6];
};

struct LargeShared {
    Slot sl[2];
    double mass0[MAXM][36];     // the iterate's world mass blocks (the oracle's lanes[0].mass)
    double pose[MAXM][8], mom[MAXM][6], massv[MAXM], tens[MAXM][9], ext[MAXM][6];
    double snap[MAXM][6], dv[MAXM][6], base[MAXM][6], best[MAXM][6];
    double jparams[MAXJ][10];
    double ratio[MAXM][2];      // per body: lin²/scale_lin, ang²/scale_ang of the slot just evaluated
    double wn[2];
    double red[WAVES];
    double seedsum[2 * MAXM];
    double bcast[8];
    int jkind[MAXJ], ja[MAXJ], jb[MAXJ];
    LargeTopo tp;
    int ibcast[8];              // [0]: a non-zero Newton correction was written; [1]: seeded flag out; [2]: non-finite
};

__device__ double block_sum(LargeShared& s, double v) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s.red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = s.red[0];
#pragma unroll
    for (int w = 1; w < WAVES; ++w) r = r + s.red[w];
    return uni(r);
}

struct LargeSolver {
    LargeShared& s;
    const IslandArgs& a;
    int m, J, n, n16, LD, nnz;
    int md;   // dynamic bodies (the unknowns), the island's first md; the rest are kinematic references
    double dt, warp, half, span;
    long long xoff;
    double *A, *X, *Xn, *T, *jac;   // workspace
    int st_spans, st_newton, st_ns, st_reseed, st_reject, st_uncommit, st_nonfinite;
    bool seeded;

    static __device__ __forceinline__ int tid() {
        int t = (int)threadIdx.x;
        asm volatile("" : "+v"(t));
        return t;
    }
    __device__ double* JAC(int j, int e, int ax) { return jac + (size_t)j * 288 + e * 144 + ax * 6; }

    // midpoint = pose ∘ exp(half·vmid) for slot l
    __device__ void pre(int l) {
        for (int i = tid(); i < m; i += NT) {
            double e[8];
            twist_exp(s.sl[l].vmid[i], half, e);
            compose(s.pose[i], e, s.sl[l].mid[i]);
        }
    }
    // Jacobian columns of every joint at the iterate; the f64 columns of axis 0 also
    // deliver the wrench values
    __device__ void joints_jacobian() {
        for (int t = tid(); t < J * 24; t += NT) {
            const int j = t / 24, ax = t % 24, ea = s.ja[j], eb = s.jb[j];
            const bool values = ax == 0;
            joint_column<double>(s.jkind[j], s.jparams[j], s.sl[0].mid[ea], s.sl[0].mid[eb], s.sl[0].vmid[ea],
                                 s.sl[0].vmid[eb], ax, dt, warp, JAC(j, 0, ax), JAC(j, 1, ax),
                                 values ? s.sl[0].conn[j][0] : nullptr, values ? s.sl[0].conn[j][1] : nullptr);
        }
    }
    __device__ void joints_plain(int l) {
        for (int j = tid(); j < J; j += NT)
            joint_wrench(s.jkind[j], s.jparams[j], s.sl[l].mid[s.ja[j]], s.sl[l].mid[s.jb[j]], s.sl[l].vmid[s.ja[j]],
                         s.sl[l].vmid[s.jb[j]], dt, warp, s.sl[l].conn[j][0], s.sl[l].conn[j][1]);
    }
    static __device__ double max2(double x, double y) { return x > y ? x : y; }
    // gather (external + incident in joint order) and body_post for slot l: eight
    // threads per body, one mass-block row each; the worst residual ratio in body
    // order (newton.rs:139-151) into s.wn[l] (after the next barrier)
    __device__ void gather_post(int l, bool store_mass) {
        const int t = tid(), i = t >> 3, r = t & 7;
        if (i < m) {
            double tot[6];
            for (int k = 0; k < 6; ++k) tot[k] = s.ext[i][k];
            for (int q = s.tp.inc_start[i]; q < s.tp.inc_start[i + 1]; ++q) {
                const int code = s.tp.inc[q];
                const double* w = s.sl[l].conn[code >> 1][code & 1];
                for (int k = 0; k < 6; ++k) tot[k] = tot[k] + w[k];
            }
            const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
            double brow[6], rhs[6], sc[2];
            body_post_row(s.sl[l].mid[i], s.sl[l].vmid[i], s.massv[i], s.tens[i], s.snap[i], tot, half, floor2,
                          r < 6 ? r : 5, brow, rhs, sc);
            if (r == 0) {
                for (int k = 0; k < 6; ++k) {
                    s.sl[l].total[i][k] = tot[k];
                    s.sl[l].rhs[i][k] = rhs[k];
                }
                s.sl[l].scale[i][0] = sc[0];
                s.sl[l].scale[i][1] = sc[1];
                const double lin2 = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
                const double ang2 = rhs[3] * rhs[3] + rhs[4] * rhs[4] + rhs[5] * rhs[5];
                s.ratio[i][0] = lin2 / sc[0];
                s.ratio[i][1] = ang2 / sc[1];
            }
            if (store_mass && r < 6)
                for (int c = 0; c < 6; ++c) s.mass0[i][r * 6 + c] = brow[c];
        }
        __syncthreads();
        if (t == 0) {
            double worst = 0.0;
            for (int q = 0; q < md; ++q) {
                worst = max2(worst, s.ratio[q][0]);
                worst = max2(worst, s.ratio[q][1]);
            }
            s.wn[l] = worst;
        }
    }
    // A = 0·A + [i=j]·𝕀 + Σ terms (−half²·∂W/∂pose − half·∂W/∂vel), non-zero blocks only
    __device__ void assemble() {
        const double pf = -(half * half), vf = -half;
        if (tid() == 0) s.ibcast[0] = 0;
        for (int e = tid(); e < nnz * 6; e += NT) {
            const int q = e / 6, r = e - 6 * q;
            const int i = s.tp.nbr_row[q], j = s.tp.nbr[q];
            const double dg = (i == j) ? 1.0 : 0.0;
            double* dst = A + (size_t)(6 * i + r) * LD + 6 * j;
            double acc[6];
            for (int c = 0; c < 6; ++c) acc[c] = 0.0 * dst[c] + dg * s.mass0[i][r * 6 + c];
            for (int k = s.tp.term_start[q]; k < s.tp.term_start[q + 1]; ++k) {
                const int code = s.tp.terms[k];
                const int jn = code >> 2, re = (code >> 1) & 1, ce = code & 1;
                const double* jr = jac + (size_t)jn * 288 + re * 144;
                for (int c = 0; c < 6; ++c) {
                    const double p = jr[(ce * 6 + c) * 6 + r];
                    const double v = jr[(12 + ce * 6 + c) * 6 + r];
                    acc[c] = acc[c] + pf * p + vf * v;
                }
            }
            for (int c = 0; c < 6; ++c) dst[c] = acc[c];
        }
    }
    // out = a·b on 16×16 FP64 MFMA tiles (lane l supplies a[l%16][l/16] and
    // b[l/16][l%16]; accumulator j holds c[l/16 + 4j][l%16]), wave w taking tiles
    // w, w + 8, …; with `sparse` the 4-column slabs of a that are zero in the slab
    // masks are skipped. With make_t the stores are T = 2I − a·b and the return is
    // this thread's share of ‖I − a·b‖² over the real n × n entries. Padding is
    // stored as 0.
    __device__ double product(const double* am, const double* bm, double* out, bool sparse, bool make_t) {
        const int wave = uni((int)(tid() >> 6)), lane = tid() & 63;
        const int lr = lane & 15, lk = lane >> 4;
        const int TT = n16 >> 4, KS = n16 >> 2;
        double part = 0.0;
        for (int tile = wave; tile < TT * TT; tile += WAVES) {
            const int ti = tile / TT, tj = tile - ti * TT;
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            const double* ar = am + (size_t)(ti * 16 + lr) * LD + lk;
            const double* bc = bm + (size_t)lk * LD + tj * 16 + lr;
            if (sparse) {
                for (int ks = 0; ks < KS; ++ks) {
                    const unsigned mk = uni((int)s.tp.kmask[ti][ks >> 5]);
                    if (!((mk >> (ks & 31)) & 1u)) continue;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[4 * ks], bc[(size_t)4 * ks * LD], acc, 0, 0, 0);
                }
            } else {
                int ks = 0;
                for (; ks + 4 <= KS; ks += 4) {
                    double av[4], bv[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        av[u] = ar[4 * (ks + u)];
                        bv[u] = bc[(size_t)4 * (ks + u) * LD];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], acc, 0, 0, 0);
                }
                for (; ks < KS; ++ks)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[4 * ks], bc[(size_t)4 * ks * LD], acc, 0, 0, 0);
            }
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int r = ti * 16 + lk + 4 * jj, c = tj * 16 + lr;
                const bool real = r < n && c < n;
                double v = acc[jj];
                if (make_t) {
                    const double d = (r == c) ? 1.0 : 0.0;
                    const double x = d - v;
                    part += real ? x * x : 0.0;
                    v = 2.0 * d - v;
                }
                out[(size_t)r * LD + c] = real ? v : 0.0;
            }
        }
        return part;
    }
    __device__ double reduce_sq(const double* mat, bool identity_diag) {
        double part = 0.0;
        for (int e = tid(); e < n * n; e += NT) {
            const int row = e / n, col = e % n;
            const double diag = (identity_diag && row == col) ? 1.0 : 0.0;
            const double x = diag - mat[(size_t)row * LD + col];
            part = part + x * x;
        }
        return block_sum(s, part);
    }
    __device__ double refine_inverse() {  // newton.rs:255-279
        double part = product(A, X, T, true, true);
        __syncthreads();
        product(X, T, Xn, false, false);
        const double before = block_sum(s, part);   // its barriers also publish X'
        double* tmp = X;
        X = Xn;
        Xn = tmp;
        st_ns++;
        return before;
    }
    __device__ void seed_transpose() {  // newton.rs:209-244
        if (tid() < 2 * md) {
            const int i = tid() % md;
            const bool col = tid() >= md;
            double sum = 0.0;
            for (int j = 0; j < md; ++j)
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c)
                        sum += fabs(col ? A[(size_t)(6 * j + r) * LD + 6 * i + c] : A[(size_t)(6 * i + r) * LD + 6 * j + c]);
            s.seedsum[tid()] = sum;
        }
        __syncthreads();
        if (tid() == 0) {
            double one = 0.0, inf = 0.0;
            for (int i = 0; i < md; ++i) {
                inf = max2(inf, s.seedsum[i]);
                one = max2(one, s.seedsum[md + i]);
            }
            s.bcast[0] = 1.0 / (one * inf);
        }
        __syncthreads();
        const double sc = uni(s.bcast[0]);
        for (int e = tid(); e < n * n; e += NT) {
            const int row = e / n, col = e % n;
            X[(size_t)row * LD + col] = A[(size_t)col * LD + row] * sc;
        }
        __syncthreads();
        seeded = true;
    }
    __device__ void run_budget(int budget, int seed_budget) {  // newton.rs:301-327
        bool has_prev = false, reseeded = false;
        double previous = 0.0;
        int remaining = budget;
        while (remaining > 0) {
            remaining -= 1;
            const double before = refine_inverse();
            if (has_prev && !(before < 4.0 * previous) && !reseeded) {
                seed_transpose();
                reseeded = true;
                has_prev = false;
                remaining = seed_budget;
                st_reseed++;
                continue;
            }
            previous = before;
            has_prev = true;
        }
    }
    // dv = 0·dv + Σ_j X[i,j]·rhs_j in block order, one thread per row; the writer of
    // each entry also copies the iterate into the line search's base
    __device__ void matvec() {
        const int t = tid();
        if (t < n) {
            const int i = t / 6, r = t % 6;
            double acc = 0.0 * s.dv[i][r];
            for (int j = 0; j < md; ++j) {
                const double* xb = X + (size_t)t * LD + 6 * j;
                double sum = 0.0;
                for (int c = 0; c < 6; ++c) sum = sum + xb[c] * s.sl[0].rhs[j][c];
                acc = acc + sum;
            }
            s.dv[i][r] = acc;
            s.base[i][r] = s.sl[0].vmid[i][r];
            if (acc != 0.0) s.ibcast[0] = 1;
        }
    }

    // ───── one Newton solve at the current `half` (newton.rs:334-530) ─────
    __device__ void newton_solve(double tol2, int ns_warm, int ns_seed, bool& converged, double& best,
                                 bool& assembled) {
        const double huge = 1000000000.0;
        for (int t = tid(); t < m; t += NT)
            world_velocity(s.pose[t], s.massv[t], s.tens[t], s.mom[t], s.sl[0].vmid[t]);
        __syncthreads();
        // kinematic bodies: the line search's base holds their prescribed velocity and
        // their dv is −0, the additive identity, so that every candidate base + α·dv is
        // that velocity bit for bit (the matvec writes the unknowns' rows only)
        for (int t = tid(); t < m * 6; t += NT) {
            s.best[t / 6][t % 6] = s.sl[0].vmid[t / 6][t % 6];
            if (t >= md * 6) {
                s.base[t / 6][t % 6] = s.sl[0].vmid[t / 6][t % 6];
                s.dv[t / 6][t % 6] = -0.0;
            }
        }
        __syncthreads();
        best = huge;
        converged = false;
        assembled = false;
        bool has_carried = false;
        double carried = 0.0;
        for (int iter = 0; iter < MAX_ITERATIONS; ++iter) {
            double rnorm;
            st_newton++;
            if (has_carried) {
                has_carried = false;
                joints_jacobian();
                __syncthreads();
                rnorm = carried;
            } else {
                pre(0);
                __syncthreads();
                joints_jacobian();
                __syncthreads();
                gather_post(0, true);
                __syncthreads();
                rnorm = uni(s.wn[0]);
            }
            if (rnorm < best) {
                best = rnorm;
                for (int t = tid(); t < m * 6; t += NT) s.best[t / 6][t % 6] = s.sl[0].vmid[t / 6][t % 6];
            }
            if (rnorm <= tol2) {
                converged = true;
                break;
            }
            if (!(rnorm < huge)) break;
            assemble();
            __syncthreads();
            assembled = true;
            int budget;
            if (seeded) budget = ns_warm;
            else {
                seed_transpose();
                budget = ns_seed;
            }
            run_budget(budget, ns_seed);
            matvec();
            __syncthreads();
            if (uni(s.ibcast[0]) == 0) break;   // all-zero correction → hand over to subdivision
            // backtracking line search from the same base, α = 2^-lane, lanes in order:
            // lane 0 is evaluated in the iterate's own slot (its mass blocks are the
            // oracle's lanes[0].mass), later lanes in slot 1
            int accepted = -1;
            double tnorm = 0.0;
            for (int lane = 0; lane < LANES && accepted < 0; ++lane) {
                const int sl = lane == 0 ? 0 : 1;
                const double alpha = ldexp(1.0, -lane);
                for (int t = tid(); t < m * 6; t += NT)
                    s.sl[sl].vmid[t / 6][t % 6] = s.base[t / 6][t % 6] + s.dv[t / 6][t % 6] * alpha;
                __syncthreads();
                pre(sl);
                __syncthreads();
                joints_plain(sl);
                __syncthreads();
                gather_post(sl, lane == 0);
                __syncthreads();
                const double w = uni(s.wn[sl]);
                if (w < rnorm) {
                    accepted = lane;
                    tnorm = w;
                }
            }
            if (accepted >= 0) {
                if (accepted > 0) {  // adopt_lane (cache.rs:1115-1136)
                    for (int t = tid(); t < m * 6; t += NT) {
                        s.sl[0].vmid[t / 6][t % 6] = s.sl[1].vmid[t / 6][t % 6];
                        s.sl[0].rhs[t / 6][t % 6] = s.sl[1].rhs[t / 6][t % 6];
                    }
                    for (int t = tid(); t < m * 8; t += NT) s.sl[0].mid[t / 8][t % 8] = s.sl[1].mid[t / 8][t % 8];
                    // the accepted lane's mass block, recomputed bit-identically from its inputs
                    for (int t = tid(); t < md; t += NT) {
                        double block[6][6], rr[6], sc[2];
                        const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
                        body_post(s.sl[1].mid[t], s.sl[1].vmid[t], s.massv[t], s.tens[t], s.snap[t], s.sl[1].total[t],
                                  half, floor2, block, rr, sc);
                        for (int r = 0; r < 6; ++r)
                            for (int c = 0; c < 6; ++c) s.mass0[t][r * 6 + c] = block[r][c];
                    }
                    __syncthreads();
                }
                has_carried = true;
                carried = tnorm;
            } else {
                for (int t = tid(); t < m * 6; t += NT)
                    s.sl[0].vmid[t / 6][t % 6] = s.base[t / 6][t % 6] + s.dv[t / 6][t % 6] * 0.0;
                __syncthreads();
                st_reject++;
                break;
            }
        }
        __syncthreads();
        for (int t = tid(); t < m * 6; t += NT) s.sl[0].vmid[t / 6][t % 6] = s.best[t / 6][t % 6];
        __syncthreads();
    }

    static __device__ bool log2_floor(double v, int& e) {  // newton.rs:101-121
        double x = v;
        e = 0;
        if (!(x > 0.0)) return false;
        while (x >= 2.0) {
            x = x / 2.0;
            e += 1;
            if (e > 250) return false;
        }
        while (x < 1.0 && e > -250) {
            x = x * 2.0;
            e -= 1;
        }
        return true;
    }
    __device__ bool signal_log2(bool assembled, int& r) {
        if (!assembled) return false;
        const double an = reduce_sq(A, false);
        double part = 0.0;
        for (int e = tid(); e < md * 36; e += NT) {
            const double v = s.mass0[e / 36][e % 36];
            part = part + v * v;
        }
        const double mn = block_sum(s, part);
        const double stiff = effective_zero(mn) ? 0.0 : an / mn;
        return log2_floor(stiff, r);
    }
    __device__ void restore_x() {
        const double* src = a.xacc + xoff;
        for (int e = tid(); e < n * n; e += NT) X[(size_t)(e / n) * LD + e % n] = src[e];
        __syncthreads();
    }
    __device__ void publish_x(bool& accepted_seeded) {
        double* dst = a.xacc + xoff;
        double part = 0.0;
        for (int e = tid(); e < n * n; e += NT) {
            const double v = X[(size_t)(e / n) * LD + e % n];
            dst[e] = v;
            const double x = 0.0 - v;
            part = part + x * x;
        }
        const double nx = block_sum(s, part);
        const double huge = 1000000000.0;
        if (nx < huge * huge) accepted_seeded = seeded;
        else {
            seeded = false;
            accepted_seeded = false;
        }
    }
    static __device__ int floor_div(int q, int d) { return (q >= 0) ? q / d : -((-q + d - 1) / d); }

    // ───── solve_step (newton.rs:543-683) ─────
    __device__ void solve_step(double tol2, int ns_warm, int ns_seed) {
        bool accepted_seeded = seeded;
        uint64_t remaining = 1ull << MAX_DEPTH;
        int depth = 0;
        while (remaining > 0) {
            uint64_t units = 1ull << (MAX_DEPTH - depth);
            while (units > remaining) {
                depth += 1;
                units >>= 1;
            }
            span = dt;
            for (int k = 0; k < depth; ++k) span = span / 2.0;
            half = span / 2.0;
            st_spans++;
            for (int t = tid(); t < m; t += NT) world_momentum(s.pose[t], s.mom[t], s.snap[t]);
            __syncthreads();
            bool converged, assembled;
            double best;
            newton_solve(tol2, ns_warm, ns_seed, converged, best, assembled);
            const bool finite = best < 1000000000.0;
            const bool last = depth == MAX_DEPTH;
            if (!converged && !last) {
                restore_x();
                seeded = accepted_seeded;
                int r, down;
                if (signal_log2(assembled, r)) down = floor_div(r - THRESHOLD_LOG2, PER_HALVING_LOG2) + 1;
                else down = 1;
                const int hi = MAX_DEPTH - depth;
                down = down < 1 ? 1 : (down > hi ? hi : down);
                depth += down;
                continue;
            }
            if (converged || finite) {
                if (!converged) st_uncommit++;
                for (int t = tid(); t < md; t += NT)   // the unknowns only (newton.rs:704-753)
                    commit_body(s.pose[t], s.mom[t], s.massv[t], s.tens[t], s.sl[0].vmid[t], s.snap[t], span, half,
                                converged);
                __syncthreads();
                publish_x(accepted_seeded);
            } else {
                restore_x();
                seeded = accepted_seeded;
                st_nonfinite++;
            }
            remaining -= units;
            auto aligned = [&](int d) {
                if (d <= 0) return false;
                const uint64_t bigger = 1ull << (MAX_DEPTH - (d - 1));
                return bigger <= remaining && remaining % bigger == 0;
            };
            int r;
            if (remaining > 0 && converged && aligned(depth) && signal_log2(assembled, r)) {
                int up = floor_div(THRESHOLD_LOG2 - r, PER_HALVING_LOG2);
                while (up > 0 && aligned(depth)) {
                    depth -= 1;
                    up -= 1;
                }
            }
        }
    }
};

__global__ __launch_bounds__(NT) void island_step_kernel_large(IslandArgs a) {
    __shared__ LargeShared s;
    if ((int)blockIdx.x >= a.n_islands) return;
    const int isl = a.island_ids ? a.island_ids[blockIdx.x] : (int)blockIdx.x;
    const int b0 = a.island_start[isl], m = a.island_end[isl] - b0;
    const int j0 = a.island_jstart[isl], J = a.island_jend[isl] - j0;
    if (m > MAXM || J > MAXJ || m <= 0) return;  // the host routes islands by size; never taken

    // ── load the island ──
    for (int t = threadIdx.x; t < m * 8; t += NT) s.pose[t / 8][t % 8] = a.pose[(size_t)(b0 + t / 8) * 8 + t % 8];
    for (int t = threadIdx.x; t < m * 6; t += NT) {
        s.mom[t / 6][t % 6] = a.mom[(size_t)(b0 + t / 6) * 6 + t % 6];
        s.ext[t / 6][t % 6] = field_add(a, a.mass[b0 + t / 6], t % 6, a.ext[(size_t)(b0 + t / 6) * 6 + t % 6]);
        s.dv[t / 6][t % 6] = 0.0;
    }
    for (int t = threadIdx.x; t < m * 9; t += NT) s.tens[t / 9][t % 9] = a.tens[(size_t)(b0 + t / 9) * 9 + t % 9];
    for (int t = threadIdx.x; t < m; t += NT) s.massv[t] = a.mass[b0 + t];
    for (int t = threadIdx.x; t < J; t += NT) {
        s.jkind[t] = a.jkind[j0 + t];
        s.ja[t] = a.ja[j0 + t] - b0;
        s.jb[t] = a.jb[j0 + t] - b0;
        for (int p = 0; p < 10; ++p) s.jparams[t][p] = a.jparams[(size_t)(j0 + t) * 10 + p];
    }
    {   // the island's topology record, word by word
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.ltopo + isl);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&s.tp);
        static_assert(sizeof(LargeTopo) % 4 == 0, "the record copies as words");
        for (int t = threadIdx.x; t < (int)(sizeof(LargeTopo) / 4); t += NT) dst[t] = src[t];
    }
    LargeSolver sv{s, a};
    sv.m = m;
    sv.J = J;
    // the unknowns: the dynamic bodies, which come first (the host checks the order)
    int md = 0;
    for (int i = 0; i < m; ++i) md += a.mass[b0 + i] != 0.0 ? 1 : 0;
    sv.md = uni(md);
    sv.n = 6 * sv.md;
    sv.n16 = ((sv.n + 15) / 16) * 16;
    sv.LD = sv.n16;
    sv.dt = a.dt;
    sv.warp = a.warp;
    sv.xoff = a.xacc_off[isl];
    {
        double* ws = a.lws + a.lws_off[isl];
        const size_t nn = (size_t)sv.n16 * sv.n16;
        sv.A = ws;
        sv.X = ws + nn;
        sv.Xn = ws + 2 * nn;
        sv.T = ws + 3 * nn;
        sv.jac = ws + 4 * nn;
        // the matrices start at zero each launch (their padding stays zero), then X_acc
        for (size_t e = threadIdx.x; e < 4 * nn; e += NT) ws[e] = 0.0;
        __syncthreads();
        const int n = sv.n;
        for (int e = threadIdx.x; e < n * n; e += NT) sv.X[(size_t)(e / n) * sv.LD + e % n] = a.xacc[sv.xoff + e];
    }
    __syncthreads();
    sv.nnz = s.tp.nbr_start[m];
    sv.st_spans = sv.st_newton = sv.st_ns = sv.st_reseed = sv.st_reject = sv.st_uncommit = sv.st_nonfinite = 0;
    sv.seeded = a.seeded[isl] != 0;
    sv.solve_step(a.tol2, a.ns_warm, a.ns_seed);
    __syncthreads();

    // ── recompute_centroid (island.rs:206-244) and publish the charges ──
    // One thread per body forms its position (into snap, whose span is over); thread 0
    // adds them in body order and forms the shift (into best: the solve is over); one
    // thread per body composes its pose with the shift and checks its state.
    double* shift_sh = &s.best[0][0];
    for (int t = threadIdx.x; t < m; t += NT) body_position(s.pose[t], s.snap[t]);
    if (threadIdx.x == 0) {
        s.ibcast[1] = sv.seeded ? 1 : 0;
        s.ibcast[2] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total_m = 0.0, wx[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < m; ++i) {
            total_m += s.massv[i];
            for (int k = 0; k < 3; ++k) wx[k] += s.snap[i][k] * s.massv[i];
        }
        const bool shifted = !effective_zero(total_m);
        if (shifted) {
            const double inv = 1.0 / total_m;
            double sh[6] = {-(wx[0] * inv), -(wx[1] * inv), -(wx[2] * inv), 0.0, 0.0, 0.0};
            twist_exp(sh, 1.0, shift_sh);
            for (int k = 0; k < 3; ++k) a.origin[isl * 3 + k] += wx[k] * inv;
        }
        s.ibcast[3] = shifted ? 1 : 0;
        for (int k = 0; k < 3; ++k) s.bcast[2 + k] = a.origin[isl * 3 + k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += NT) {
        if (s.ibcast[3]) {
            double np[8];
            compose(s.pose[i], shift_sh, np);
            for (int k = 0; k < 8; ++k) s.pose[i][k] = np[k];
        }
        bool bad = false;
        for (int k = 0; k < 8; ++k) bad |= !isfinite(s.pose[i][k]);
        for (int k = 0; k < 6; ++k) bad |= !isfinite(s.mom[i][k]);
        if (bad) s.ibcast[2] = 1;
    }
    __syncthreads();
    const double org[3] = {s.bcast[2], s.bcast[3], s.bcast[4]};
    for (int t = threadIdx.x; t < m; t += NT) {
        double p[3];
        body_position(s.pose[t], p);
        double* c = a.charges_out + (size_t)(b0 + t) * 8;
        for (int k = 0; k < 3; ++k) {
            c[k] = p[k] + org[k];
            c[3 + k] = p[k];
        }
        c[6] = (a.G_of_island ? a.G_of_island[isl] : a.G) * s.massv[t];
        c[7] = a.size[b0 + t];
    }
    for (int t = threadIdx.x; t < m * 8; t += NT) a.pose[(size_t)(b0 + t / 8) * 8 + t % 8] = s.pose[t / 8][t % 8];
    for (int t = threadIdx.x; t < m * 6; t += NT) a.mom[(size_t)(b0 + t / 6) * 6 + t % 6] = s.mom[t / 6][t % 6];
    if (threadIdx.x == 0) {
        a.seeded[isl] = s.ibcast[1];
        int* st = a.stats + (size_t)isl * ISLAND_STATS;
        st[0] += sv.st_spans;
        st[1] += sv.st_newton;
        st[2] += sv.st_ns;
        st[3] += sv.st_reseed;
        st[4] += sv.st_reject;
        st[5] += sv.st_uncommit;
        st[STAT_NONFINITE_SPANS] += sv.st_nonfinite;
        if (s.ibcast[2]) mark_fatal(a, isl);
    }
}

}  // namespace

void launch_island_large(const IslandArgs& a, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    if (a.n_islands <= 0) return;
    launch_k(island_step_kernel_large, dim3(a.n_islands), dim3(NT), stream, start, stop, a);
}

}  // namespace exg
