// One implicit Newton step of single-body islands without joints: the control flow of
// island.hip (crates/newton/src/integrator/implicit/newton.rs:334-683, 209-327,
// 688-753) restated for a single 6×6 block, in the reference's own summation order
// (the oracle's term order for gemm, reduce, matvec, seed).
//
// Free bodies are most of the batched and served worlds (60 % of their bodies). Two
// kernels share them:
//
// * singles_step_kernel, one lane per island: its body state in registers, its world
//   mass block (which is also A: the block is assembled with no joint terms) and its
//   approximate inverse X in lane-private LDS columns (entry e of lane l at [e][l], so
//   a wave's access to one entry is 64 consecutive doubles; 36.9 KB per wave, four
//   waves per CU), and, inside a Newton–Schulz iteration, X and the product T in
//   registers. Each lane runs the reference's sequential arithmetic on its own body,
//   so a wave holds 64 islands and no lane waits on another for data. A body whose
//   first span does not converge (its step subdivides), or needs more than BAIL_WORK
//   residual evaluations, is given up before anything of it is written and its island
//   goes on a requeue list: a wave lasts as long as its slowest island, and one such
//   body would hold 63 others.
// * singles_requeue_kernel, eight lanes per island over the requeue list: lane r (r <
//   6; lanes 6 and 7 mirror row 5) holds row r of A, X, T and the candidate-0 mass
//   block in registers and forms that row of every product, the other rows coming from
//   the group by lane shuffles; the body's own arithmetic (midpoint, residual, commit)
//   runs in every lane on the same values. Eight lanes finish one subdividing body
//   sooner than one lane does. Its last workgroup clears the list's counters for the
//   next step (no memset: the step stays capturable in a graph).
//
// Both solve the whole step from its start with the same arithmetic in the same order;
// the compiler fuses multiply-adds differently in the two kernels, so they agree to
// rounding, not bit for bit. Which kernel finishes a body is decided by that body's own
// state alone, so a world's results do not depend on what else shares the launch.
// Without a requeue list (IslandArgs::requeue null) the one-lane kernel solves every
// step itself.
//
// Line-search lanes are evaluated in order with early exit: the first lane whose
// residual drops below the current one is accepted, which is the lane the reference's
// side-by-side evaluation accepts; lane 0's mass block is kept for signal_log2 when
// every lane fails, as in the reference. In the one-lane kernel a later lane is
// evaluated without keeping its block, and the block is formed again (the same
// arithmetic, the same bits) only when that lane is accepted.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "island.h"
#include "physics_dev.h"
#include "body_rows.h"

namespace exg {

namespace {

constexpr int SB = 64;           // one-lane kernel: threads per workgroup, one wave of 64 islands
constexpr int SG = 8;            // requeue kernel: lanes per island
constexpr int SB8 = 256;         // requeue kernel: threads per workgroup (32 islands)
constexpr int MAX_IT = 32;
// One-lane kernel: residual evaluations (Newton iterations and line-search lanes) a
// first span may take before its island is handed to the requeue kernel. A free body
// converges in about four Newton iterations (eight evaluations); bodies in close
// encounters take many more, and one such lane would hold its 63 wave-mates.
constexpr int BAIL_WORK = 16;
constexpr int DEPTH = 8;
constexpr int THRESH = 8;
constexpr int PER_HALVING = 4;
constexpr int NLANES = 8;

__device__ inline double mx2(double x, double y) { return x > y ? x : y; }

// Lane-private LDS columns of the workgroup's 64 islands: entry e of lane l's world
// mass block at s_mass[e][l], of its approximate inverse at s_inv[e][l]. Declared at
// file scope so that every access is a direct LDS access (an immediate offset from one
// lane address), not a flat access through a pointer.
__shared__ double s_mass[36][SB];
__shared__ double s_inv[36][SB];

// Σ over the 36 entries, row-major from 0, of (diag·δ − b)² (fn_reduce_sq_term).
__device__ __forceinline__ double reduce_sq(const double b[36], double diag) {
    double out = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double v = b[r * 6 + c];
            const double e = (r == c) ? diag - v : -v;
            out += e * e;
        }
    return 0.0 + out;
}
template <bool INV>
__device__ __forceinline__ double reduce_sq_lds(int ln, double diag) {
    double out = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double v = INV ? s_inv[r * 6 + c][ln] : s_mass[r * 6 + c][ln];
            const double e = (r == c) ? diag - v : -v;
            out += e * e;
        }
    return 0.0 + out;
}

// The step's counters of one island (ISLAND_STATS slots 0-5 and 7).
struct StepCounts {
    int spans = 0, newton = 0, ns = 0, reseed = 0, reject = 0, uncommit = 0, nonfinite = 0;
};

// recompute_centroid (island.rs:206-244) for one body, the published charge, the state
// write-back, the counters and the fatal mark. Every caller lane computes; `writer`
// stores.
__device__ __forceinline__ void finish_island(const IslandArgs& a, int isl, int b, double pose[8], const double mom[6],
                                              double mass, bool seeded, const StepCounts& n, bool writer) {
    double p[3];
    body_position(pose, p);
    const double total_m = mass;
    double wx[3] = {0.0 + p[0] * mass, 0.0 + p[1] * mass, 0.0 + p[2] * mass};
    double org[3] = {a.origin[isl * 3 + 0], a.origin[isl * 3 + 1], a.origin[isl * 3 + 2]};
    if (!effective_zero(total_m)) {
        double inv = 1.0 / total_m;
        double sh[6] = {-(wx[0] * inv), -(wx[1] * inv), -(wx[2] * inv), 0.0, 0.0, 0.0};
        double shift[8], np[8];
        twist_exp(sh, 1.0, shift);
        compose(pose, shift, np);
#pragma unroll
        for (int q = 0; q < 8; ++q) pose[q] = np[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) org[q] += wx[q] * inv;
    }
    if (!writer) return;
#pragma unroll
    for (int q = 0; q < 3; ++q) a.origin[isl * 3 + q] = org[q];
    body_position(pose, p);
    double* c = a.charges_out + (size_t)b * 8;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c[q] = p[q] + org[q];
        c[3 + q] = p[q];
    }
    c[6] = (a.G_of_island ? a.G_of_island[isl] : a.G) * mass;
    c[7] = a.size[b];
#pragma unroll
    for (int q = 0; q < 8; ++q) a.pose[(size_t)b * 8 + q] = pose[q];
#pragma unroll
    for (int q = 0; q < 6; ++q) a.mom[(size_t)b * 6 + q] = mom[q];
    a.seeded[isl] = seeded ? 1 : 0;
    int* st = a.stats + (size_t)isl * ISLAND_STATS;
    st[0] += n.spans;
    st[1] += n.newton;
    st[2] += n.ns;
    st[3] += n.reseed;
    st[4] += n.reject;
    st[5] += n.uncommit;
    st[STAT_NONFINITE_SPANS] += n.nonfinite;
    bool bad = false;
#pragma unroll
    for (int q = 0; q < 8; ++q) bad |= !isfinite(pose[q]);
#pragma unroll
    for (int q = 0; q < 6; ++q) bad |= !isfinite(mom[q]);
    if (bad) mark_fatal(a, isl);
}

// body_post (functions/src/lib.rs:170-247) for one body without keeping the mass
// block: each column of Ad*_M ∘ I ∘ Ad_{M⁻¹} is formed as body_post forms it and
// enters P = block·V at once (P[r] sums its columns in order from 0, as there).
__device__ __forceinline__ void body_post_momentum(const double pose_mid[8], const double vel[6], double mass,
                                                   const double t9[9], double p[6]) {
    double inv[8], adi[6][6], cad[6][6];
    motor_rev(pose_mid, inv);
    ad_matrix(inv, adi);
    coad_matrix(pose_mid, cad);
#pragma unroll
    for (int r = 0; r < 6; ++r) p[r] = 0.0;
#pragma unroll
    for (int col = 0; col < 6; ++col) {
        double x[6];
#pragma unroll
        for (int i = 0; i < 3; ++i) x[i] = adi[i][col] * mass;
#pragma unroll
        for (int r = 0; r < 3; ++r)
            x[3 + r] = t9[3 * r] * adi[3][col] + t9[3 * r + 1] * adi[4][col] + t9[3 * r + 2] * adi[5][col];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double acc = cad[r][0] * x[0];
#pragma unroll
            for (int k = 1; k < 6; ++k) acc = acc + cad[r][k] * x[k];
            p[r] = p[r] + acc * vel[col];
        }
    }
}

// newton.rs:209-244 for one block, X = Aᵀ/(‖A‖₁‖A‖∞), on lane ln's LDS columns.
// Out of line (it touches LDS and scalars only): one copy for its two call sites.
__device__ __attribute__((noinline)) void seed_inverse(int ln) {
    double row_s = 0.0, col_s = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double v = fabs(s_mass[r * 6 + c][ln]);
            row_s += v;
            col_s += v;
        }
    const double inf = mx2(0.0, row_s), one = mx2(0.0, col_s);
    const double scale = 1.0 / (one * inf);
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) s_inv[r * 6 + c][ln] = s_mass[c * 6 + r][ln] * scale;
}
// T = A·X, ‖I − T‖², X ← 0·X + X·(2I − T) (newton.rs:255-279) on lane ln's LDS
// columns, the products term for term as the oracle's gemm: out = 0·out +
// a·(sign·b + diag·2I). T starts from zero at each product (0·T + s = s for the
// finite T this kernel ever forms). Returns ‖I − T‖². Out of line, as seed_inverse.
__device__ __attribute__((noinline)) double refine_inverse_lds(int ln) {
    const double two = 1.0 + 1.0;
    double Xr[36], T[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) Xr[e] = s_inv[e][ln];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        double ar[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) ar[k] = s_mass[r * 6 + k][ln];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                double right = 1.0 * Xr[k * 6 + c];
                if (k == c) right += 0.0 * two;
                s += ar[k] * right;
            }
            T[r * 6 + c] = s;
        }
    }
    const double before = reduce_sq(T, 1.0);
    // X's rows are replaced one at a time (row r of the product reads row r of the old
    // X, held in registers, and all of T)
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        double xn[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                double right = -1.0 * T[k * 6 + c];
                if (k == c) right += 1.0 * two;
                s += Xr[r * 6 + k] * right;
            }
            xn[c] = 0.0 * Xr[r * 6 + c] + s;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) s_inv[r * 6 + c][ln] = xn[c];
    }
    return before;
}

struct Free {
    const IslandArgs& a;
    int isl;
    double pose[8], mom[6], mass, tens[9], ext[6];
    double snap[6], vmid[6], rhs[6], best_v[6], dv[6];
    // M: the world mass block of the current iterate; assembly makes it A (there are no
    // joint terms), and it stays A through the Newton–Schulz iterations and the matvec.
    // The line search then writes candidate 0's block over it (A is not read again;
    // signal_log2 takes ‖A‖² from `an`, kept at assembly).
    // (M and X live in s_mass and s_inv, column `ln`)
    int ln;
    double an;
    double dt, warp, half, span;
    bool seeded;
    int bail_work = 0;      // > 0: give the span up after this many residual evaluations
    bool bailed = false;
    int st_spans = 0, st_newton = 0, st_ns = 0, st_reseed = 0, st_reject = 0, st_uncommit = 0, st_nonfinite = 0;

    // The residual of a candidate velocity (pre + gather + body_post + wnorm for one
    // body without joints; functions/src/lib.rs:170-247, newton.rs:139-151) from its
    // momentum P, and the worst residual ratio.
    __device__ __forceinline__ double residual(const double p[6], double rhs_out[6]) {
        const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
        double hh[6], sc[2];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            hh[r] = ext[r] * half;
            rhs_out[r] = -(p[r] - snap[r] - hh[r]);
        }
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const int b = 3 * o;
            const double sp_ = p[b] * p[b] + p[b + 1] * p[b + 1] + p[b + 2] * p[b + 2];
            const double sn = snap[b] * snap[b] + snap[b + 1] * snap[b + 1] + snap[b + 2] * snap[b + 2];
            const double sh = hh[b] * hh[b] + hh[b + 1] * hh[b + 1] + hh[b + 2] * hh[b + 2];
            sc[o] = fmax(fmax(fmax(sp_, sn), sh), floor2);
        }
        const double lin2 = rhs_out[0] * rhs_out[0] + rhs_out[1] * rhs_out[1] + rhs_out[2] * rhs_out[2];
        const double ang2 = rhs_out[3] * rhs_out[3] + rhs_out[4] * rhs_out[4] + rhs_out[5] * rhs_out[5];
        double worst = 0.0;
        worst = mx2(worst, lin2 / sc[0]);
        worst = mx2(worst, ang2 / sc[1]);
        return worst;
    }
    // candidate v: midpoint, world mass block into M, residual; returns the ratio
    __device__ __forceinline__ double evaluate_block(const double v[6], double rhs_out[6]) {
        double e[8], mid[8], p[6];
        twist_exp(v, half, e);
        compose(pose, e, mid);
        double inv[8], adi[6][6], cad[6][6];
        motor_rev(mid, inv);
        ad_matrix(inv, adi);
        coad_matrix(mid, cad);
#pragma unroll
        for (int col = 0; col < 6; ++col) {
            double x[6];
#pragma unroll
            for (int i = 0; i < 3; ++i) x[i] = adi[i][col] * mass;
#pragma unroll
            for (int r = 0; r < 3; ++r)
                x[3 + r] = tens[3 * r] * adi[3][col] + tens[3 * r + 1] * adi[4][col] + tens[3 * r + 2] * adi[5][col];
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                double acc = cad[r][0] * x[0];
#pragma unroll
                for (int k = 1; k < 6; ++k) acc = acc + cad[r][k] * x[k];
                s_mass[r * 6 + col][ln] = acc;
            }
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc = acc + s_mass[r * 6 + c][ln] * v[c];
            p[r] = acc;
        }
        return residual(p, rhs_out);
    }
    // candidate v without keeping its block
    __device__ __forceinline__ double evaluate_ratio(const double v[6], double rhs_out[6]) {
        double e[8], mid[8], p[6];
        twist_exp(v, half, e);
        compose(pose, e, mid);
        body_post_momentum(mid, v, mass, tens, p);
        return residual(p, rhs_out);
    }
    __device__ __forceinline__ void seed_transpose() {
        seed_inverse(ln);
        seeded = true;
    }
    __device__ __forceinline__ double refine_inverse() {
        st_ns++;
        return refine_inverse_lds(ln);
    }
    __device__ __forceinline__ void run_budget(int budget, int seed_budget) {  // newton.rs:301-327
        bool has_prev = false, reseeded = false;
        double previous = 0.0;
        int remaining = budget;
#pragma clang loop unroll(disable)
        while (remaining > 0) {
            remaining -= 1;
            double before = refine_inverse();
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
    __device__ __forceinline__ void newton_solve(double tol2, int ns_warm, int ns_seed, bool& converged, double& best,
                                                 bool& assembled) {  // newton.rs:334-530
        const double huge = 1000000000.0;
        world_velocity(pose, mass, tens, mom, vmid);
#pragma unroll
        for (int k = 0; k < 6; ++k) best_v[k] = vmid[k];
        best = huge;
        converged = false;
        assembled = false;
        bool has_carried = false;
        double carried = 0.0;
        int work = 0;
#pragma clang loop unroll(disable)
        for (int iter = 0; iter < MAX_IT; ++iter) {
            if (bail_work > 0 && work >= bail_work) {
                bailed = true;
                break;
            }
            double rnorm;
            st_newton++;
            if (has_carried) {
                has_carried = false;
                rnorm = carried;
            } else {
                rnorm = evaluate_block(vmid, rhs);
                work++;
            }
            if (rnorm < best) {
                best = rnorm;
#pragma unroll
                for (int k = 0; k < 6; ++k) best_v[k] = vmid[k];
            }
            if (rnorm <= tol2) {
                converged = true;
                break;
            }
            if (!(rnorm < huge)) break;
            // assemble: A = 0·A + 1·mass block with no joint terms, i.e. M itself (A is
            // finite whenever it is assembled: a non-finite block stops the solve first)
            an = 0.0 + reduce_sq_lds<false>(ln, 0.0);
            assembled = true;
            int budget;
            if (seeded) budget = ns_warm;
            else {
                seed_transpose();
                budget = ns_seed;
            }
            run_budget(budget, ns_seed);
            // dv = 0·dv + X·rhs
            bool nz = false;
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                double sum = 0.0;
#pragma unroll
                for (int c = 0; c < 6; ++c) sum += s_inv[r * 6 + c][ln] * rhs[c];
                dv[r] = 0.0 * dv[r] + sum;
                nz |= dv[r] != 0.0;
            }
            if (!nz) break;
            // line search from base = vmid (vmid changes only when a lane is accepted)
            int accepted = -1;
            double tnorm = 0.0;
#pragma clang loop unroll(disable)
            for (int l = 0; l < NLANES; ++l) {
                double v[6], r_l[6];
                const double alpha = ldexp(1.0, -l);
#pragma unroll
                for (int k = 0; k < 6; ++k) v[k] = vmid[k] + dv[k] * alpha;
                // lane 0's block is the iterate's from here on (kept for signal_log2 when
                // every lane fails); a later lane's block is formed again on adoption
                const double w = l == 0 ? evaluate_block(v, r_l) : evaluate_ratio(v, r_l);
                work++;
                if (w < rnorm) {
                    accepted = l;
                    tnorm = w;
                    if (l > 0) {
                        double r_again[6];
                        evaluate_block(v, r_again);  // adopt_lane (cache.rs:1115-1136)
                    }
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        vmid[k] = v[k];
                        rhs[k] = r_l[k];
                    }
                    break;
                }
            }
            if (accepted >= 0) {
                has_carried = true;
                carried = tnorm;
            } else {
#pragma unroll
                for (int k = 0; k < 6; ++k) vmid[k] = vmid[k] + dv[k] * 0.0;
                st_reject++;
                break;
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) vmid[k] = best_v[k];
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
    __device__ __forceinline__ bool signal_log2(bool assembled, int& r) {  // newton.rs:688-702
        if (!assembled) return false;
        const double mn = 0.0 + reduce_sq_lds<false>(ln, 0.0);
        const double stiff = effective_zero(mn) ? 0.0 : an / mn;
        return log2_floor(stiff, r);
    }
    static __device__ int floor_div(int q, int d) { return (q >= 0) ? q / d : -((-q + d - 1) / d); }
    __device__ __forceinline__ void load_x(long long xoff) {
#pragma unroll
        for (int e = 0; e < 36; ++e) s_inv[e][ln] = a.xacc[xoff + e];
    }
    // With `bail`, a first span that does not converge, or that takes more than
    // BAIL_WORK residual evaluations, ends the solve before anything is committed or
    // written (returns false: the island goes to the requeue kernel, which solves the
    // step from its start with the same arithmetic).
    __device__ __forceinline__ bool solve_step(double tol2, int ns_warm, int ns_seed, bool bail) {  // newton.rs:543-683
        const long long xoff = a.xacc_off[isl];
        bool accepted_seeded = seeded;
        uint64_t remaining = 1ull << DEPTH;
        int depth = 0;
        while (remaining > 0) {
            uint64_t units = 1ull << (DEPTH - depth);
            while (units > remaining) {
                depth += 1;
                units >>= 1;
            }
            span = dt;
            for (int k = 0; k < depth; ++k) span = span / 2.0;
            half = span / 2.0;
            st_spans++;
            world_momentum(pose, mom, snap);
            bool converged, assembled;
            double best;
            bail_work = bail && depth == 0 ? BAIL_WORK : 0;
            newton_solve(tol2, ns_warm, ns_seed, converged, best, assembled);
            if (bailed) return false;
            const bool finite = best < 1000000000.0;
            const bool last = depth == DEPTH;
            if (!converged && !last) {
                if (bail && depth == 0) return false;
                load_x(xoff);  // restore X_acc
                seeded = accepted_seeded;
                int r, down;
                if (signal_log2(assembled, r)) down = floor_div(r - THRESH, PER_HALVING) + 1;
                else down = 1;
                const int hi = DEPTH - depth;
                down = down < 1 ? 1 : (down > hi ? hi : down);
                depth += down;
                continue;
            }
            if (converged || finite) {
                if (!converged) st_uncommit++;
                commit_body(pose, mom, mass, tens, vmid, snap, span, half, converged);
#pragma unroll
                for (int e = 0; e < 36; ++e) a.xacc[xoff + e] = s_inv[e][ln];  // publish X_acc
                const double nx = 0.0 + reduce_sq_lds<true>(ln, 0.0);
                const double huge = 1000000000.0;
                if (nx < huge * huge) accepted_seeded = seeded;
                else {
                    seeded = false;
                    accepted_seeded = false;
                }
            } else {
                load_x(xoff);
                seeded = accepted_seeded;
                st_nonfinite++;
            }
            remaining -= units;
            auto aligned = [&](int d) {
                if (d <= 0) return false;
                uint64_t bigger = 1ull << (DEPTH - (d - 1));
                return bigger <= remaining && remaining % bigger == 0;
            };
            int r;
            if (remaining > 0 && converged && aligned(depth) && signal_log2(assembled, r)) {
                int up = floor_div(THRESH - r, PER_HALVING);
                while (up > 0 && aligned(depth)) {
                    depth -= 1;
                    up -= 1;
                }
            }
        }
        return true;
    }
};

__global__ __launch_bounds__(SB) void singles_step_kernel(IslandArgs a) {
    const int k = (int)(blockIdx.x * SB + threadIdx.x);
    if (k >= a.n_islands) return;
    const int isl = a.island_ids ? a.island_ids[k] : k;
    const int b = a.island_start[isl];
    if (a.island_end[isl] - b != 1 || a.island_jend[isl] != a.island_jstart[isl]) return;  // routed by class; never taken
    Free sg{a, isl};
    sg.ln = (int)threadIdx.x;
#pragma unroll
    for (int q = 0; q < 8; ++q) sg.pose[q] = a.pose[(size_t)b * 8 + q];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        sg.mom[q] = a.mom[(size_t)b * 6 + q];
        sg.ext[q] = field_add(a, a.mass[b], q, a.ext[(size_t)b * 6 + q]);
        sg.dv[q] = 0.0;
    }
#pragma unroll
    for (int e = 0; e < 36; ++e) s_mass[e][sg.ln] = 0.0;
#pragma unroll
    for (int q = 0; q < 9; ++q) sg.tens[q] = a.tens[(size_t)b * 9 + q];
    sg.mass = a.mass[b];
    if (sg.mass == 0.0) {
        // a kinematic body alone: no unknowns, so the step is one span whose first
        // Newton iteration finds the empty residual converged, and nothing is committed
        // (newton.rs:806-830, 704-753); its pose and velocity stay as they are
        StepCounts n;
        n.spans = 1;
        n.newton = 1;
        n.ns = n.reseed = n.reject = n.uncommit = n.nonfinite = 0;
        finish_island(a, isl, b, sg.pose, sg.mom, sg.mass, a.seeded[isl] != 0, n, true);
        return;
    }
    sg.an = 0.0;
    sg.load_x(a.xacc_off[isl]);
    sg.dt = a.dt;
    sg.warp = a.warp;
    sg.seeded = a.seeded[isl] != 0;
    if (!sg.solve_step(a.tol2, a.ns_warm, a.ns_seed, a.requeue != nullptr)) {
        a.requeue[atomicAdd(&a.requeue_ctl[0], 1u)] = isl;
        return;
    }
    StepCounts n;
    n.spans = sg.st_spans;
    n.newton = sg.st_newton;
    n.ns = sg.st_ns;
    n.reseed = sg.st_reseed;
    n.reject = sg.st_reject;
    n.uncommit = sg.st_uncommit;
    n.nonfinite = sg.st_nonfinite;
    finish_island(a, isl, b, sg.pose, sg.mom, sg.mass, sg.seeded, n, true);
}

__device__ __forceinline__ double from(double v, int j) { return group8(v, j); }

// Σ over the six rows held by the group (row r in lane r) of (diag·δ − b)², from 0,
// row-major (fn_reduce_sq_term): every lane gathers the 36 terms and folds them in
// order, so every lane holds the same sum.
__device__ __forceinline__ double reduce_rows(const double br[6], double diag) {
    double out = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double v = from(br[c], r);
            const double e = (r == c) ? diag - v : -v;
            out += e * e;
        }
    return 0.0 + out;
}

// ───── the requeue kernel's solver: eight lanes per island ─────
struct Single {
    const IslandArgs& a;
    int row, isl;
    double pose[8], mom[6], mass, tens[9], ext[6];
    double snap[6], vmid[6], mid[8], rhs[6], best_v[6], dv[6];
    double Ar[6], Xr[6], Tr[6], Mr[6];  // row `row` of A, X, T and the candidate-0 mass block
    double dt, warp, half, span;
    bool seeded;
    int st_spans = 0, st_newton = 0, st_ns = 0, st_reseed = 0, st_reject = 0, st_uncommit = 0, st_nonfinite = 0;

    // Midpoint, residual and residual ratio of a candidate velocity (pre + gather +
    // body_post + wnorm for one body without joints; functions/src/lib.rs:170-247,
    // newton.rs:139-151), and row `row` of its world mass block Ad*_M ∘ I ∘ Ad_{M⁻¹}.
    // Each block entry is formed as body_post forms it; the momentum P = block·V
    // gathers its rows from the group.
    __device__ __forceinline__ double evaluate(const double v[6], double mid_out[8], double rhs_out[6], double brow[6]) {
        double e[8], sc[2];
        twist_exp(v, half, e);
        compose(pose, e, mid_out);
        const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
        body_post_row(mid_out, v, mass, tens, snap, ext, half, floor2, row, brow, rhs_out, sc);
        const double lin2 = rhs_out[0] * rhs_out[0] + rhs_out[1] * rhs_out[1] + rhs_out[2] * rhs_out[2];
        const double ang2 = rhs_out[3] * rhs_out[3] + rhs_out[4] * rhs_out[4] + rhs_out[5] * rhs_out[5];
        double worst = 0.0;
        worst = mx2(worst, lin2 / sc[0]);
        worst = mx2(worst, ang2 / sc[1]);
        return worst;
    }
    __device__ __forceinline__ void seed_transpose() {  // newton.rs:209-244, one block: X = Aᵀ/(‖A‖₁‖A‖∞)
        double row_s = 0.0, col_s = 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const double v = fabs(from(Ar[c], r));
                row_s += v;
                col_s += v;
            }
        const double inf = mx2(0.0, row_s), one = mx2(0.0, col_s);
        const double scale = 1.0 / (one * inf);
#pragma unroll
        for (int c = 0; c < 6; ++c) {  // X[row][c] = A[c][row]·scale
            Xr[c] = pick6(from(Ar[0], c), from(Ar[1], c), from(Ar[2], c), from(Ar[3], c), from(Ar[4], c),
                          from(Ar[5], c), row) * scale;
        }
        seeded = true;
    }
    // T = 0·T + A·X, ‖I − T‖², X ← 0·X + X·(2I − T) (newton.rs:255-279), the products
    // term for term as the oracle's gemm: out = 0·out + a·(sign·b + diag·2I).
    __device__ __forceinline__ double refine_inverse() {
        const double two = 1.0 + 1.0;
        double tn[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                double right = 1.0 * from(Xr[c], k);
                if (k == c) right += 0.0 * two;
                s += Ar[k] * right;
            }
            tn[c] = 0.0 * Tr[c] + s;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) Tr[c] = tn[c];
        const double before = reduce_rows(Tr, 1.0);
        double xn[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                double right = -1.0 * from(Tr[c], k);
                if (k == c) right += 1.0 * two;
                s += Xr[k] * right;
            }
            xn[c] = 0.0 * Xr[c] + s;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) Xr[c] = xn[c];
        st_ns++;
        return before;
    }
    __device__ __forceinline__ void run_budget(int budget, int seed_budget) {  // newton.rs:301-327
        bool has_prev = false, reseeded = false;
        double previous = 0.0;
        int remaining = budget;
        while (remaining > 0) {
            remaining -= 1;
            double before = refine_inverse();
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
    __device__ __forceinline__ void newton_solve(double tol2, int ns_warm, int ns_seed, bool& converged, double& best,
                                 bool& assembled) {  // newton.rs:334-530
        const double huge = 1000000000.0;
        world_velocity(pose, mass, tens, mom, vmid);
        for (int k = 0; k < 6; ++k) best_v[k] = vmid[k];
        best = huge;
        converged = false;
        assembled = false;
        bool has_carried = false;
        double carried = 0.0;
        for (int iter = 0; iter < MAX_IT; ++iter) {
            double rnorm;
            st_newton++;
            if (has_carried) {
                has_carried = false;
                rnorm = carried;
            } else {
                rnorm = evaluate(vmid, mid, rhs, Mr);
            }
            if (rnorm < best) {
                best = rnorm;
                for (int k = 0; k < 6; ++k) best_v[k] = vmid[k];
            }
            if (rnorm <= tol2) {
                converged = true;
                break;
            }
            if (!(rnorm < huge)) break;
            // assemble: A = 0·A + 1·mass block (no joint terms)
            for (int c = 0; c < 6; ++c) Ar[c] = 0.0 * Ar[c] + 1.0 * Mr[c];
            assembled = true;
            int budget;
            if (seeded) budget = ns_warm;
            else {
                seed_transpose();
                budget = ns_seed;
            }
            run_budget(budget, ns_seed);
            // dv = 0·dv + X·rhs: row `row` here, the others from the group
            double sum = 0.0;
            for (int c = 0; c < 6; ++c) sum += Xr[c] * rhs[c];
            const double mine = 0.0 * pick6(dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], row) + sum;
            bool nz = false;
            for (int r = 0; r < 6; ++r) {
                dv[r] = from(mine, r);
                nz |= dv[r] != 0.0;
            }
            if (!nz) break;
            double base[6];
            for (int k = 0; k < 6; ++k) base[k] = vmid[k];
            int accepted = -1;
            double tnorm = 0.0;
            for (int l = 0; l < NLANES; ++l) {
                double v[6], m_l[8], r_l[6], brow[6];
                const double alpha = ldexp(1.0, -l);
                for (int k = 0; k < 6; ++k) v[k] = base[k] + dv[k] * alpha;
                const double w = evaluate(v, m_l, r_l, brow);
                if (l == 0)
                    for (int c = 0; c < 6; ++c) Mr[c] = brow[c];  // the iterate's block is lane 0's from here on
                if (w < rnorm) {
                    accepted = l;
                    tnorm = w;
                    for (int k = 0; k < 6; ++k) {
                        vmid[k] = v[k];
                        rhs[k] = r_l[k];
                    }
                    for (int k = 0; k < 8; ++k) mid[k] = m_l[k];
                    if (l > 0)
                        for (int c = 0; c < 6; ++c) Mr[c] = brow[c];  // adopt_lane (cache.rs:1115-1136)
                    break;
                }
            }
            if (accepted >= 0) {
                has_carried = true;
                carried = tnorm;
            } else {
                for (int k = 0; k < 6; ++k) vmid[k] = base[k] + dv[k] * 0.0;
                st_reject++;
                break;
            }
        }
        for (int k = 0; k < 6; ++k) vmid[k] = best_v[k];
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
    __device__ __forceinline__ bool signal_log2(bool assembled, int& r) {  // newton.rs:688-702
        if (!assembled) return false;
        const double an = 0.0 + reduce_rows(Ar, 0.0);
        const double mn = 0.0 + reduce_rows(Mr, 0.0);
        const double stiff = effective_zero(mn) ? 0.0 : an / mn;
        return log2_floor(stiff, r);
    }
    static __device__ int floor_div(int q, int d) { return (q >= 0) ? q / d : -((-q + d - 1) / d); }
    __device__ __forceinline__ void load_x(long long xoff) {
        for (int c = 0; c < 6; ++c) Xr[c] = a.xacc[xoff + row * 6 + c];
    }
    __device__ __forceinline__ void solve_step(double tol2, int ns_warm, int ns_seed, bool writer) {  // newton.rs:543-683
        const long long xoff = a.xacc_off[isl];
        bool accepted_seeded = seeded;
        uint64_t remaining = 1ull << DEPTH;
        int depth = 0;
        while (remaining > 0) {
            uint64_t units = 1ull << (DEPTH - depth);
            while (units > remaining) {
                depth += 1;
                units >>= 1;
            }
            span = dt;
            for (int k = 0; k < depth; ++k) span = span / 2.0;
            half = span / 2.0;
            st_spans++;
            world_momentum(pose, mom, snap);
            bool converged, assembled;
            double best;
            newton_solve(tol2, ns_warm, ns_seed, converged, best, assembled);
            const bool finite = best < 1000000000.0;
            const bool last = depth == DEPTH;
            if (!converged && !last) {
                load_x(xoff);  // restore X_acc
                seeded = accepted_seeded;
                int r, down;
                if (signal_log2(assembled, r)) down = floor_div(r - THRESH, PER_HALVING) + 1;
                else down = 1;
                const int hi = DEPTH - depth;
                down = down < 1 ? 1 : (down > hi ? hi : down);
                depth += down;
                continue;
            }
            if (converged || finite) {
                if (!converged) st_uncommit++;
                commit_body(pose, mom, mass, tens, vmid, snap, span, half, converged);
                if (writer)
                    for (int c = 0; c < 6; ++c) a.xacc[xoff + row * 6 + c] = Xr[c];  // publish X_acc
                const double nx = 0.0 + reduce_rows(Xr, 0.0);
                const double huge = 1000000000.0;
                if (nx < huge * huge) accepted_seeded = seeded;
                else {
                    seeded = false;
                    accepted_seeded = false;
                }
            } else {
                load_x(xoff);
                seeded = accepted_seeded;
                st_nonfinite++;
            }
            remaining -= units;
            auto aligned = [&](int d) {
                if (d <= 0) return false;
                uint64_t bigger = 1ull << (DEPTH - (d - 1));
                return bigger <= remaining && remaining % bigger == 0;
            };
            int r;
            if (remaining > 0 && converged && aligned(depth) && signal_log2(assembled, r)) {
                int up = floor_div(THRESH - r, PER_HALVING);
                while (up > 0 && aligned(depth)) {
                    depth -= 1;
                    up -= 1;
                }
            }
        }
    }
};

// Islands of the requeue list, one group of eight lanes each (every group of the grid
// walks the list by the grid's group count). The list's length is read at the start;
// the workgroup that finishes last clears the length and its own counter.
__global__ __launch_bounds__(SB8) void singles_requeue_kernel(IslandArgs a) {
    const int j = threadIdx.x & (SG - 1);
    const unsigned count = a.requeue_ctl[0];
    const unsigned groups = gridDim.x * (SB8 / SG);
    for (unsigned q = (blockIdx.x * SB8 + threadIdx.x) / SG; q < count; q += groups) {
        const int isl = a.requeue[q];
        const int b = a.island_start[isl];
        const bool writer = j < 6;  // lanes 6 and 7 mirror row 5
        Single sg{a, j < 6 ? j : 5, isl};
        for (int t = 0; t < 8; ++t) sg.pose[t] = a.pose[(size_t)b * 8 + t];
        for (int t = 0; t < 6; ++t) {
            sg.mom[t] = a.mom[(size_t)b * 6 + t];
            sg.ext[t] = field_add(a, a.mass[b], t, a.ext[(size_t)b * 6 + t]);
            sg.dv[t] = 0.0;
            sg.Ar[t] = 0.0;
            sg.Tr[t] = 0.0;
            sg.Mr[t] = 0.0;
        }
        for (int t = 0; t < 9; ++t) sg.tens[t] = a.tens[(size_t)b * 9 + t];
        sg.mass = a.mass[b];
        sg.load_x(a.xacc_off[isl]);
        sg.dt = a.dt;
        sg.warp = a.warp;
        sg.seeded = a.seeded[isl] != 0;
        sg.solve_step(a.tol2, a.ns_warm, a.ns_seed, writer);
        StepCounts n;
        n.spans = sg.st_spans;
        n.newton = sg.st_newton;
        n.ns = sg.st_ns;
        n.reseed = sg.st_reseed;
        n.reject = sg.st_reject;
        n.uncommit = sg.st_uncommit;
        n.nonfinite = sg.st_nonfinite;
        finish_island(a, isl, b, sg.pose, sg.mom, sg.mass, sg.seeded, n, j == 0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&a.requeue_ctl[1], 1u) == gridDim.x - 1) {  // every workgroup has read the length
            a.requeue_ctl[0] = 0;
            a.requeue_ctl[1] = 0;
        }
    }
}

}  // namespace

void launch_singles(const IslandArgs& a, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    if (a.n_islands <= 0) return;
    const bool requeue = a.requeue && a.requeue_ctl;
    launch_k(singles_step_kernel, dim3((unsigned)((a.n_islands + SB - 1) / SB)), dim3(SB), stream, start,
             requeue ? nullptr : stop, a);
    if (!requeue) return;
    // a grid for the worst case up to one group per island, at most 1,024 workgroups
    const long long groups = std::min<long long>(a.n_islands, 1024LL * (SB8 / SG));
    launch_k(singles_requeue_kernel, dim3((unsigned)((groups * SG + SB8 - 1) / SB8)), dim3(SB8), stream, nullptr,
             stop, a);
}

}  // namespace exg
