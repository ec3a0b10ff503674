// One implicit Newton step of single-body islands without joints: the control flow of
// island.hip (crates/newton/src/integrator/implicit/newton.rs:334-683, 209-327,
// 688-753) restated for a single 6×6 block, in the reference's own summation order
// (the oracle's term order for gemm, reduce, matvec, seed).
//
// Free bodies are most of the batched and served worlds (60 % of their bodies). Eight
// lanes serve one island: lane r (r < 6; lanes 6 and 7 mirror row 5 and store
// nothing) holds row r of A, X, the product T of a Newton–Schulz iteration and the
// candidate-0 mass block in registers and forms that row of every product; the other
// rows come from the group by lane shuffles. The body's own arithmetic (midpoint,
// residual, commit) runs in every lane on the same values, so the eight lanes take
// the same decisions. A wave holds eight islands and takes as long as the slowest of
// them: a body whose step subdivides holds up seven others, not sixty-three.
//
// Line-search lanes are evaluated in order with early exit: the first lane whose
// residual drops below the current one is accepted, which is the lane the reference's
// side-by-side evaluation accepts; lane 0's mass block is kept for signal_log2 when
// every lane fails, as in the reference.
#include <hip/hip_runtime.h>
#include "island.h"
#include "physics_dev.h"
#include "body_rows.h"

namespace exg {

namespace {

constexpr int SG = 8;            // lanes per island
constexpr int SB = 256;          // threads per workgroup (32 islands)
constexpr int MAX_IT = 32;
constexpr int DEPTH = 8;
constexpr int THRESH = 8;
constexpr int PER_HALVING = 4;
constexpr int NLANES = 8;

__device__ inline double mx2(double x, double y) { return x > y ? x : y; }
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

__global__ __launch_bounds__(SB) void singles_step_kernel(IslandArgs a) {
    const int j = threadIdx.x & (SG - 1);
    const int k = (int)((blockIdx.x * SB + threadIdx.x) / SG);
    // whole groups leave together: the eight lanes of a group share k
    if (k >= a.n_islands) return;
    const int isl = a.island_ids ? a.island_ids[k] : k;
    const int b = a.island_start[isl];
    if (a.island_end[isl] - b != 1 || a.island_jend[isl] != a.island_jstart[isl]) return;  // routed by class; never taken
    const bool writer = j < 6;  // lanes 6 and 7 mirror row 5
    Single sg{a, j < 6 ? j : 5, isl};
    for (int q = 0; q < 8; ++q) sg.pose[q] = a.pose[(size_t)b * 8 + q];
    for (int q = 0; q < 6; ++q) {
        sg.mom[q] = a.mom[(size_t)b * 6 + q];
        sg.ext[q] = a.ext[(size_t)b * 6 + q];
        sg.dv[q] = 0.0;
        sg.Ar[q] = 0.0;
        sg.Tr[q] = 0.0;
        sg.Mr[q] = 0.0;
    }
    for (int q = 0; q < 9; ++q) sg.tens[q] = a.tens[(size_t)b * 9 + q];
    sg.mass = a.mass[b];
    sg.load_x(a.xacc_off[isl]);
    sg.dt = a.dt;
    sg.warp = a.warp;
    sg.seeded = a.seeded[isl] != 0;
    sg.solve_step(a.tol2, a.ns_warm, a.ns_seed, writer);

    // recompute_centroid (island.rs:206-244) for one body, then the charge
    double p[3];
    body_position(sg.pose, p);
    const double total_m = sg.mass;
    double wx[3] = {0.0 + p[0] * sg.mass, 0.0 + p[1] * sg.mass, 0.0 + p[2] * sg.mass};
    double org[3] = {a.origin[isl * 3 + 0], a.origin[isl * 3 + 1], a.origin[isl * 3 + 2]};
    if (!effective_zero(total_m)) {
        double inv = 1.0 / total_m;
        double sh[6] = {-(wx[0] * inv), -(wx[1] * inv), -(wx[2] * inv), 0.0, 0.0, 0.0};
        double shift[8], np[8];
        twist_exp(sh, 1.0, shift);
        compose(sg.pose, shift, np);
        for (int q = 0; q < 8; ++q) sg.pose[q] = np[q];
        for (int q = 0; q < 3; ++q) org[q] += wx[q] * inv;
    }
    if (j != 0) return;
    for (int q = 0; q < 3; ++q) a.origin[isl * 3 + q] = org[q];
    body_position(sg.pose, p);
    double* c = a.charges_out + (size_t)b * 8;
    for (int q = 0; q < 3; ++q) {
        c[q] = p[q] + org[q];
        c[3 + q] = p[q];
    }
    c[6] = (a.G_of_island ? a.G_of_island[isl] : a.G) * sg.mass;
    c[7] = a.size[b];
    for (int q = 0; q < 8; ++q) a.pose[(size_t)b * 8 + q] = sg.pose[q];
    for (int q = 0; q < 6; ++q) a.mom[(size_t)b * 6 + q] = sg.mom[q];
    a.seeded[isl] = sg.seeded ? 1 : 0;
    int* st = a.stats + (size_t)isl * ISLAND_STATS;
    st[0] += sg.st_spans;
    st[1] += sg.st_newton;
    st[2] += sg.st_ns;
    st[3] += sg.st_reseed;
    st[4] += sg.st_reject;
    st[5] += sg.st_uncommit;
    st[STAT_NONFINITE_SPANS] += sg.st_nonfinite;
    bool bad = false;
    for (int q = 0; q < 8; ++q) bad |= !isfinite(sg.pose[q]);
    for (int q = 0; q < 6; ++q) bad |= !isfinite(sg.mom[q]);
    if (bad) mark_fatal(a, isl);
}

}  // namespace

void launch_singles(const IslandArgs& a, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    if (a.n_islands <= 0) return;
    const long long lanes = (long long)a.n_islands * SG;
    launch_k(singles_step_kernel, dim3((unsigned)((lanes + SB - 1) / SB)), dim3(SB), stream, start, stop, a);
}

}  // namespace exg
