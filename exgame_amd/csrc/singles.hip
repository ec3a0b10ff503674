// One implicit Newton step of single-body islands without joints, one thread per
// island: the control flow of island.hip (crates/newton/src/integrator/implicit/newton.rs:
// 334-683, 209-327, 688-753) restated for a single 6×6 block, in the reference's own
// summation order (the oracle's term order for gemm, reduce, matvec, seed).
//
// Free bodies are most of the batched and served worlds (60 % of their bodies); a
// workgroup per such island would leave 63 of 64 lanes idle. Here 64 islands share a
// wave and each lane runs its island's whole step. The four 6×6 matrices of a lane
// (A, X, T = scratch product, the candidate-0 mass block) live in LDS laid out
// element-major ([36][64] per matrix), so the 64 lanes touch 64 consecutive doubles
// at every access.
//
// Line-search lanes are evaluated in order with early exit: the first lane whose
// residual drops below the current one is accepted, which is the lane the reference's
// side-by-side evaluation accepts; lane 0's mass block is kept for signal_log2 when
// every lane fails, as in the reference.
#include <hip/hip_runtime.h>
#include "island.h"
#include "physics_dev.h"

namespace exg {

namespace {

constexpr int ST = 64;           // threads (islands) per workgroup
constexpr int MAX_IT = 32;
constexpr int DEPTH = 8;
constexpr int THRESH = 8;
constexpr int PER_HALVING = 4;
constexpr int NLANES = 8;

struct SingleShared {
    double A[36][ST];
    double X[36][ST];
    double T[36][ST];
    double M0[36][ST];   // candidate-0 (iterate) world mass block
};

__device__ inline double mx2(double x, double y) { return x > y ? x : y; }

// Σ over one block of (diag·δ − b)², from 0, row-major (fn_reduce_sq_term)
__device__ inline double reduce_block(const double (*b)[ST], int t, double diag) {
    double out = 0.0;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double v = b[r * 6 + c][t];
            double e = (r == c) ? diag - v : -v;
            out += e * e;
        }
    return 0.0 + out;
}

// out = 0·out + a·(sign·b + diag·2I)  (fn_gemm_term with one neighbour block),
// a row at a time. `out` may be `a` (not `b`): row r of the result needs only row r
// of `a`, which is read before the row is stored.
// The right operand (sign·b + diag·2I) is formed once in registers.
__device__ inline void gemm_block(const double (*a)[ST], const double (*b)[ST], double (*out)[ST], int t,
                                  bool two_minus) {
    const double sign = two_minus ? -1.0 : 1.0, diag = two_minus ? 1.0 : 0.0, two = 1.0 + 1.0;
    double right[36];
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double v = sign * b[k * 6 + c][t];
            if (k == c) v += diag * two;
            right[k * 6 + c] = v;
        }
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        double ar[6], res[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) ar[k] = a[r * 6 + k][t];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += ar[k] * right[k * 6 + c];
            res[c] = 0.0 * out[r * 6 + c][t] + s;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) out[r * 6 + c][t] = res[c];
    }
}

// Midpoint, residual, mass block and residual ratio of one candidate velocity
// (pre + gather + body_post + wnorm for one body without joints). Out of line: one
// copy serves the iterate and every line-search lane; it touches no LDS.
__device__ __forceinline__ double evaluate_single(const double pose[8], const double v[6], double mass,
                                               const double tens[9], const double snap[6], const double ext[6],
                                               double half, double mid_out[8], double rhs_out[6],
                                               double block[6][6]) {
    double e[8], sc[2];
    twist_exp(v, half, e);
    compose(pose, e, mid_out);
    const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
    body_post(mid_out, v, mass, tens, snap, ext, half, floor2, block, rhs_out, sc);
    double lin2 = rhs_out[0] * rhs_out[0] + rhs_out[1] * rhs_out[1] + rhs_out[2] * rhs_out[2];
    double ang2 = rhs_out[3] * rhs_out[3] + rhs_out[4] * rhs_out[4] + rhs_out[5] * rhs_out[5];
    double worst = 0.0;
    worst = mx2(worst, lin2 / sc[0]);
    worst = mx2(worst, ang2 / sc[1]);
    return worst;
}

struct Single {
    SingleShared& s;
    const IslandArgs& a;
    int t, isl;
    double pose[8], mom[6], mass, tens[9], ext[6];
    double snap[6], vmid[6], mid[8], rhs[6], best_v[6], dv[6];
    double dt, warp, half, span;
    bool seeded;
    int st_spans = 0, st_newton = 0, st_ns = 0, st_reseed = 0, st_reject = 0, st_uncommit = 0;

    // midpoint, residual and mass block of a candidate velocity (wnorm: newton.rs:139-151)
    __device__ double evaluate(const double v[6], double mid_out[8], double rhs_out[6], double block[6][6]) {
        return evaluate_single(pose, v, mass, tens, snap, ext, half, mid_out, rhs_out, block);
    }
    __device__ void store_m0(const double block[6][6]) {
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) s.M0[r * 6 + c][t] = block[r][c];
    }
    __device__ void seed_transpose() {  // newton.rs:209-244
        double row = 0.0, col = 0.0;
        for (int e = 0; e < 36; ++e) {
            row += fabs(s.A[e][t]);
            col += fabs(s.A[e][t]);
        }
        double inf = mx2(0.0, row), one = mx2(0.0, col);
        double scale = 1.0 / (one * inf);
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) s.X[r * 6 + c][t] = s.A[c * 6 + r][t] * scale;
        seeded = true;
    }
    __device__ double refine_inverse() {  // newton.rs:255-279
        gemm_block(s.A, s.X, s.T, t, false);
        double before = reduce_block(s.T, t, 1.0);
        gemm_block(s.X, s.T, s.X, t, true);  // in place: X is the left operand
        st_ns++;
        return before;
    }
    __device__ void run_budget(int budget, int seed_budget) {  // newton.rs:301-327
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
    __device__ void newton_solve(double tol2, int ns_warm, int ns_seed, bool& converged, double& best,
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
                double block[6][6];
                rnorm = evaluate(vmid, mid, rhs, block);
                store_m0(block);
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
            for (int e = 0; e < 36; ++e) s.A[e][t] = 0.0 * s.A[e][t] + 1.0 * s.M0[e][t];
            assembled = true;
            int budget;
            if (seeded) budget = ns_warm;
            else {
                seed_transpose();
                budget = ns_seed;
            }
            run_budget(budget, ns_seed);
            bool nz = false;
            for (int r = 0; r < 6; ++r) {  // dv = 0·dv + X·rhs
                double sum = 0.0;
                for (int c = 0; c < 6; ++c) sum += s.X[r * 6 + c][t] * rhs[c];
                dv[r] = 0.0 * dv[r] + sum;
                nz |= dv[r] != 0.0;
            }
            if (!nz) break;
            double base[6];
            for (int k = 0; k < 6; ++k) base[k] = vmid[k];
            int accepted = -1;
            double tnorm = 0.0;
            for (int l = 0; l < NLANES; ++l) {
                double v[6], m_l[8], r_l[6], block[6][6];
                const double alpha = ldexp(1.0, -l);
                for (int k = 0; k < 6; ++k) v[k] = base[k] + dv[k] * alpha;
                double w = evaluate(v, m_l, r_l, block);
                if (l == 0) store_m0(block);  // the iterate's storage holds lane 0 from here on
                if (w < rnorm) {
                    accepted = l;
                    tnorm = w;
                    for (int k = 0; k < 6; ++k) {
                        vmid[k] = v[k];
                        rhs[k] = r_l[k];
                    }
                    for (int k = 0; k < 8; ++k) mid[k] = m_l[k];
                    if (l > 0) store_m0(block);  // adopt_lane (cache.rs:1115-1136)
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
    __device__ bool signal_log2(bool assembled, int& r) {  // newton.rs:688-702
        if (!assembled) return false;
        double an = 0.0 + reduce_block(s.A, t, 0.0);
        double mn = 0.0 + reduce_block(s.M0, t, 0.0);
        double stiff = effective_zero(mn) ? 0.0 : an / mn;
        return log2_floor(stiff, r);
    }
    static __device__ int floor_div(int q, int d) { return (q >= 0) ? q / d : -((-q + d - 1) / d); }
    __device__ void solve_step(double tol2, int ns_warm, int ns_seed) {  // newton.rs:543-683
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
            bool finite = best < 1000000000.0;
            bool last = depth == DEPTH;
            if (!converged && !last) {
                for (int e = 0; e < 36; ++e) s.X[e][t] = a.xacc[xoff + e];  // restore X_acc
                seeded = accepted_seeded;
                int r, down;
                if (signal_log2(assembled, r)) down = floor_div(r - THRESH, PER_HALVING) + 1;
                else down = 1;
                int hi = DEPTH - depth;
                down = down < 1 ? 1 : (down > hi ? hi : down);
                depth += down;
                continue;
            }
            if (converged || finite) {
                if (!converged) st_uncommit++;
                commit_body(pose, mom, mass, tens, vmid, snap, span, half, converged);
                for (int e = 0; e < 36; ++e) a.xacc[xoff + e] = s.X[e][t];  // publish X_acc
                double nx = 0.0 + reduce_block(s.X, t, 0.0);
                const double huge = 1000000000.0;
                if (nx < huge * huge) accepted_seeded = seeded;
                else {
                    seeded = false;
                    accepted_seeded = false;
                }
            } else {
                for (int e = 0; e < 36; ++e) s.X[e][t] = a.xacc[xoff + e];
                seeded = accepted_seeded;
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

__global__ __launch_bounds__(ST) void singles_step_kernel(IslandArgs a) {
    __shared__ SingleShared s;
    const int t = threadIdx.x;
    const int k = blockIdx.x * ST + t;
    if (k >= a.n_islands) return;  // no barriers below: each lane is independent
    const int isl = a.island_ids ? a.island_ids[k] : k;
    const int b = a.island_start[isl];
    if (a.island_end[isl] - b != 1 || a.island_jend[isl] != a.island_jstart[isl]) return;  // routed by class; never taken
    Single sg{s, a, t, isl};
    for (int q = 0; q < 8; ++q) sg.pose[q] = a.pose[(size_t)b * 8 + q];
    for (int q = 0; q < 6; ++q) {
        sg.mom[q] = a.mom[(size_t)b * 6 + q];
        sg.ext[q] = a.ext[(size_t)b * 6 + q];
        sg.dv[q] = 0.0;
    }
    for (int q = 0; q < 9; ++q) sg.tens[q] = a.tens[(size_t)b * 9 + q];
    sg.mass = a.mass[b];
    const long long xoff = a.xacc_off[isl];
    for (int e = 0; e < 36; ++e) {
        s.A[e][t] = 0.0;
        s.T[e][t] = 0.0;
        s.M0[e][t] = 0.0;
        s.X[e][t] = a.xacc[xoff + e];
    }
    sg.dt = a.dt;
    sg.warp = a.warp;
    sg.seeded = a.seeded[isl] != 0;
    sg.solve_step(a.tol2, a.ns_warm, a.ns_seed);

    // recompute_centroid (island.rs:206-244) for one body, then the charge
    double p[3];
    body_position(sg.pose, p);
    const double total_m = sg.mass;
    double wx[3] = {0.0 + p[0] * sg.mass, 0.0 + p[1] * sg.mass, 0.0 + p[2] * sg.mass};
    if (!effective_zero(total_m)) {
        double inv = 1.0 / total_m;
        double sh[6] = {-(wx[0] * inv), -(wx[1] * inv), -(wx[2] * inv), 0.0, 0.0, 0.0};
        double shift[8], np[8];
        twist_exp(sh, 1.0, shift);
        compose(sg.pose, shift, np);
        for (int q = 0; q < 8; ++q) sg.pose[q] = np[q];
        for (int q = 0; q < 3; ++q) a.origin[isl * 3 + q] += wx[q] * inv;
    }
    body_position(sg.pose, p);
    double* c = a.charges_out + (size_t)b * 8;
    for (int q = 0; q < 3; ++q) {
        c[q] = p[q] + a.origin[isl * 3 + q];
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
}

}  // namespace

void launch_singles(const IslandArgs& a, hipStream_t stream) {
    if (a.n_islands <= 0) return;
    hipLaunchKernelGGL(singles_step_kernel, dim3((a.n_islands + ST - 1) / ST), dim3(ST), 0, stream, a);
}

}  // namespace exg
