// ORACLE — test infrastructure only. CPU f64 restatement of the reference's
// implicit-midpoint Newton solver with its full host control flow, one island at
// a time, plus the island re-anchoring and the world step with N-body gravity.
//
//   Newton::step_all / solve_step / newton_solve   crates/newton/src/integrator/implicit/newton.rs:334-861
//   run_budget / refine_inverse / seed_transpose   newton.rs:209-327
//   commit_span / signal_log2 / log2_floor         newton.rs:688-753, 101-121
//   NewtonCache sync / lanes / checkpoint          crates/newton/src/integrator/implicit/cache.rs:566-1166
//   baked row semantics (a1·out + …)               crates/newton/build.rs:1156-1403
//   Island::recompute_centroid                     crates/newton/src/mechanism/island.rs:206-244
//   Mechanism::step phase order                    crates/newton/src/mechanism/mechanism.rs:526-600
//
// The World slots of the reference become plain arrays; lane 0 aliases the main
// iterate as it does there.
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>
#include "physics.hpp"

namespace oracle {

struct NewtonParams {
    double tol2 = 1e-8;          // (1e-4)², newton.rs:802-805; the f64 workloads pass their own
    int ns_warm = 2;             // NS_ITERATIONS_DEFAULT, newton.rs:58
    int ns_seed = 24;            // NS_SEED_ITERATIONS_DEFAULT, newton.rs:61
};

constexpr int LANES = 8;                 // cache.rs:69
constexpr int MAX_ITERATIONS = 32;       // newton.rs:355
constexpr int MAX_DEPTH = 8;             // newton.rs:557
constexpr int THRESHOLD_LOG2 = 8;        // newton.rs:570
constexpr int PER_HALVING_LOG2 = 4;      // newton.rs:571
constexpr int REDUCE_MAX_TERMS = 126;    // rows/reduce.rs: ROW − HEAD

struct BlockTerm { int joint; int row_end; int col_end; };

struct SolveStats {
    int spans = 0;
    int newton_iterations = 0;
    int ns_iterations = 0;
    int reseeds = 0;
    int line_search_rejects = 0;
    int unconverged_commits = 0;
};

// One attempted span of solve_step: its depth, whether Newton converged, whether the
// span was committed, and the stiffness signal ⌊log₂(‖A‖²/‖𝕀‖²)⌋ read after it (for a
// rejected span: the deepening input; for a committed one: the growth input), or
// NO_SIGNAL when none was read.
constexpr int NO_SIGNAL = -100000;
struct SpanEvent { int depth, converged, committed, signal; };

struct Island {
    std::vector<Body> bodies;   // dynamic and kinematic bodies, in island order
    std::vector<Joint> joints;
    double origin[3] = {0.0, 0.0, 0.0};

    // NewtonCache (cache.rs:200-320): the unknowns are the dynamic bodies, in island
    // order (`order`, newton.rs:806-830); block row p belongs to body order[p]
    int m = -1;
    std::vector<int> order, pos;   // pos[body]: its block row, −1 for a kinematic body
    std::vector<Block> A, X, Xn, Xacc, R;
    std::vector<std::vector<int>> neighbours;
    std::vector<std::vector<BlockTerm>> block_terms;
    bool seeded = false, accepted_seeded = false;
    SolveStats stats;
    std::vector<SpanEvent> trace;   // the spans of the last step
};

// Per body (every body of the island): vmid, midpoint, solve_vel, total; per block
// row (the unknowns): rhs, scale, mass.
struct Lane {
    std::vector<M16> vmid, midpoint, solve_vel;
    std::vector<std::array<std::array<double, 6>, 2>> conn;
    std::vector<std::array<double, 6>> total, rhs, scale;
    std::vector<Block> mass;
    void resize(int nb, int m, int nj) {
        vmid.resize(nb); midpoint.resize(nb); solve_vel.resize(nb); conn.resize(nj);
        total.resize(nb); rhs.resize(m); scale.resize(m); mass.assign(m, zero_block());
    }
};

// cache.sync: dimension latch, sparsity and the assembly table (cache.rs:566-788)
inline void cache_sync(Island& isl) {
    const int nb = (int)isl.bodies.size();
    isl.order.clear();
    isl.pos.assign(nb, -1);
    for (int i = 0; i < nb; ++i)
        if (!is_kinematic(isl.bodies[i].inertia)) {
            isl.pos[i] = (int)isl.order.size();
            isl.order.push_back(i);
        }
    int m = (int)isl.order.size();
    if (m != isl.m) {
        isl.A.assign(m * m, zero_block());
        isl.X.assign(m * m, zero_block());
        isl.Xn.assign(m * m, zero_block());
        isl.Xacc.assign(m * m, zero_block());
        isl.R.assign(m * m, zero_block());
        isl.seeded = false;
        isl.accepted_seeded = false;
        isl.m = m;
    }
    isl.neighbours.assign(m, {});
    for (int i = 0; i < m; ++i) isl.neighbours[i].push_back(i);
    isl.block_terms.assign(m * m, {});
    for (int jn = 0; jn < (int)isl.joints.size(); ++jn) {
        const Joint& e = isl.joints[jn];
        int ia = isl.pos[e.a], ib = isl.pos[e.b];
        if (ia >= 0 && ib >= 0) {   // cache.rs:700-707: neighbours among the unknowns
            auto& na = isl.neighbours[ia];
            if (std::find(na.begin(), na.end(), ib) == na.end()) na.push_back(ib);
            auto& nbr = isl.neighbours[ib];
            if (std::find(nbr.begin(), nbr.end(), ia) == nbr.end()) nbr.push_back(ia);
        }
        int ends[2] = {ia, ib};
        for (int re = 0; re < 2; ++re)        // cache.rs:708-720: no row or column for a kinematic end
            for (int ce = 0; ce < 2; ++ce)
                if (ends[re] >= 0 && ends[ce] >= 0) isl.block_terms[ends[re] * m + ends[ce]].push_back({jn, re, ce});
    }
}

struct NewtonCtx {
    Island& isl;
    const NewtonParams& prm;
    double dt, warp;
    const std::vector<std::array<double, 6>>& ext;   // external (field) wrenches, island frame
    std::vector<std::array<double, 6>> snap_mom;      // world momentum at span start
    std::vector<Lane> lanes;
    std::vector<JointJac> jac;
    std::vector<std::array<double, 6>> dv;
    double half = 0.0;

    NewtonCtx(Island& i, const NewtonParams& p, double dt_, double warp_, const std::vector<std::array<double, 6>>& e)
        : isl(i), prm(p), dt(dt_), warp(warp_), ext(e) {
        int nb = (int)isl.bodies.size(), nj = (int)isl.joints.size();
        lanes.resize(LANES);
        for (auto& l : lanes) l.resize(nb, isl.m, nj);
        jac.resize(nj);
        dv.assign(isl.m, {});
        snap_mom.assign(nb, {});
    }
    int m() const { return isl.m; }
    int nb() const { return (int)isl.bodies.size(); }

    // PRE + connection kernels + GATHER for one lane (accelerator/mod.rs:649-670)
    void dispatch(Lane& L, bool jacobian) {
        const int nb = this->nb();
        for (int i = 0; i < nb; ++i) {
            L.midpoint[i] = fn_pre(isl.bodies[i].pose, L.vmid[i], half);
            L.solve_vel[i] = L.vmid[i];
        }
        eval_joints(L, jacobian);
        for (int i = 0; i < nb; ++i) L.total[i] = ext[i];
        for (int jn = 0; jn < (int)isl.joints.size(); ++jn) {
            const Joint& e = isl.joints[jn];
            for (int s = 0; s < 6; ++s) L.total[e.a][s] = L.total[e.a][s] + L.conn[jn][0][s];
            for (int s = 0; s < 6; ++s) L.total[e.b][s] = L.total[e.b][s] + L.conn[jn][1][s];
        }
    }
    void eval_joints(Lane& L, bool jacobian) {
        for (int jn = 0; jn < (int)isl.joints.size(); ++jn) {
            const Joint& e = isl.joints[jn];
            M16 poses[2] = {L.midpoint[e.a], L.midpoint[e.b]};
            M16 vels[2] = {L.solve_vel[e.a], L.solve_vel[e.b]};
            if (jacobian) {
                joint_jacobian(e, poses, vels, dt, warp, jac[jn]);
                for (int t = 0; t < 2; ++t)
                    for (int s = 0; s < 6; ++s) L.conn[jn][t][s] = jac[jn].value[t][s];
            } else {
                double out[2][6];
                joint_plain(e, poses, vels, dt, warp, out);
                for (int t = 0; t < 2; ++t)
                    for (int s = 0; s < 6; ++s) L.conn[jn][t][s] = out[t][s];
            }
        }
    }
    void body_post(Lane& L) {
        const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
        for (int p = 0; p < m(); ++p) {   // the unknowns only (body_post.rs:105-118)
            const int i = isl.order[p];
            PostOut o = fn_body_post(L.midpoint[i], L.solve_vel[i], isl.bodies[i].inertia, snap_mom[i].data(),
                                     L.total[i].data(), half, floor2);
            L.mass[p] = o.block;
            for (int s = 0; s < 6; ++s) { L.rhs[p][s] = o.rhs[s]; L.scale[p][s] = o.scale[s]; }
        }
    }
    static double max2(double a, double b) { return a > b ? a : b; }  // newton.rs:123-126
    double wnorm(const Lane& L) const {  // newton.rs:139-151
        double worst = 0.0;
        for (int i = 0; i < m(); ++i) {
            const auto& r = L.rhs[i];
            const auto& s = L.scale[i];
            double lin2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            double ang2 = r[3] * r[3] + r[4] * r[4] + r[5] * r[5];
            worst = max2(worst, lin2 / s[0]);
            worst = max2(worst, ang2 / s[3]);
        }
        return worst;
    }
    // AssembleBlock rows (build.rs:1310-1403): acc = a1·out + mass_scale·mass, then terms
    void assemble() {
        int m = this->m();
        double pf = -(half * half), vf = -half;
        const Lane& L0 = lanes[0];
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                Block& out = isl.A[i * m + j];
                double ms = (i == j) ? 1.0 : 0.0;
                Block acc;
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) acc[r][c] = 0.0 * out[r][c] + ms * L0.mass[i][r][c];
                for (const BlockTerm& t : isl.block_terms[i * m + j]) {
                    double pose[6][6], vel[6][6];
                    for (int col = 0; col < 6; ++col)
                        for (int s = 0; s < 6; ++s) {
                            pose[col][s] = jac[t.joint].jac[t.row_end][t.col_end * 6 + col][s];
                            vel[col][s] = jac[t.joint].jac[t.row_end][12 + t.col_end * 6 + col][s];
                        }
                    fn_assemble_term(acc, pose, vel, pf, vf);
                }
                out = acc;
            }
    }
    // Gemm rows: out = a1·out + Σ_k a[i,k]·(sign·b[k,j] + diag·2I) (build.rs:1239-1295)
    static void gemm(int m, const std::vector<Block>& a, const std::vector<Block>& b, std::vector<Block>& out,
                     const std::vector<std::vector<int>>& rows_of, bool two_minus) {
        double sign = two_minus ? -1.0 : 1.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                Block acc;
                const Block& old = out[i * m + j];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) acc[r][c] = 0.0 * old[r][c];
                for (int k : rows_of[i]) {
                    double diag = (two_minus && k == j) ? 1.0 : 0.0;
                    fn_gemm_term(acc, a[i * m + k], b[k * m + j], sign, diag);
                }
                out[i * m + j] = acc;
            }
    }
    // BlockReduce rows: partials over chunks of 126 blocks, summed in row order on the
    // host (rows/reduce.rs, cache.rs:494-498)
    static double reduce(const std::vector<Block>& blocks, int m, bool identity_diag) {
        double total = 0.0;
        int n = (int)blocks.size();
        int nchunks = std::max(1, (n + REDUCE_MAX_TERMS - 1) / REDUCE_MAX_TERMS);
        for (int ch = 0; ch < nchunks; ++ch) {
            double part = 0.0;
            for (int k = ch * REDUCE_MAX_TERMS; k < std::min(n, (ch + 1) * REDUCE_MAX_TERMS); ++k) {
                double diag = (identity_diag && (k / std::max(m, 1) == k % std::max(m, 1))) ? 1.0 : 0.0;
                part = fn_reduce_sq_term(part, blocks[k], diag);
            }
            total = total + part;
        }
        return total;
    }
    double refine_inverse() {  // newton.rs:255-279
        int m = this->m();
        if (m == 0) return 0.0;
        gemm(m, isl.A, isl.X, isl.R, isl.neighbours, false);
        double before = reduce(isl.R, m, true);
        std::vector<std::vector<int>> dense(m);
        for (int i = 0; i < m; ++i) for (int k = 0; k < m; ++k) dense[i].push_back(k);
        gemm(m, isl.X, isl.R, isl.Xn, dense, true);
        std::swap(isl.X, isl.Xn);
        isl.stats.ns_iterations++;
        return before;
    }
    void seed_transpose() {  // newton.rs:209-244
        int m = this->m();
        double one = 0.0, inf = 0.0;
        for (int i = 0; i < m; ++i) {
            double row = 0.0, col = 0.0;
            for (int j = 0; j < m; ++j) {
                const Block& rb = isl.A[i * m + j];
                const Block& cb = isl.A[j * m + i];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) {
                        row += std::fabs(rb[r][c]);
                        col += std::fabs(cb[r][c]);
                    }
            }
            inf = max2(inf, row);
            one = max2(one, col);
        }
        double scale = 1.0 / (one * inf);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const Block& src = isl.A[j * m + i];
                Block dst = zero_block();
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) dst[r][c] = src[c][r] * scale;
                isl.X[i * m + j] = dst;
            }
        isl.seeded = true;
    }
    void run_budget(int budget, int seed_budget) {  // newton.rs:301-327
        const double margin = 4.0;
        bool has_prev = false;
        double previous = 0.0;
        bool reseeded = false;
        int remaining = budget;
        while (remaining > 0) {
            remaining -= 1;
            double before = refine_inverse();
            if (has_prev && !(before < margin * previous) && !reseeded) {
                seed_transpose();
                reseeded = true;
                has_prev = false;
                remaining = seed_budget;
                isl.stats.reseeds++;
                continue;
            }
            previous = before;
            has_prev = true;
        }
    }
    // BlockMatVec rows: dv_i = a1·dv_i + Σ_j X[i,j]·rhs_j (build.rs:1161-1232)
    void block_matvec() {
        int m = this->m();
        const Lane& L0 = lanes[0];
        for (int i = 0; i < m; ++i) {
            double acc[6];
            for (int s = 0; s < 6; ++s) acc[s] = 0.0 * dv[i][s];
            for (int j = 0; j < m; ++j) fn_matvec_term(acc, isl.X[i * m + j], L0.rhs[j].data());
            for (int s = 0; s < 6; ++s) dv[i][s] = acc[s];
        }
    }
    // newton_solve (newton.rs:334-530); returns converged, best, assembled; leaves
    // the best iterate in lane 0's vmid.
    void newton_solve(bool& converged, double& best, bool& assembled) {
        int m = this->m();
        const double huge = 1000000000.0;
        Lane& L0 = lanes[0];
        // every body's world velocity, a kinematic one's prescribed (newton.rs:357-376)
        for (int i = 0; i < nb(); ++i) L0.vmid[i] = world_velocity(isl.bodies[i]);
        for (int l = 1; l < LANES; ++l)
            for (int i = 0; i < nb(); ++i) lanes[l].vmid[i] = L0.vmid[i];
        best = huge;
        std::vector<M16> best_vals = L0.vmid;
        converged = false;
        bool has_carried = false;
        double carried = 0.0;
        assembled = false;
        for (int iter = 0; iter < MAX_ITERATIONS; ++iter) {
            double rnorm;
            isl.stats.newton_iterations++;
            if (has_carried) {
                has_carried = false;
                eval_joints(L0, true);
                rnorm = carried;
            } else {
                dispatch(L0, true);
                body_post(L0);
                rnorm = wnorm(L0);
            }
            if (rnorm < best) { best = rnorm; best_vals = L0.vmid; }
            if (rnorm <= prm.tol2) { converged = true; break; }
            if (!(rnorm < huge)) break;
            assemble();
            assembled = true;
            int budget;
            if (isl.seeded) budget = prm.ns_warm;
            else { seed_transpose(); budget = prm.ns_seed; }
            run_budget(budget, prm.ns_seed);
            block_matvec();
            bool no_correction = true;
            for (int i = 0; i < m && no_correction; ++i)
                for (int s = 0; s < 6; ++s)
                    if (dv[i][s] != 0.0) { no_correction = false; break; }
            if (no_correction) break;
            std::vector<M16> base = L0.vmid;
            double alpha = 1.0;
            for (int l = 0; l < LANES; ++l) {
                for (int p = 0; p < m; ++p) {
                    const int i = isl.order[p];
                    M16 d6 = screw_from6(dv[p].data());
                    lanes[l].vmid[i] = base[i] + d6.scale(alpha);
                }
                alpha = alpha / 2.0;
            }
            for (int l = 0; l < LANES; ++l) dispatch(lanes[l], false);
            for (int l = 0; l < LANES; ++l) body_post(lanes[l]);
            int accepted = -1;
            double tnorm = 0.0;
            for (int l = 0; l < LANES; ++l) {
                double t = wnorm(lanes[l]);
                if (t < rnorm) { accepted = l; tnorm = t; break; }
            }
            if (accepted >= 0) {
                if (accepted != 0) {  // adopt_lane (cache.rs:1115-1136)
                    Lane& S = lanes[accepted];
                    L0.vmid = S.vmid;
                    L0.midpoint = S.midpoint;
                    L0.solve_vel = S.solve_vel;
                    L0.rhs = S.rhs;
                    L0.mass = S.mass;
                }
                has_carried = true;
                carried = tnorm;
            } else {
                for (int p = 0; p < m; ++p) {
                    const int i = isl.order[p];
                    M16 d6 = screw_from6(dv[p].data());
                    L0.vmid[i] = base[i] + d6.scale(0.0);
                }
                isl.stats.line_search_rejects++;
                break;
            }
        }
        L0.vmid = best_vals;
    }
    static bool log2_floor(double v, int& e) {  // newton.rs:101-121
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
    bool signal_log2(bool assembled, int& r) {  // newton.rs:688-702
        if (!assembled) return false;
        double a = reduce(isl.A, m(), false);
        double mass = reduce(lanes[0].mass, m(), false);
        double stiff = is_effective_zero(mass) ? 0.0 : a / mass;  // cache.rs:462-469
        return log2_floor(stiff, r);
    }
    void commit_span(double span, double half_, bool converged) {  // newton.rs:708-753
        const Lane& L0 = lanes[0];
        for (int p = 0; p < m(); ++p) {   // the unknowns only (newton.rs:704-753)
            const int i = isl.order[p];
            Body& b = isl.bodies[i];
            const M16& vs = L0.vmid[i];
            M16 m_start = b.pose;
            M16 pose_mid = compose(m_start, twist_exp(vs, half_));
            M16 v_body = transport_contra(vs, motor_inverse(pose_mid));
            M16 p_s_mid = transport_co(inertia_apply(b.inertia, v_body), pose_mid);
            M16 p_s_new = converged ? p_s_mid.scale(2.0) - screw_from6(snap_mom[i].data()) : p_s_mid;
            M16 m_new = compose(m_start, twist_exp(vs, span));
            b.pose = m_new;
            b.momentum = transport_co(p_s_new, motor_inverse(m_new));
        }
    }
    // solve_step (newton.rs:543-683)
    void solve_step() {
        uint64_t remaining = 1ull << MAX_DEPTH;
        int depth = 0;
        isl.trace.clear();
        while (remaining > 0) {
            uint64_t units = 1ull << (MAX_DEPTH - depth);
            while (units > remaining) { depth += 1; units >>= 1; }
            double span = dt;
            for (int k = 0; k < depth; ++k) span = span / 2.0;
            half = span / 2.0;
            isl.stats.spans++;
            for (int i = 0; i < nb(); ++i) screw_to6(world_momentum(isl.bodies[i]), snap_mom[i].data());
            bool converged, assembled;
            double best;
            newton_solve(converged, best, assembled);
            bool finite = best < 1000000000.0;
            bool last = depth == MAX_DEPTH;
            if (!converged && !last) {
                isl.X = isl.Xacc;                       // restore_rows
                isl.seeded = isl.accepted_seeded;       // finish_rollback
                int r, down;
                isl.trace.push_back({depth, 0, 0, NO_SIGNAL});
                if (signal_log2(assembled, r)) {
                    isl.trace.back().signal = r;
                    int q = r - THRESHOLD_LOG2;
                    int fl = (q >= 0) ? q / PER_HALVING_LOG2 : -((-q + PER_HALVING_LOG2 - 1) / PER_HALVING_LOG2);
                    down = fl + 1;
                } else down = 1;
                int hi = MAX_DEPTH - depth;
                down = std::max(1, std::min(down, hi));
                depth += down;
                continue;
            }
            isl.trace.push_back({depth, converged ? 1 : 0, (converged || finite) ? 1 : 0, NO_SIGNAL});
            if (converged || finite) {
                if (!converged) isl.stats.unconverged_commits++;
                commit_span(span, half, converged);
                isl.Xacc = isl.X;                        // publish_rows
                double n = reduce(isl.X, m(), false);    // x_reduce
                const double huge = 1000000000.0;
                if (n < huge * huge) isl.accepted_seeded = isl.seeded;  // finish_checkpoint
                else { isl.seeded = false; isl.accepted_seeded = false; }
            } else {
                isl.X = isl.Xacc;
                isl.seeded = isl.accepted_seeded;
            }
            remaining -= units;
            auto aligned = [&](int d) {
                if (d <= 0) return false;
                uint64_t bigger = 1ull << (MAX_DEPTH - (d - 1));
                return bigger <= remaining && remaining % bigger == 0;
            };
            int r;
            if (remaining > 0 && converged && aligned(depth) && signal_log2(assembled, r)) {
                isl.trace.back().signal = r;
                int q = THRESHOLD_LOG2 - r;
                int up = (q >= 0) ? q / PER_HALVING_LOG2 : -((-q + PER_HALVING_LOG2 - 1) / PER_HALVING_LOG2);
                while (up > 0 && aligned(depth)) { depth -= 1; up -= 1; }
            }
        }
    }
};

// Newton::step_all for one island (newton.rs:789-861).
inline void newton_step_island(Island& isl, const std::vector<std::array<double, 6>>& ext, double dt, double warp,
                               const NewtonParams& prm) {
    cache_sync(isl);
    NewtonCtx ctx(isl, prm, dt, warp, ext);
    ctx.solve_step();
}

// Explicit family through the bridge (bridge.rs:19-63): world wrench = external +
// joints at the current pose and world velocity, pulled back to the body frame.
enum ExplicitKind { EXPLICIT_EULER = 0, SYMPLECTIC_EULER = 1, LIE_EULER = 3 };
inline void explicit_step_island(Island& isl, const std::vector<std::array<double, 6>>& ext, double dt, double warp,
                                 int kind) {
    int m = (int)isl.bodies.size();
    std::vector<M16> wv(m);
    for (int i = 0; i < m; ++i) wv[i] = world_velocity(isl.bodies[i]);
    std::vector<std::array<double, 6>> total = ext;
    for (const Joint& e : isl.joints) {
        M16 poses[2] = {fn_pre(isl.bodies[e.a].pose, wv[e.a], 0.0), fn_pre(isl.bodies[e.b].pose, wv[e.b], 0.0)};
        M16 vels[2] = {wv[e.a], wv[e.b]};
        double out[2][6];
        joint_plain(e, poses, vels, dt, warp, out);
        for (int s = 0; s < 6; ++s) { total[e.a][s] += out[0][s]; total[e.b][s] += out[1][s]; }
    }
    for (int i = 0; i < m; ++i) {
        Body& b = isl.bodies[i];
        M16 w_body = transport_co(screw_from6(total[i].data()), motor_inverse(b.pose));
        if (kind == SYMPLECTIC_EULER) symplectic_euler_step(b, w_body, dt);
        else if (kind == LIE_EULER) lie_euler_step(b, w_body, dt);
        else explicit_euler_step(b, w_body, dt);
    }
}

// Island::recompute_centroid (island.rs:206-244)
inline void recompute_centroid(Island& isl) {
    if (isl.bodies.empty()) { isl.origin[0] = isl.origin[1] = isl.origin[2] = 0.0; return; }
    double total_m = 0.0;
    double wx[3] = {0.0, 0.0, 0.0};
    for (const Body& b : isl.bodies) {
        double mm = b.inertia.mass;
        V3<double> r = body_position(b);
        total_m += mm;
        for (int k = 0; k < 3; ++k) wx[k] += r[k] * mm;
    }
    if (is_effective_zero(total_m)) return;
    double inv = 1.0 / total_m;
    V3<double> delta(wx[0] * inv, wx[1] * inv, wx[2] * inv);
    M16 shift = twist_exp(screw_new(-delta, V3<double>()), 1.0);
    for (Body& b : isl.bodies) b.pose = compose(b.pose, shift);
    for (int k = 0; k < 3; ++k) isl.origin[k] += delta[k];
}

// A world: islands plus one gravity propagator over every body (slot = global index
// in island order). `front` holds the charges the next accumulate reads.
struct World {
    std::vector<Island> islands;
    double G = 1.0;
    bool gravity = true;
    double uniform[3] = {0.0, 0.0, 0.0};   // UniformField acceleration (field.rs:16-44), 0 = off
    std::vector<Charge> front;
    int integrator = 2;                     // 0 ExplicitEuler, 1 SymplecticEuler, 2 Newton, 3 LieEuler
    NewtonParams prm;
    int64_t steps = 0;

    int n_bodies() const { int n = 0; for (auto& i : islands) n += (int)i.bodies.size(); return n; }
    void init_front() { front.assign(n_bodies(), zero_charge()); }

    Charge charge_of(const Island& isl, const Body& b) const {  // gravity.rs:485-496
        Charge c;
        V3<double> local = body_position(b);
        for (int k = 0; k < 3; ++k) { c.local[k] = local[k]; c.com[k] = local[k] + isl.origin[k]; }
        c.grav = G * b.inertia.mass;
        c.size = b.size;
        return c;
    }

    // One Mechanism::step over every island then advance_epoch (mechanism.rs:526-600,
    // gravity.rs:392-397). Islands only read `front` during the step, so they are
    // independent; OpenMP spreads them (and the gravity rows) over host cores.
    void step(double dt, double warp) {
        if ((int)front.size() != n_bodies()) init_front();
        int nb = n_bodies();
        std::vector<int> base(islands.size() + 1, 0);
        for (size_t k = 0; k < islands.size(); ++k) base[k + 1] = base[k] + (int)islands[k].bodies.size();
        std::vector<std::array<double, 6>> ext_all(nb);
        double g_inv = 1.0 / G;
        #pragma omp parallel for schedule(dynamic, 16)
        for (int gi = 0; gi < nb; ++gi) {
            double w[6] = {0, 0, 0, 0, 0, 0};
            if (gravity) gravity_accumulate_for(gi, front, g_inv, w);
            for (int s = 0; s < 6; ++s) ext_all[gi][s] = 0.0 + w[s];
        }
        std::vector<Charge> back(front.size(), zero_charge());
        #pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < (int)islands.size(); ++k) {
            Island& isl = islands[k];
            int m = (int)isl.bodies.size();
            std::vector<std::array<double, 6>> ext(m);
            for (int i = 0; i < m; ++i) {
                ext[i] = ext_all[base[k] + i];
                if (uniform[0] != 0.0 || uniform[1] != 0.0 || uniform[2] != 0.0) {
                    // UniformField (field.rs:28-44): F = m·a, torque 0, added after gravity
                    double mm = isl.bodies[i].inertia.mass;
                    for (int s = 0; s < 3; ++s) { ext[i][s] += uniform[s] * mm; ext[i][3 + s] += 0.0; }
                }
            }
            if (integrator == 2) newton_step_island(isl, ext, dt, warp, prm);
            else explicit_step_island(isl, ext, dt, warp, integrator);
            recompute_centroid(isl);
            for (int i = 0; i < m; ++i) back[base[k] + i] = charge_of(isl, isl.bodies[i]);
        }
        front = back;
        steps++;
    }
};

}  // namespace oracle
